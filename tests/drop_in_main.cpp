// A driver with the call sequence of the reference src/main.cu:14-322 (loader, csr_to_csc, ColSort,
// formatTransform, FastLoad_cpu, the naive CSC and vendor comparators, FastLoad_spmv, Tile_destroy),
// compiled against include/fastload_compat.hpp and linked with libfastload.so and
// libfastload_compare.so by tests/test_gpu_parity.py::test_reference_style_driver.
#include <cstdlib>
#include <cstring>
#include "fastload_compat.hpp"
int main(int argc, char **argv) {
  int rowA, colA, nnz, sym; int *rp, *ci; double *v;
  if (mmio_allinone(&rowA, &colA, &nnz, &sym, &rp, &ci, &v, argv[1]) != 0) return 1;
  for (int i = 0; i < nnz; i++) v[i] = i % 10;
  int *cr = (int*)malloc(4*nnz), *cp = (int*)malloc(4*(colA+1)), *npc = (int*)malloc(4*colA);
  double *cv = (double*)malloc(8*nnz); double t;
  csr_to_csc(t, rowA, colA, nnz, v, rp, ci, cv, cr, cp, npc);
  double *x = (double*)malloc(8*colA), *sx = (double*)malloc(8*colA);
  for (int i = 0; i < colA; i++) x[i] = i % 10;
  int *tr = (int*)malloc(4*nnz), *sn = (int*)malloc(4*colA); double *tv = (double*)malloc(8*nnz);
  ColSort(t, rowA, colA, nnz, npc, cp, cr, cv, tr, tv, sn, x, sx);
  slide_matrix A; int h; double t2;
  formatTransform(t, t2, &A, tr, tv, sn, nnz, rowA, colA, h);
  double *y = (double*)calloc(rowA, 8), *g = (double*)calloc(rowA, 8);
  for (int c = 0; c < colA; c++) for (int j = cp[c]; j < cp[c+1]; j++) g[cr[j]] += cv[j] * x[c];
  FastLoad_cpu(&A, nnz, rowA, colA, sx, y, g);  // src/main.cu:263-277 (prints only on a mismatch)
  std::memset(y, 0, 8 * (size_t)rowA);
  csc_spmv(argv[1], rowA, colA, nnz, cp, cr, cv, x, y, g);  // src/main.cu:288-298
  std::memset(y, 0, 8 * (size_t)rowA);
  call_cusparse_spmv(argv[1], nnz, rowA, colA, rp, ci, v, x, y, g);  // src/main.cu:300-310
  std::memset(y, 0, 8 * (size_t)rowA);
  FastLoad_spmv(argv[1], &A, nnz, colA, rowA, sx, y, g);
  Tile_destroy(&A);
  return 0;
}
