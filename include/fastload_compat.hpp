// fastload_compat.hpp -- the reference's C++ host API (skyler-ruiter/FastLoad, src/) with its
// exact names and reference-parameter signatures, as inline wrappers over the C ABI in
// fastload.h. A driver written against the reference (src/main.cu) compiles against this header
// and links libfastload.so instead of the CUDA sources. As in the reference the wrappers return
// void; failures are printed to stderr (the C ABI returns status codes, fastload.h).
#ifndef FASTLOAD_COMPAT_HPP
#define FASTLOAD_COMPAT_HPP

#include <cstdio>

#include "fastload.h"
#include "fastload_compare.h"

namespace fastload_compat_detail {
inline void report(int rc, const char *what) {
    if (rc < 0) std::fprintf(stderr, "FastLoad %s failed (%d): %s\n", what, rc, fl_last_error());
}
}  // namespace fastload_compat_detail

// src/auxiliary.h:65-67
inline int mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric, int **csrrowptr, int **csrcolidx,
                         double **csrval, char *filename) {
    return fl_mmio_allinone(m, n, nnz, isSymmetric, csrrowptr, csrcolidx, csrval, filename);
}

// src/csr2csc_cuda.h:51-53
inline void csr_to_csc(double &time_csr2csc, int m, int n, int nnz, const double *csr_val, const int *csr_row_ptr,
                       const int *csr_col_ind, double *csc_val, int *csc_row_ind, int *csc_col_ptr, int *nnzpercol) {
    fastload_compat_detail::report(fl_csr_to_csc(&time_csr2csc, m, n, nnz, csr_val, csr_row_ptr, csr_col_ind, csc_val,
                                                 csc_row_ind, csc_col_ptr, nnzpercol),
                                   "csr_to_csc");
}

// src/ColSort_GPU.h:46-58
inline void ColSort(double &timeForSort, int rowA, int colA, int nnz, int *nnzPerCol, int *cscColPtr, int *cscRowIdx,
                    double *cscVal, int *sortrowidxTmp, double *sortvalTmp, int *sortnnzTmp, double *x,
                    double *sortx) {
    fastload_compat_detail::report(fl_ColSort(&timeForSort, rowA, colA, nnz, nnzPerCol, cscColPtr, cscRowIdx, cscVal,
                                              sortrowidxTmp, sortvalTmp, sortnnzTmp, x, sortx),
                                   "ColSort");
}

// src/formatTransform_GPU.h:289-298
inline void formatTransform(double &timeFormatTran, double &timeFormatclas, slide_matrix *matrix,
                            int *sortrowidx_tmp, double *sortval_tmp, int *sortnnz_tmp, int nnz, int rowA, int colA,
                            int &h_count) {
    fastload_compat_detail::report(fl_formatTransform(&timeFormatTran, &timeFormatclas, matrix, sortrowidx_tmp,
                                                      sortval_tmp, sortnnz_tmp, nnz, rowA, colA, &h_count),
                                   "formatTransform");
}

// src/FastLoad_GPU.h:275-282 (note the colA, rowA order)
inline void FastLoad_spmv(char *filename, slide_matrix *matrix, int nnz, int colA, int rowA, double *x, double *y,
                          double *y_golden) {
    fastload_compat_detail::report(fl_FastLoad_spmv(filename, matrix, nnz, colA, rowA, x, y, y_golden, nullptr),
                                   "FastLoad_spmv");
}

// src/FastLoad_CPU.h:2-9 (the reference's serial host loop: one thread, same roundings)
inline void FastLoad_cpu(slide_matrix *matrix, int nnz, int rowA, int colA, double *x, double *y,
                         double *y_goldencsc) {
    fastload_compat_detail::report(fl_FastLoad_cpu(matrix, nnz, rowA, colA, x, y, y_goldencsc, 1, nullptr),
                                   "FastLoad_cpu");
}

// src/format.h:29-42
inline void Tile_destroy(slide_matrix *matrix) { fl_Tile_destroy(matrix); }

// The two comparators src/main.cu times beside FastLoad (measurement only; a caller of these two
// links libfastload_compare.so as well).
// src/CSCSpMV.h:29-40
inline void csc_spmv(char *filename, int m, int n, int nnz, int *csc_ptr, int *csc_rowidx, double *csc_val, double *x,
                     double *y, double *y_golden) {
    const int rc = flc_csc_spmv(filename, m, n, nnz, csc_ptr, csc_rowidx, csc_val, x, y, y_golden, nullptr);
    if (rc < 0) std::fprintf(stderr, "FastLoad csc_spmv failed (%d): %s\n", rc, flc_last_error());
}

// src/cusparse_cuda.h:27-36 (rocSPARSE in place of cuSPARSE)
inline void call_cusparse_spmv(char *filename, const int A_nnz, const int A_num_rows, const int A_num_cols,
                               int *hA_csrOffsets, int *hA_columns, double *hA_values, double *hX, double *hY,
                               double *hY_result) {
    const int rc = flc_rocsparse_spmv(filename, A_nnz, A_num_rows, A_num_cols, hA_csrOffsets, hA_columns, hA_values,
                                      hX, hY, hY_result, nullptr);
    if (rc < 0) std::fprintf(stderr, "FastLoad call_cusparse_spmv failed (%d): %s\n", rc, flc_last_error());
}

#endif  // FASTLOAD_COMPAT_HPP
