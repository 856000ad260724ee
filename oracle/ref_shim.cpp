// Build recipe helper for oracle/_ref -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the reference's own host-only headers where they lie (-I <reference>/src):
// src/coo2csc.h (csr2csc, col_sort, formattransation, tile_classification), which pulls in
// src/auxiliary.h (mmio_allinone, exclusive_scan), src/mmio.h and src/format.h, plus
// src/FastLoad_CPU.h (FastLoad_cpu). No reference source is copied into this repository; this
// file only gives the compiled functions C linkage so the tests can call them through ctypes.
// The tile width comes from the reference's own `slidesize` macro (format.h:44-46), set on
// the compiler command line (oracle/Makefile builds a 32-wide and a 64-wide library).
#include <cstring>
#include "coo2csc.h"
#include "FastLoad_CPU.h"

extern "C" {

int ref_slidesize() { return slidesize; }

int ref_mmio_allinone(int *m, int *n, int *nnz, int *isSym, int **rowptr, int **colidx, double **val,
                      char *filename)
{
    return mmio_allinone(m, n, nnz, isSym, rowptr, colidx, val, filename);
}

void ref_csr2csc(int nnz, int m, int n, int *csrcolidx, int *csrrowptr, double *csrval, int **cscrowidx,
                 int **csccolptr, int **nnzpercol, double **cscval)
{
    csr2csc(nnz, m, n, csrcolidx, csrrowptr, csrval, cscrowidx, csccolptr, nnzpercol, cscval);
}

void ref_col_sort(int n, int *nnzpercol, int *colptr, int *rowidx, double *val, int *tmp_row, double *tmp_val,
                  int *sortnnz, double *x, double *sortx, int nnz)
{
    col_sort(n, nnzpercol, colptr, rowidx, val, tmp_row, tmp_val, sortnnz, x, sortx, nnz);
}

// formattransation followed by tile_classification on a zero-initialised slide_matrix.
slide_matrix *ref_transform_classify(int *tmp_row, double *tmp_val, int *sortnnz, int nnz, int m, int n)
{
    slide_matrix *A = (slide_matrix *)calloc(1, sizeof(slide_matrix));
    formattransation(A, tmp_row, tmp_val, sortnnz, nnz, m, n, 0);
    tile_classification(A, nnz, m, n);
    return A;
}

void ref_destroy(slide_matrix *A)
{
    Tile_destroy(A);
    free(A);
}

// FastLoad_cpu with y accumulated in place; the golden argument is y itself so the built-in
// exact check never reports (the tests compare results on their own).
void ref_fastload_cpu(slide_matrix *A, int nnz, int m, int n, double *sortx, double *y)
{
    FastLoad_cpu(A, nnz, m, n, sortx, y, y);
}

void ref_free(void *p) { free(p); }

}  // extern "C"
