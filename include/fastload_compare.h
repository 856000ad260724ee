/*
 * fastload_compare.h -- comparators for measurement only (libfastload_compare.so): the reference's
 * naive CSC kernel and a vendor CSR SpMV, which the reference times beside FastLoad
 * (src/main.cu:289-311). Neither is ever a dispatch target of libfastload.so; they exist so a
 * benchmark can put a naive and a vendor number next to FastLoad's. Plain pointers and sizes, the
 * status codes of fastload.h (FL_OK, FL_ERR_*), messages via flc_last_error().
 */
#ifndef FASTLOAD_COMPARE_H
#define FASTLOAD_COMPARE_H

#include "fastload.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message describing the last error of this library on the calling thread ("" if none). */
const char *flc_last_error(void);

/* Replaces csc_spmv (src/CSCSpMV.h:29-131), host pointers: the naive CSC kernel of
 * src/CSCSpMV.h:4-27 (one lane per column, grid-stride over columns, one atomic per entry into y)
 * launched 1000 times, each timed as the host clock around launch + synchronize with y cleared
 * before it; y receives A*x; a 1%-relative check against y_golden (skipped if NULL) printed as
 * "BASIC CSC GPU PASS! (t ms g GFlops)" or "BASIC CSC check NO PASS! error_count_cuda = N".
 * *time_ms (if not NULL) receives the mean. Returns FL_OK, or 1 when the check fails. filename is
 * only used in the printout (may be NULL). */
int flc_csc_spmv(const char *filename, int m, int n, int nnz, const int *csc_ptr, const int *csc_rowidx,
                 const double *csc_val, const double *x, double *y, const double *y_golden, double *time_ms);

/* The same naive kernel on device-resident arrays: y += A*x (the caller clears y). fp64 = 1:
 * double values, x and y; 0: float. */
int flc_csc_spmv_device(int m, int n, const int *d_colptr, const int *d_rowidx, const void *d_val, const void *d_x,
                        void *d_y, int fp64, void *stream);

/* Replaces call_cusparse_spmv (src/cusparse_cuda.h:27-155) with rocSPARSE, host pointers:
 * y = A*x for a 0-based CSR matrix in FP64, timed as the reference times it (one call including
 * the library's set-up, host clock around it); y compared exactly with y_result (skipped if NULL),
 * printed as "rocSPARSE test PASSED! (t ms g GFlops)" or "rocSPARSE test FAILED: wrong result".
 * Returns FL_OK, 1 when the check fails, FL_ERR_HIP on a rocSPARSE error. */
int flc_rocsparse_spmv(const char *filename, int nnz, int rows, int cols, const int *csr_ptr, const int *csr_col,
                       const double *csr_val, const double *x, double *y, const double *y_result, double *time_ms);

/* rocSPARSE CSR SpMV on device-resident arrays (the adaptive CSR kernel, analysed once at
 * creation): flc_rocsparse_run computes y = A*x on the given stream. fp64 as above. */
typedef struct flc_rocsparse flc_rocsparse;
int flc_rocsparse_create(flc_rocsparse **h, int rows, int cols, int nnz, const int *d_rowptr, const int *d_col,
                         const void *d_val, int fp64, void *stream);
int flc_rocsparse_run(flc_rocsparse *h, const void *d_x, void *d_y, void *stream);
int flc_rocsparse_destroy(flc_rocsparse *h);

#ifdef __cplusplus
}
#endif

#endif /* FASTLOAD_COMPARE_H */
