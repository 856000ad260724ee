/*
 * fastload.h -- C ABI of the MI355X-native FastLoad (gfx950, one GPU).
 *
 * FastLoad is a CSC sparse matrix-vector product y = A*x: columns (and x) are sorted by
 * their non-zero count, and the sorted non-zeros are laid out in lane-wide tiles so that
 * matrix and x loads coalesce. This header is the drop-in boundary for the reference's host
 * API (skyler-ruiter/FastLoad, src/): every declaration cites the reference interface it
 * replaces. Plain pointers and sizes only; no HIP or torch types. include/fastload_compat.hpp
 * adds inline C++ wrappers with the reference's exact names and reference-parameter
 * signatures.
 *
 * Errors: every int-returning entry point returns FL_OK (0) or a negative FL_ERR_* code and
 * records a message readable with fl_last_error(). (The reference prints CUDA errors and
 * continues, src/FastLoad_GPU.h:4-27.)
 */
#ifndef FASTLOAD_H
#define FASTLOAD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- status codes */
#define FL_OK 0
#define FL_ERR_OPEN (-1)        /* mmio_allinone: fopen failed            (src/auxiliary.h:81-82) */
#define FL_ERR_BANNER (-2)      /* mmio_allinone: bad Matrix Market banner (src/auxiliary.h:84-88) */
#define FL_ERR_TRUNC (-3)       /* mmio_allinone: fewer entry lines than announced */
#define FL_ERR_SIZE (-4)        /* mmio_allinone: missing/invalid size line (src/auxiliary.h:96-98) */
#define FL_ERR_RANGE (-5)       /* an index outside the matrix */
#define FL_ERR_UNSUPPORTED (-6) /* dense ("array") Matrix Market files */
#define FL_ERR_ALLOC (-7)       /* host or device allocation failed */
#define FL_ERR_INVALID (-8)     /* invalid argument (negative size, NULL pointer, unsorted input) */
#define FL_ERR_HIP (-9)         /* a HIP runtime call failed */

/* Message describing the last error on the calling thread ("" if none). */
const char *fl_last_error(void);

/* ---------------------------------------------------------- tiled-CSC container
 * Field order and types of slide_matrix (src/format.h:6-20). Arrays are host memory,
 * malloc-allocated by fl_formatTransform and released by fl_Tile_destroy.
 *   sortval/sortrowidx/sortrowindex : nnz entries in tile order
 *   tile_ptr : tilenum+1, tile_len/tile_colidx/tile_format : tilenum
 *   countrow : tilenum+1 (exclusive scan of segments per tile), segmentoffset : segsum+1
 *   tile_tall : not set (NULL), as on the reference's GPU path (src/formatTransform_GPU.h:507-546)
 */
typedef struct {
    double *sortval;
    int *sortrowidx;
    int *tile_ptr;
    int *tile_len;
    int *tile_colidx;
    int *tile_tall;
    int *tile_format;
    int *countrow;
    int *segmentoffset;
    int *sortrowindex;
    int tilenum;
    int segsum;
} slide_matrix;

/* ------------------------------------------------- reference host API (host pointers)
 * Each call uploads its inputs, computes on the GPU, and downloads its outputs, as the
 * reference does. Times are in milliseconds: GPU-event times of the device work, except for
 * fl_FastLoad_spmv, which measures as the reference does (see there).
 */

/* Replaces mmio_allinone (src/auxiliary.h:65-232): Matrix Market file -> 0-based CSR with
 * symmetric/hermitian mirroring; the three arrays are malloc-allocated (free with fl_free). */
int fl_mmio_allinone(int *m, int *n, int *nnz, int *isSymmetric, int **csrrowptr, int **csrcolidx,
                     double **csrval, const char *filename);

/* Frees memory returned by fl_mmio_allinone. */
void fl_free(void *p);

/* Replaces csr_to_csc (src/csr2csc_cuda.h:51-135): CSR -> CSC (rows ascending inside each
 * column, ties by CSR position) plus per-column counts. Outputs are caller-allocated:
 * csc_val/csc_row_ind [nnz], csc_col_ptr [n+1], nnzpercol [n]. */
int fl_csr_to_csc(double *time_ms, int m, int n, int nnz, const double *csr_val, const int *csr_row_ptr,
                  const int *csr_col_ind, double *csc_val, int *csc_row_ind, int *csc_col_ptr, int *nnzpercol);

/* Replaces ColSort (src/ColSort_GPU.h:46-137): stable ascending sort of columns by nnzPerCol;
 * outputs the concatenated sorted columns sortrowidxTmp/sortvalTmp [nnz], sortnnzTmp [colA]
 * and sortx [colA] = x[perm]. rowA is unused, as in the reference. */
int fl_ColSort(double *time_ms, int rowA, int colA, int nnz, const int *nnzPerCol, const int *cscColPtr,
               const int *cscRowIdx, const double *cscVal, int *sortrowidxTmp, double *sortvalTmp,
               int *sortnnzTmp, const double *x, double *sortx);

/* Replaces formatTransform (src/formatTransform_GPU.h:289-596): levels, 32-wide tiles, tile
 * data layout, segments and tile formats. Fills *matrix (malloc'd arrays) and *h_count (the
 * number of levels). sortnnz_tmp has colA entries. rowA is unused, as in the reference. */
int fl_formatTransform(double *timeFormatTran, double *timeFormatclas, slide_matrix *matrix,
                       const int *sortrowidx_tmp, const double *sortval_tmp, const int *sortnnz_tmp, int nnz,
                       int rowA, int colA, int *h_count);

/* Replaces FastLoad_spmv (src/FastLoad_GPU.h:275-409). Note the (colA, rowA) argument order.
 * x is the permuted vector (sortx, colA entries); y (rowA entries, original row order)
 * receives A*x. As in the reference: one dry run, then 100 timed launches (y zeroed before
 * each), a 1%-relative check of y against y_golden (skipped if y_golden is NULL) printed as
 * PASS/NO PASS. *time_ms (if not NULL) receives the mean time of one launch, measured as the
 * reference does (host clock around launch + synchronize, src/FastLoad_GPU.h:349-374). Returns
 * FL_OK, or 1 when the check fails. filename is only used in the printout (may be NULL). */
int fl_FastLoad_spmv(const char *filename, const slide_matrix *matrix, int nnz, int colA, int rowA, const double *x,
                     double *y, const double *y_golden, double *time_ms);

/* Replaces FastLoad_cpu (src/FastLoad_CPU.h:2-46; called at src/main.cu:267): the host loop over
 * the tiles in tile order, y[sortrowidx[j]] += sortval[j] * x[tile_colidx[t] + (j - tile_ptr[t]) %
 * tile_len[t]] with x = sortx (colA entries); y (rowA entries) accumulates, the caller zeroes it.
 * threads = 1 runs the reference's serial loop with its roundings (bit-identical y); threads = 0
 * uses every core the process may run on, threads > 1 that many, splitting the tiles into
 * contiguous ranges and adding into y atomically (same sum up to the order of additions).
 * When y_goldencsc is not NULL, y is compared with it for exact equality as the reference does
 * (src/FastLoad_CPU.h:33-45; meaningful on integer-valued data): mismatches are counted into
 * *errors (if not NULL), printed as "Check NO PASS! error_count_check_cpu= N", and the call
 * returns 1. Returns FL_OK, FL_ERR_INVALID (bad layout), or FL_ERR_RANGE (a row index outside
 * [0, rowA); those elements are skipped). */
int fl_FastLoad_cpu(const slide_matrix *matrix, int nnz, int rowA, int colA, const double *x, double *y,
                    const double *y_goldencsc, int threads, int *errors);
/* The same loop in FP32 (no reference counterpart; the reference is FP64 only): values sortval
 * [nnz] in tile order, x and y float, products and sums rounded to FP32 in the same order. */
int fl_FastLoad_cpu_f32(const slide_matrix *matrix, const float *sortval, int nnz, int rowA, int colA,
                        const float *x, float *y, int threads);

/* Replaces Tile_destroy (src/format.h:29-42): frees the arrays fl_formatTransform set
 * (NULL fields are skipped) and zeroes the struct. */
void fl_Tile_destroy(slide_matrix *matrix);

/* ------------------------------------------------- on-disk cache of the preprocessing
 * No reference counterpart: the reference rebuilds the layout on every run (src/main.cu:151-248).
 * A versioned binary file ("FLMATRIX", version 1, little-endian) holds the tiled layout -- the
 * slide_matrix fields of src/format.h:6-20 and the column order -- and, when written from a
 * device-resident matrix, every prepared schedule. Every section carries a 64-bit hash; a reader
 * refuses a file whose magic, version, byte order, section bounds, lengths or hashes do not match
 * (FL_ERR_INVALID; FL_ERR_OPEN when the file cannot be opened, read or written).
 *
 * fl_slide_save writes a host slide_matrix (as fl_formatTransform returns it) with its sizes, the
 * level count and, when perm is not NULL, the column order (colA entries: sortx = x[perm]).
 * fl_slide_load reads the slide_matrix sections of any such file (also one written by
 * fl_matrix_save) into malloc'd arrays (free with fl_Tile_destroy; *perm, if asked for and
 * present, with fl_free; else set to NULL). */
int fl_slide_save(const char *path, const slide_matrix *matrix, int rowA, int colA, int nnz, int h_count,
                  const int *perm);
int fl_slide_load(const char *path, slide_matrix *matrix, int *rowA, int *colA, int *nnz, int *h_count, int **perm);

/* ------------------------------------------------------------ device memory (plumbing)
 * Minimal device-memory helpers so callers without a GPU framework can keep inputs resident.
 * Streams are passed as void* (a hipStream_t, or NULL for the default stream). */
int fl_device_count(int *count);
int fl_malloc(void **dptr, size_t bytes);
int fl_free_device(void *dptr);
int fl_memcpy_h2d(void *dst, const void *src, size_t bytes);
int fl_memcpy_d2h(void *dst, const void *src, size_t bytes);
int fl_memcpy_d2d(void *dst, const void *src, size_t bytes);
int fl_memset_device(void *dptr, int value, size_t bytes);
int fl_synchronize(void);
int fl_stream_create(void **stream);
int fl_stream_destroy(void *stream);
int fl_stream_synchronize(void *stream);

/* Event-timed region helpers (ms between two events recorded on a stream). */
int fl_event_create(void **ev);
int fl_event_destroy(void *ev);
int fl_event_record(void *ev, void *stream);
int fl_event_elapsed(float *ms, void *start, void *stop);

/* ------------------------------------------------ device-resident FastLoad matrix
 * The whole reference pipeline (ColSort_GPU + formatTransform_GPU) run on device buffers,
 * with no host round trip of the matrix, and kept resident for repeated SpMVs -- the
 * reference re-uploads everything per call (src/FastLoad_GPU.h:304-328).
 *   width: tile width W (32 = the reference's hard-coded layout, bit-exact; 64 = wave-wide).
 * The input CSC (rows ascending inside columns) is only read during fl_matrix_create.
 */
typedef struct fl_matrix fl_matrix;

typedef struct {
    int m, n, nnz;
    int width;    /* W */
    int levels;   /* h_count */
    int tilenum;
    int segsum;
    int max_col_nnz;
} fl_matrix_info;

/* Value storage kinds for the precision variants (no reference counterpart: FP64 only). */
#define FL_VAL_F64 0    /* double values (drop-in FP64 mode), x/y double */
#define FL_VAL_F32 1    /* float values, x/y float */
#define FL_VAL_BF16 2   /* bfloat16 values, FP32 accumulation, x/y float */
#define FL_VAL_BF16X2 3 /* split hi+lo bfloat16 values (hi = bf16(v), lo = bf16(v - hi)); x float,
                           y double. hi*x and lo*x are FP32 products, recombined in FP64: per row
                           by the FAST kernel (two FP32 row sums), per entry by WIN/WINR (FP64 LDS
                           sums) and by PB (the FP64 sum of the pair rounded to one FP32 product,
                           relative error 2^-24, before the FP64 band sums) */
#define FL_VAL_PACKED 4 /* batched schedules (WIN/WINR/PB) only: each entry one 32-bit word
                           holding its value and its 16-bit schedule slot, the value rounded
                           to the mantissa bits left over (31 - slot bits - exponent-code bits,
                           at least 12; refused otherwise) with an exponent code sized to the
                           matrix's exponent range; x and y float, FP64 LDS sums */

int fl_matrix_create(fl_matrix **A, int m, int n, int nnz, const int *d_colptr, const int *d_rowidx,
                     const double *d_val, int width, void *stream);
int fl_matrix_destroy(fl_matrix *A);
int fl_matrix_get_info(const fl_matrix *A, fl_matrix_info *info);

/* Copies the device layout to caller-allocated host arrays (any pointer may be NULL):
 * perm/sortnnz [n], tile_ptr [tilenum+1], tile_len/tile_colidx/tile_format [tilenum],
 * sortrowidx/sortrowindex [nnz], sortval [nnz], countrow [tilenum+1], segmentoffset [segsum+1],
 * lvl_col/lvl_base/lvl_tall [levels]. */
int fl_matrix_export(const fl_matrix *A, int *perm, int *sortnnz, int *tile_ptr, int *tile_len, int *tile_colidx,
                     int *tile_format, int *sortrowidx, int *sortrowindex, double *sortval, int *countrow,
                     int *segmentoffset, int *lvl_col, int *lvl_base, int *lvl_tall);

/* sortx[i] = x[perm[i]] on the device (Pre4sortedx, src/ColSort_GPU.h:9-19). */
int fl_permute_x_f64(const fl_matrix *A, const double *d_x, double *d_sortx, void *stream);
int fl_permute_x_f32(const fl_matrix *A, const float *d_x, float *d_sortx, void *stream);

/* Builds (once) the value copy of the given kind from the FP64 values. FL_VAL_BF16X2 keeps the
 * hi and lo arrays (FAST kernel) and the pair packed as one 32-bit word per entry (hi << 16 | lo,
 * read by the batched schedules). FL_VAL_PACKED builds the value + slot words of every batched
 * schedule prepared so far and of those prepared later (each value rounded to nearest even, from
 * FP64, to the mantissa bits left over; magnitudes below 2^-126 stored as zero). */
int fl_matrix_prepare_values(fl_matrix *A, int val_kind, void *stream);
/* The format of a schedule's FL_VAL_PACKED words: mantissa, exponent-code and slot bits (their
 * sum plus the sign bit is at most 32). The relative rounding error of a value is at most
 * 2^-(mantissa_bits+1). FL_KERNEL_HYB: the words of its PB sub-matrix (the heavy part keeps FP32
 * values). FL_ERR_INVALID when the schedule has no packed words. */
int fl_matrix_pack_info(const fl_matrix *A, int kernel, int *mantissa_bits, int *exponent_bits, int *index_bits);

/* SpMV kernels:
 *   FL_KERNEL_DROPIN: the reference per-tile algorithm (FastLoad___kernel,
 *     src/FastLoad_GPU.h:146-273) over the reference layout: reads tile_ptr/len/colidx/format and
 *     sortrowindex; formats 0/1/2 (per-lane atomics / one atomic per tile / segmented sums).
 *   FL_KERNEL_FAST: same tiles, geometry derived from the per-level table instead of per-tile
 *     metadata, segment heads derived by comparing neighbouring row indices (no sortrowindex),
 *     several tiles per wave in flight. */
#define FL_KERNEL_DROPIN 0
#define FL_KERNEL_FAST 1
/*   FL_KERNEL_WIN: batches of consecutive tiles per workgroup; the distinct rows of a batch are
 *     compacted to LDS slots (2-byte slot per entry instead of a 4-byte row), products are summed
 *     per slot in FP64 LDS atomics and each slot is flushed once: a row no other batch touches
 *     with a plain store (fl_spmv_assign) or read-modify-write (fl_spmv), a shared row with one
 *     global atomic. For rows with locality (stencils, bands, blocks).
 *   FL_KERNEL_PB: two-phase propagation blocking for rows without locality: phase 1 stages the
 *     batch's products in LDS by row band and copies them band-major; phase 2 sums each band of
 *     rows in FP64 LDS and updates y once per row. No global atomics. Matrices of >= 2^25 entries
 *     run in two chunks of batches so that a chunk's products stay in the Infinity Cache. */
#define FL_KERNEL_WIN 2
#define FL_KERNEL_PB 3
/*   FL_KERNEL_WINR: the WIN kernel on tiles grouped by the 64-row window of their first entry
 *     instead of by column: on stencils and bands most rows then belong to a single batch and are
 *     written without an atomic. */
#define FL_KERNEL_WINR 4
/*   FL_KERNEL_HYB (no reference counterpart): the longest sorted columns ("heavy", at most 38912,
 *     or 19456 in FP64) multiply x staged once per workgroup in LDS, row by row in jagged-diagonal
 *     slices of 64 rows (value + 16-bit column per entry, no index traffic for y); the other
 *     columns run the PB schedule as their own sub-matrix, whose phase 2 also adds the heavy
 *     part's y. For matrices whose few longest columns hold most entries (power laws). Value kinds
 *     FL_VAL_F64, FL_VAL_F32 and FL_VAL_PACKED (the light part's PB words; heavy values FP32). */
#define FL_KERNEL_HYB 5

/* y += A * sortx (y must be zeroed by the caller for y = A*x, as src/main.cu:288-313 does).
 * FL_KERNEL_WIN/FL_KERNEL_PB need fl_matrix_prepare_schedule first; they write FL_VAL_BF16X2
 * results to a double y directly and need no work buffer.
 * val_kind selects the value copy (see fl_matrix_prepare_values); pointer types:
 *   FL_VAL_F64: x,y double.  FL_VAL_F32/FL_VAL_BF16: x,y float.
 *   FL_VAL_BF16X2: x float; y double (recombined); needs d_work of 2*m floats (zeroed inside). */
int fl_spmv(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *d_work, void *stream);
/* y = A * sortx (no clearing needed; BLAS beta = 0). FL_KERNEL_PB writes every row of y once in
 * its second phase; the other kernels clear y with a fill kernel, then accumulate (WIN/WINR
 * store the rows a single batch owns instead of reading them back). */
int fl_spmv_assign(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *d_work,
                   void *stream);

/* Cache policy of the read-once streams (values, slots, PB products, HYB slots) of the WIN/WINR,
 * PB and HYB schedules; no reference counterpart. FL_STREAM_AUTO (the default): the streaming
 * (nontemporal) hint when a schedule streams more than the 256 MB Infinity Cache per SpMV,
 * ordinary loads otherwise; FL_STREAM_NEVER / FL_STREAM_ALWAYS force one of the two. Applies
 * to every later call in the process; y does not depend on it, only the speed does. Returns
 * FL_ERR_INVALID for another value. */
#define FL_STREAM_AUTO (-1)
#define FL_STREAM_NEVER 0
#define FL_STREAM_ALWAYS 1
int fl_set_stream_policy(int policy);

/* fl_matrix_save writes a device-resident matrix: its layout (width, column order, tiles,
 * elements, classification), which value copies were prepared, and the prepared WIN/WINR, PB and
 * row-order schedules. fl_matrix_load reads one back into a new handle: every row and column
 * index, pointer array and schedule slot or run is checked before use; the level table, tile
 * metadata, classification, batch membership and value copies are recomputed on the device (the
 * stored tile metadata must agree with them); the schedule sorts are not redone. */
int fl_matrix_save(const fl_matrix *A, const char *path, void *stream);
int fl_matrix_load(fl_matrix **A, const char *path, void *stream);

/* y = A * sortx like fl_spmv_assign, allocating nothing: no reference counterpart. For callers
 * that cannot allocate device memory during the call (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The PB schedule's product array and the FAST
 * kernel's split-BF16 work buffer live in a static 64 MiB device array per GPU instead of memory
 * the library or the caller allocates; a call needing more (PB: the matrix's product ring, up to
 * nnz products of 4 bytes, 8 in FP64; FAST split BF16: 2*m floats) returns FL_ERR_INVALID before
 * anything is enqueued. Schedules and value copies must be prepared beforehand (those calls do
 * allocate). All calls on one GPU share the array: calls that overlap on one GPU race. */
int fl_spmv_assign_noalloc(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *stream);

/* Builds (once per parameter set) the preprocessing of a batched schedule on a 64-wide matrix:
 *   kernel FL_KERNEL_WIN / FL_KERNEL_WINR: batch_tiles in {64, 128, 256} (LDS <= batch_tiles*512 bytes);
 *   kernel FL_KERNEL_PB : batch_tiles as above, band_rows in [1024, 16384] (LDS = band_rows*8 bytes).
 * The PB schedule keeps a product buffer of nnz floats (doubles in FP64 mode) allocated on its
 * first use. */
int fl_matrix_prepare_schedule(fl_matrix *A, int kernel, int batch_tiles, int band_rows, void *stream);
/* Builds the FL_KERNEL_HYB schedule (no reference counterpart): the last heavy_cols sorted columns
 * in [0, min(n, 38912)] become the heavy part, the others a sub-matrix with the PB schedule of
 * batch_tiles and band_rows (as above). fl_matrix_prepare_schedule(A, FL_KERNEL_HYB, ...) uses
 * heavy_cols = min(n, 32768). Value copies prepared before or after are built for both parts.
 * Extra memory: the heavy entries (8 + 2 bytes each, + 4 with FP32 values), four arrays of m
 * or m/64 entries, the sub-matrix and its PB schedule. */
int fl_matrix_prepare_hyb(fl_matrix *A, int heavy_cols, int batch_tiles, int band_rows, void *stream);
/* The prepared HYB split: heavy columns and heavy entries (the rest are in the PB sub-matrix). */
int fl_matrix_hyb_info(const fl_matrix *A, int *heavy_cols, long long *heavy_nnz);

/* ------------------------------------ multiple right-hand sides (no reference counterpart)
 * Y (m x k, row-major) = A * X  (accumulate = 0)  or  Y += A * X  (accumulate = 1), where X is
 * n x k, row-major, in the ORIGINAL column order (no permutation). k in {1,2,4,8,16,32,64};
 * val_kind FL_VAL_F64 (double X/Y, FP64 sums) or FL_VAL_F32 (float X/Y, FP32 sums).
 * The entries are re-ordered by row once (fl_matrix_prepare_rows); a few lanes own a row (k = 1:
 * four lanes, each summing every fourth entry; k >= 2: each lane gathers up to 16 bytes of an X
 * row, 16-byte lanes need X and Y 16-byte aligned), and each Y row is written once, without
 * atomics. Results do not depend on the lane width or on the other rows of the matrix. */
int fl_matrix_prepare_rows(fl_matrix *A, void *stream);
int fl_spmm(fl_matrix *A, int val_kind, int k, const void *d_X, void *d_Y, int accumulate, void *stream);

/* ------------------------------------------ iterative solver (no reference counterpart)
 * Conjugate gradients for A x = b on a square, symmetric positive definite device-resident matrix
 * -- the caller that repeats the SpMV (the reference repeats it only to time it,
 * src/FastLoad_GPU.h:349-374). The product uses one prepared schedule: kernel as for fl_spmv, or
 * FL_KERNEL_ROW (the row-ordered kernel, after fl_matrix_prepare_rows); val_kind FL_VAL_F64
 * (double b, x and vectors) or FL_VAL_F32 (float; dot products summed in FP64). A must outlive the
 * solver. fl_cg_solve starts from the x it is given, runs iterations in groups of check_every
 * (rounded up to even) captured once as a hipGraph, and after each group stops when
 * ||r|| / ||b|| <= rtol or max_iter iterations have run: *iters receives the iterations run (a
 * multiple of the group), *rel_res the final ||r|| / ||b|| (recursive residual). */
typedef struct fl_cg fl_cg;
#define FL_KERNEL_ROW (-1)
int fl_cg_create(fl_cg **cg, fl_matrix *A, int kernel, int val_kind);
int fl_cg_solve(fl_cg *cg, const void *d_b, void *d_x, int max_iter, double rtol, int check_every, int *iters,
                double *rel_res);
int fl_cg_destroy(fl_cg *cg);

/* --------------------------------------------- batches of independent matrices
 * Block-diagonal assembly A = diag(A_0, ..., A_{count-1}) of host CSC matrices: one SpMV/SpMM
 * with A computes every y_i = A_i x_i, so a whole batch runs through the same preprocessing and
 * kernels in one launch per phase. Outputs are malloc'd (fl_free); row_off/col_off have
 * count+1 entries (y_i = y[row_off[i] .. row_off[i+1]), x_i likewise). */
int fl_batch_assemble(int count, const int *m, const int *n, const int *const *colptr, const int *const *rowidx,
                      const double *const *val, int *M, int *N, int **colptr_out, int **rowidx_out,
                      double **val_out, int *row_off, int *col_off);

/* --------------------------------------------------------- hipGraph capture (plumbing)
 * Capture the kernels enqueued on a created stream between begin and end into an executable
 * graph; replay it with fl_graph_launch. The SpMV/SpMM entry points enqueue no memset nodes.
 * Prepare values/schedules and run once eagerly before capturing (first calls allocate). */
int fl_graph_begin_capture(void *stream);
int fl_graph_end_capture(void *stream, void **graph_exec);
int fl_graph_launch(void *graph_exec, void *stream);
int fl_graph_destroy(void *graph_exec);
/* Zero `bytes` of device memory with a kernel (capturable, unlike a memset node). */
int fl_zero_device(void *d, size_t bytes, void *stream);

/* ------------------------------------------ learned schedule selector (no reference counterpart)
 * The reference runs one fixed schedule; its nearest hook is the unused column statistics of
 * src/auxiliary.h:6-25 (mean, standard deviation, ratio of column lengths).
 *
 * fl_features: 40 features of a device CSC matrix, written to host `features`:
 *   [0..31] fraction of columns whose length falls in bin b (b = 0: empty; b >= 1: length in
 *           [2^(b-1), 2^b); bin 31 open-ended);
 *   [32] log2(m+1)/24  [33] log2(n+1)/24  [34] log2(nnz+1)/32  [35] log2(mean row length+1)/8
 *   [36] coefficient of variation of row lengths  [37] log2(max/mean row length + 1)/8
 *   [38] fraction of empty rows  [39] mean over non-empty columns of (last row - first row)/m.
 * Rows must be ascending inside each column (as fl_matrix_create requires). */
int fl_features(int m, int n, int nnz, const int *d_colptr, const int *d_rowidx, float *features, void *stream);

/* A 3-layer perceptron in -> h1 -> h2 -> out (ReLU hidden layers, logits out) whose matrix
 * products run on MFMA with BF16 operands and FP32 accumulation; FP32 master weights, Adam.
 * Parameters (fl_mlp_get_params / set_params / gradients order): W1 (in x h1, row-major), b1,
 * W2, b2, W3, b3 -- fl_mlp_param_count floats. Inputs d_X (batch x in) and soft targets d_T
 * (batch x out, rows summing to 1) are device FP32, row-major. Loss: mean over the batch of
 * -sum_c T log softmax(logits). */
typedef struct fl_mlp fl_mlp;
int fl_mlp_create(fl_mlp **out, int n_in, int n_h1, int n_h2, int n_out, uint64_t seed);
int fl_mlp_destroy(fl_mlp *P);
int fl_mlp_param_count(const fl_mlp *P);
int fl_mlp_get_params(const fl_mlp *P, float *host);
int fl_mlp_set_params(fl_mlp *P, const float *host);
/* logits (batch x out) to device d_logits */
int fl_mlp_forward(fl_mlp *P, const float *d_X, int batch, float *d_logits, void *stream);
/* argmax class per row to host classes[batch], its softmax probability to host prob (optional) */
int fl_mlp_predict(fl_mlp *P, const float *d_X, int batch, int *classes, float *prob, void *stream);
/* gradients of the loss at the current parameters to host (param_count floats); loss optional */
int fl_mlp_gradients(fl_mlp *P, const float *d_X, const float *d_T, int batch, float *host_grads, float *loss,
                     void *stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8, learning rate lr); loss before the step */
int fl_mlp_train_step(fl_mlp *P, const float *d_X, const float *d_T, int batch, float lr, float *loss, void *stream);

/* The 13 schedules the selector chooses from: class -> tile width (fl_matrix_create), kernel
 * (fl_spmv; FL_KERNEL_ROW = fl_spmm with k = 1 after fl_matrix_prepare_rows), batch tiles and
 * band rows (fl_matrix_prepare_schedule; for FL_KERNEL_HYB those of its PB sub-matrix, heavy
 * columns min(n, 32768)). Class 0 is the reference schedule. */
#define FL_SCHEDULES 13
int fl_schedule_info(int cls, int *width, int *kernel, int *batch_tiles, int *band_rows);
/* fl_features + fl_mlp_predict on a device CSC matrix: the network's class and its probability.
 * Falls back to class 0 (probability 0) when net is NULL, the matrix has no non-zeros or a
 * feature is not finite. */
int fl_select_schedule(fl_mlp *net, int m, int n, int nnz, const int *d_colptr, const int *d_rowidx, int *cls,
                       float *prob, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTLOAD_H */
