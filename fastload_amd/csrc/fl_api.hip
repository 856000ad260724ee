// extern "C" entry points of libfastload.so (declared in include/fastload.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fl_matrix.h"

namespace fl {
void csr2csc_device(int m, int n, int nnz, const int *d_rowptr, const int *d_colind, const double *d_val,
                    double *d_cscval, int *d_cscrow, int *d_colptr, int *d_nnzpercol, hipStream_t s);
void sorted_copy(int n, const int *d_perm, const int *d_sortnnz, const int *d_P, const int *d_colptr,
                 const int *d_rowidx, const double *d_val, int *d_tmp_row, double *d_tmp_val, hipStream_t s);
int check_compressed(const int *d_ptr, int count, const int *d_idx, int nnz, int limit, hipStream_t s);
void prepare_values(fl_matrix *A, int kind, hipStream_t s);
void permute_x(const fl_matrix *A, const void *x, void *sx, bool dbl, hipStream_t s);
void permute_raw(const int *perm, int n, const void *x, void *sx, bool dbl, hipStream_t s);
void win_prepare(fl_matrix *A, int bt, bool rows, hipStream_t s);
void pb_prepare(fl_matrix *A, int bt, int R, hipStream_t s);

template <typename T>
static void upload(DevBuf<T> &d, const T *h, size_t count, hipStream_t s) {
    d.alloc(std::max<size_t>(count, 1));
    if (count) FL_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

template <typename T>
static void download(T *h, const T *d, size_t count, hipStream_t s) {
    if (count && h) FL_HIP(hipMemcpyAsync(h, d, sizeof(T) * count, hipMemcpyDeviceToHost, s));
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        FL_HIP(hipEventCreate(&a));
        FL_HIP(hipEventCreate(&b));
        FL_HIP(hipEventRecord(a, s));
    }
    float stop() {
        FL_HIP(hipEventRecord(b, s));
        FL_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        FL_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

static int invalid(const char *msg) {
    set_error("%s", msg);
    return FL_ERR_INVALID;
}

static bool check_ptr_host(const int *ptr, int count, int nnz) {
    if (ptr[0] != 0 || ptr[count] != nnz) return false;
    for (int i = 0; i < count; ++i)
        if (ptr[i] > ptr[i + 1]) return false;
    return true;
}

static bool check_idx_host(const int *idx, int nnz, int limit) {
    for (int i = 0; i < nnz; ++i)
        if (idx[i] < 0 || idx[i] >= limit) return false;
    return true;
}

template <typename T>
static T *host_copy(const DevBuf<T> &d, size_t count, hipStream_t s) {
    T *h = (T *)malloc(sizeof(T) * std::max<size_t>(count, 1));
    if (!h) throw std::bad_alloc();
    download(h, d.p, count, s);
    return h;
}

}  // namespace fl

using namespace fl;

extern "C" {

// ------------------------------------------------------------ reference host API
int fl_csr_to_csc(double *time_ms, int m, int n, int nnz, const double *csr_val, const int *csr_row_ptr,
                  const int *csr_col_ind, double *csc_val, int *csc_row_ind, int *csc_col_ptr, int *nnzpercol) {
    FL_API_BEGIN
    if (m < 0 || n < 0 || nnz < 0) return invalid("fl_csr_to_csc: negative size");
    if (!csr_row_ptr || !csc_col_ptr || !nnzpercol || (nnz && (!csr_val || !csr_col_ind || !csc_val || !csc_row_ind)))
        return invalid("fl_csr_to_csc: NULL pointer");
    if (!check_ptr_host(csr_row_ptr, m, nnz)) return invalid("fl_csr_to_csc: csr_row_ptr is not a valid row pointer");
    if (!check_idx_host(csr_col_ind, nnz, n)) {
        set_error("fl_csr_to_csc: column index out of range");
        return FL_ERR_RANGE;
    }
    hipStream_t s = nullptr;
    DevBuf<int> rp, ci, crow, cptr((size_t)n + 1), npc(std::max(n, 1));
    DevBuf<double> v, cv(std::max(nnz, 1));
    upload(rp, csr_row_ptr, (size_t)m + 1, s);
    upload(ci, csr_col_ind, nnz, s);
    upload(v, csr_val, nnz, s);
    crow.alloc(std::max(nnz, 1));
    FL_HIP(hipStreamSynchronize(s));
    Timer tm(s);
    csr2csc_device(m, n, nnz, rp.p, ci.p, v.p, cv.p, crow.p, cptr.p, npc.p, s);
    const float ms = tm.stop();
    download(csc_val, cv.p, nnz, s);
    download(csc_row_ind, crow.p, nnz, s);
    download(csc_col_ptr, cptr.p, (size_t)n + 1, s);
    download(nnzpercol, npc.p, n, s);
    FL_HIP(hipStreamSynchronize(s));
    if (time_ms) *time_ms = ms;
    return FL_OK;
    FL_API_END
}

int fl_ColSort(double *time_ms, int rowA, int colA, int nnz, const int *nnzPerCol, const int *cscColPtr,
               const int *cscRowIdx, const double *cscVal, int *sortrowidxTmp, double *sortvalTmp, int *sortnnzTmp,
               const double *x, double *sortx) {
    FL_API_BEGIN
    (void)rowA;
    if (colA < 0 || nnz < 0) return invalid("fl_ColSort: negative size");
    if (!nnzPerCol || !cscColPtr || !sortnnzTmp || !x || !sortx || (nnz && (!cscRowIdx || !cscVal || !sortrowidxTmp || !sortvalTmp)))
        return invalid("fl_ColSort: NULL pointer");
    if (!check_ptr_host(cscColPtr, colA, nnz)) return invalid("fl_ColSort: cscColPtr is not a valid column pointer");
    int maxk = 0;
    for (int c = 0; c < colA; ++c) {
        if (nnzPerCol[c] != cscColPtr[c + 1] - cscColPtr[c]) return invalid("fl_ColSort: nnzPerCol disagrees with cscColPtr");
        maxk = std::max(maxk, nnzPerCol[c]);
    }
    hipStream_t s = nullptr;
    DevBuf<int> cnt, cp, ri, perm(std::max(colA, 1)), sn(std::max(colA, 1)), P((size_t)colA + 1), trow(std::max(nnz, 1));
    DevBuf<double> cv, dx, dsx(std::max(colA, 1)), tval(std::max(nnz, 1));
    upload(cnt, nnzPerCol, colA, s);
    upload(cp, cscColPtr, (size_t)colA + 1, s);
    upload(ri, cscRowIdx, nnz, s);
    upload(cv, cscVal, nnz, s);
    upload(dx, x, colA, s);
    FL_HIP(hipStreamSynchronize(s));
    Timer tm(s);
    colsort_device(colA, cnt.p, perm.p, sn.p, maxk, s);
    exclusive_scan_i32(sn.p, P.p, colA, s);
    sorted_copy(colA, perm.p, sn.p, P.p, cp.p, ri.p, cv.p, trow.p, tval.p, s);
    permute_raw(perm.p, colA, dx.p, dsx.p, true, s);
    const float ms = tm.stop();
    download(sortrowidxTmp, trow.p, nnz, s);
    download(sortvalTmp, tval.p, nnz, s);
    download(sortnnzTmp, sn.p, colA, s);
    download(sortx, dsx.p, colA, s);
    FL_HIP(hipStreamSynchronize(s));
    if (time_ms) *time_ms = ms;
    return FL_OK;
    FL_API_END
}

int fl_formatTransform(double *timeFormatTran, double *timeFormatclas, slide_matrix *matrix,
                       const int *sortrowidx_tmp, const double *sortval_tmp, const int *sortnnz_tmp, int nnz, int rowA,
                       int colA, int *h_count) {
    FL_API_BEGIN
    (void)rowA;
    if (!matrix || !sortnnz_tmp || (nnz && (!sortrowidx_tmp || !sortval_tmp))) return invalid("fl_formatTransform: NULL pointer");
    if (colA < 0 || nnz < 0) return invalid("fl_formatTransform: negative size");
    long long sum = 0;
    for (int i = 0; i < colA; ++i) {
        if (sortnnz_tmp[i] < 0 || (i && sortnnz_tmp[i] < sortnnz_tmp[i - 1]))
            return invalid("fl_formatTransform: sortnnz_tmp must be non-negative and ascending");
        sum += sortnnz_tmp[i];
    }
    if (sum != nnz) return invalid("fl_formatTransform: sum(sortnnz_tmp) != nnz");
    memset(matrix, 0, sizeof(*matrix));
    hipStream_t s = nullptr;
    fl_matrix A;
    A.m = rowA;
    A.n = colA;
    A.nnz = nnz;
    A.W = 32;
    DevBuf<int> sn, trow, P((size_t)colA + 1);
    DevBuf<double> tval;
    upload(sn, sortnnz_tmp, colA, s);
    upload(trow, sortrowidx_tmp, nnz, s);
    upload(tval, sortval_tmp, nnz, s);
    ElemSource src;
    src.rowidx = trow.p;
    src.val = tval.p;
    FL_HIP(hipStreamSynchronize(s));
    Timer tm(s);
    exclusive_scan_i32(sn.p, P.p, colA, s);
    src.P = P.p;
    float t_tran = 0.f, t_clas = 0.f;
    transform_device(&A, sn.p, src, s, &t_tran, &t_clas);
    const float scan_ms = tm.stop() - t_tran - t_clas;
    const int T = A.tilenum;
    matrix->tile_ptr = host_copy(A.tile_ptr, (size_t)T + 1, s);
    matrix->tile_len = host_copy(A.tile_len, T, s);
    matrix->tile_colidx = host_copy(A.tile_colidx, T, s);
    matrix->tile_format = host_copy(A.tile_format, T, s);
    matrix->sortval = host_copy(A.val64, nnz, s);
    matrix->sortrowidx = host_copy(A.sortrowidx, nnz, s);
    matrix->sortrowindex = host_copy(A.sortrowindex, nnz, s);
    matrix->countrow = host_copy(A.countrow, (size_t)T + 1, s);
    matrix->segmentoffset = host_copy(A.segmentoffset, (size_t)A.segsum + 1, s);
    matrix->tile_tall = nullptr;
    matrix->tilenum = T;
    matrix->segsum = A.segsum;
    FL_HIP(hipStreamSynchronize(s));
    if (h_count) *h_count = A.L;
    if (timeFormatTran) *timeFormatTran = t_tran + std::max(scan_ms, 0.f);
    if (timeFormatclas) *timeFormatclas = t_clas;
    return FL_OK;
    FL_API_END
}

void fl_Tile_destroy(slide_matrix *M) {
    if (!M) return;
    free(M->sortval);
    free(M->sortrowidx);
    free(M->tile_ptr);
    free(M->tile_len);
    free(M->tile_colidx);
    free(M->tile_tall);
    free(M->tile_format);
    free(M->countrow);
    free(M->segmentoffset);
    free(M->sortrowindex);
    memset(M, 0, sizeof(*M));
}

int fl_FastLoad_spmv(const char *filename, const slide_matrix *M, int nnz, int colA, int rowA, const double *x,
                     double *y, const double *y_golden, double *time_ms) {
    FL_API_BEGIN
    if (!M || !x || !y) return invalid("fl_FastLoad_spmv: NULL pointer");
    const int T = M->tilenum;
    if (T < 0 || nnz < 0 || colA < 0 || rowA < 0) return invalid("fl_FastLoad_spmv: negative size");
    if (T > 0 && (!M->tile_ptr || !M->tile_len || !M->tile_colidx || !M->tile_format || !M->sortrowidx || !M->sortval ||
                  !M->sortrowindex))
        return invalid("fl_FastLoad_spmv: slide_matrix arrays missing");
    // Host-side validation of the layout so that no kernel can read or write out of bounds.
    int W = 32;
    if (T > 0) {
        if (!check_ptr_host(M->tile_ptr, T, nnz)) return invalid("fl_FastLoad_spmv: tile_ptr invalid");
        int maxlen = 0;
        for (int t = 0; t < T; ++t) {
            const int len = M->tile_len[t];
            if (len < 0 || M->tile_ptr[t + 1] - M->tile_ptr[t] != len || M->tile_colidx[t] < 0 ||
                (long long)M->tile_colidx[t] + len > colA || M->tile_format[t] < 0 || M->tile_format[t] > 2)
                return invalid("fl_FastLoad_spmv: tile metadata invalid");
            maxlen = std::max(maxlen, len);
        }
        if (maxlen > 64) return invalid("fl_FastLoad_spmv: tiles wider than 64");
        W = maxlen > 32 ? 64 : 32;
        if (!check_idx_host(M->sortrowidx, nnz, rowA)) {
            set_error("fl_FastLoad_spmv: row index out of range");
            return FL_ERR_RANGE;
        }
    }
    hipStream_t s = nullptr;
    fl_matrix A;
    A.m = rowA;
    A.n = colA;
    A.nnz = nnz;
    A.W = W;
    A.tilenum = T;
    upload(A.tile_ptr, M->tile_ptr, (size_t)T + 1, s);
    upload(A.tile_len, M->tile_len, T, s);
    upload(A.tile_colidx, M->tile_colidx, T, s);
    upload(A.tile_format, M->tile_format, T, s);
    upload(A.sortrowidx, M->sortrowidx, nnz, s);
    upload(A.sortrowindex, M->sortrowindex, nnz, s);
    upload(A.val64, M->sortval, nnz, s);
    DevBuf<double> dx, dy(std::max(rowA, 1));
    upload(dx, x, colA, s);
    // dry run (FastLoad_GPU.h:334-344)
    FL_HIP(hipMemsetAsync(dy.p, 0, sizeof(double) * std::max(rowA, 1), s));
    int rc = spmv(&A, FL_KERNEL_DROPIN, FL_VAL_F64, dx.p, dy.p, nullptr, false, s);
    if (rc) return rc;
    // FastLoad_GPU.h:349-374: 100 x (clear y, then the host clock around launch + synchronize)
    double total = 0.0;
    const int reps = 100;
    for (int r = 0; r < reps; ++r) {
        FL_HIP(hipMemsetAsync(dy.p, 0, sizeof(double) * std::max(rowA, 1), s));
        FL_HIP(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        rc = spmv(&A, FL_KERNEL_DROPIN, FL_VAL_F64, dx.p, dy.p, nullptr, false, s);
        FL_HIP(hipStreamSynchronize(s));
        total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc) return rc;
    }
    const double mean = total / reps;
    download(y, dy.p, rowA, s);
    FL_HIP(hipStreamSynchronize(s));
    if (time_ms) *time_ms = mean;
    const double gflops = mean > 0 ? 2.0 * (double)nnz * 1.0e-6 / mean : 0.0;
    if (!y_golden) return FL_OK;
    int errors = 0;  // FastLoad_GPU.h:383-389: |golden - y| > 1% |y|
    for (int i = 0; i < rowA; ++i)
        if (std::fabs(y_golden[i] - y[i]) > 0.01 * std::fabs(y[i])) ++errors;
    if (errors == 0) {
        printf("Check FASTLOAD GPU PASS! (%f ms , %4.2f GFlops)%s%s\n", mean, gflops, filename ? " " : "",
               filename ? filename : "");
        return FL_OK;
    }
    printf("Check NO PASS! error_count_slide_cuda = %d \n", errors);
    return 1;
    FL_API_END
}

// ---------------------------------------------------------------- device memory helpers
int fl_device_count(int *count) {
    FL_API_BEGIN
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (count) *count = c;
    return FL_OK;
    FL_API_END
}

int fl_malloc(void **dptr, size_t bytes) {
    FL_API_BEGIN
    if (!dptr) return invalid("fl_malloc: NULL");
    *dptr = nullptr;
    FL_HIP(hipMalloc(dptr, std::max<size_t>(bytes, 1)));
    return FL_OK;
    FL_API_END
}

int fl_free_device(void *dptr) {
    FL_API_BEGIN
    if (dptr) FL_HIP(hipFree(dptr));
    return FL_OK;
    FL_API_END
}

int fl_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    FL_API_BEGIN
    if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FL_OK;
    FL_API_END
}

int fl_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    FL_API_BEGIN
    if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FL_OK;
    FL_API_END
}

int fl_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    FL_API_BEGIN
    if (bytes) FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return FL_OK;
    FL_API_END
}

int fl_memset_device(void *dptr, int value, size_t bytes) {
    FL_API_BEGIN
    if (bytes) FL_HIP(hipMemset(dptr, value, bytes));
    return FL_OK;
    FL_API_END
}

int fl_synchronize(void) {
    FL_API_BEGIN
    FL_HIP(hipDeviceSynchronize());
    return FL_OK;
    FL_API_END
}

int fl_stream_create(void **stream) {
    FL_API_BEGIN
    hipStream_t st;
    FL_HIP(hipStreamCreate(&st));
    *stream = st;
    return FL_OK;
    FL_API_END
}

int fl_stream_destroy(void *stream) {
    FL_API_BEGIN
    if (stream) FL_HIP(hipStreamDestroy((hipStream_t)stream));
    return FL_OK;
    FL_API_END
}

int fl_stream_synchronize(void *stream) {
    FL_API_BEGIN
    FL_HIP(hipStreamSynchronize((hipStream_t)stream));
    return FL_OK;
    FL_API_END
}

int fl_event_create(void **ev) {
    FL_API_BEGIN
    hipEvent_t e;
    FL_HIP(hipEventCreate(&e));
    *ev = e;
    return FL_OK;
    FL_API_END
}

int fl_event_destroy(void *ev) {
    FL_API_BEGIN
    if (ev) FL_HIP(hipEventDestroy((hipEvent_t)ev));
    return FL_OK;
    FL_API_END
}

int fl_event_record(void *ev, void *stream) {
    FL_API_BEGIN
    FL_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return FL_OK;
    FL_API_END
}

int fl_event_elapsed(float *ms, void *start, void *stop) {
    FL_API_BEGIN
    FL_HIP(hipEventSynchronize((hipEvent_t)stop));
    FL_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return FL_OK;
    FL_API_END
}

// ------------------------------------------------------------ device-resident matrix
__global__ void k_col_counts(const int *__restrict__ colptr, int n, int *__restrict__ counts) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        counts[i] = colptr[i + 1] - colptr[i];
}

int fl_matrix_create(fl_matrix **out, int m, int n, int nnz, const int *d_colptr, const int *d_rowidx,
                     const double *d_val, int width, void *stream) {
    FL_API_BEGIN
    if (!out) return invalid("fl_matrix_create: NULL handle");
    *out = nullptr;
    if (m < 0 || n < 0 || nnz < 0) return invalid("fl_matrix_create: negative size");
    if (width != 32 && width != 64) return invalid("fl_matrix_create: width must be 32 or 64");
    if (!d_colptr || (nnz && (!d_rowidx || !d_val))) return invalid("fl_matrix_create: NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    const int bad = check_compressed(d_colptr, n, d_rowidx, nnz, m, s);
    if (bad & 2) {
        set_error("fl_matrix_create: row index out of range");
        return FL_ERR_RANGE;
    }
    if (bad) return invalid("fl_matrix_create: colptr is not a valid column pointer");
    fl_matrix *A = new fl_matrix();
    try {
        A->m = m;
        A->n = n;
        A->nnz = nnz;
        A->W = width;
        DevBuf<int> counts(std::max(n, 1));
        A->perm.alloc(std::max(n, 1));
        A->sortnnz.alloc(std::max(n, 1));
        if (n > 0) {
            k_col_counts<<<std::max(1u, std::min<unsigned>(grid_for(n), 65536u)), kBlock, 0, s>>>(d_colptr, n, counts.p);
            FL_LAUNCH_CHECK();
        }
        A->max_col_nnz = device_max_i32(counts.p, n, s);
        colsort_device(n, counts.p, A->perm.p, A->sortnnz.p, A->max_col_nnz, s);
        ElemSource src;
        src.colptr = d_colptr;
        src.perm = A->perm.p;
        src.rowidx = d_rowidx;
        src.val = d_val;
        transform_device(A, A->sortnnz.p, src, s, nullptr, nullptr);
    } catch (...) {
        delete A;
        throw;
    }
    *out = A;
    return FL_OK;
    FL_API_END
}

int fl_matrix_destroy(fl_matrix *A) {
    FL_API_BEGIN
    delete A;
    return FL_OK;
    FL_API_END
}

int fl_matrix_get_info(const fl_matrix *A, fl_matrix_info *info) {
    if (!A || !info) return invalid("fl_matrix_get_info: NULL");
    info->m = A->m;
    info->n = A->n;
    info->nnz = A->nnz;
    info->width = A->W;
    info->levels = A->L;
    info->tilenum = A->tilenum;
    info->segsum = A->segsum;
    info->max_col_nnz = A->max_col_nnz;
    return FL_OK;
}

int fl_matrix_export(const fl_matrix *A, int *perm, int *sortnnz, int *tile_ptr, int *tile_len, int *tile_colidx,
                     int *tile_format, int *sortrowidx, int *sortrowindex, double *sortval, int *countrow,
                     int *segmentoffset, int *lvl_col, int *lvl_base, int *lvl_tall) {
    FL_API_BEGIN
    if (!A) return invalid("fl_matrix_export: NULL");
    hipStream_t s = nullptr;
    const size_t T = A->tilenum, nz = A->nnz, n = A->n, L = A->L;
    download(perm, A->perm.p, n, s);
    download(sortnnz, A->sortnnz.p, n, s);
    download(tile_ptr, A->tile_ptr.p, T + 1, s);
    download(tile_len, A->tile_len.p, T, s);
    download(tile_colidx, A->tile_colidx.p, T, s);
    download(tile_format, A->tile_format.p, T, s);
    download(sortrowidx, A->sortrowidx.p, nz, s);
    download(sortrowindex, A->sortrowindex.p, nz, s);
    download(sortval, A->val64.p, nz, s);
    download(countrow, A->countrow.p, T + 1, s);
    download(segmentoffset, A->segmentoffset.p, (size_t)A->segsum + 1, s);
    download(lvl_col, A->lvl_col.p, L, s);
    download(lvl_base, A->lvl_base.p, L, s);
    download(lvl_tall, A->lvl_tall.p, L, s);
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
    FL_API_END
}

int fl_permute_x_f64(const fl_matrix *A, const double *d_x, double *d_sortx, void *stream) {
    FL_API_BEGIN
    if (!A) return invalid("fl_permute_x_f64: NULL");
    permute_x(A, d_x, d_sortx, true, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_permute_x_f32(const fl_matrix *A, const float *d_x, float *d_sortx, void *stream) {
    FL_API_BEGIN
    if (!A) return invalid("fl_permute_x_f32: NULL");
    permute_x(A, d_x, d_sortx, false, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_matrix_prepare_values(fl_matrix *A, int val_kind, void *stream) {
    FL_API_BEGIN
    if (!A) return invalid("fl_matrix_prepare_values: NULL");
    if (val_kind < FL_VAL_F64 || val_kind > FL_VAL_PACKED) return invalid("fl_matrix_prepare_values: bad kind");
    prepare_values(A, val_kind, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_matrix_pack_info(const fl_matrix *A, int kernel, int *mantissa_bits, int *exponent_bits, int *index_bits) {
    if (!A) return invalid("fl_matrix_pack_info: NULL");
    if (kernel == FL_KERNEL_HYB) {  // the PB words of the HYB schedule's sub-matrix
        if (!A->hyb_ready) return invalid("fl_matrix_pack_info: HYB schedule not prepared");
        return fl_matrix_pack_info(A->hyb_light.get(), FL_KERNEL_PB, mantissa_bits, exponent_bits, index_bits);
    }
    const bool pb = kernel == FL_KERNEL_PB;
    if (!pb && kernel != FL_KERNEL_WIN && kernel != FL_KERNEL_WINR)
        return invalid("fl_matrix_pack_info: kernel must be WIN, WINR, PB or HYB");
    if (!(pb ? A->pb_pack.p : A->win_pack.p) || (!pb && A->ord_kind != (kernel == FL_KERNEL_WINR ? 1 : 0)))
        return invalid("fl_matrix_pack_info: no packed values for this schedule");
    const PackFmt &f = pb ? A->pb_pf : A->win_pf;
    if (mantissa_bits) *mantissa_bits = f.m;
    if (exponent_bits) *exponent_bits = f.e;
    if (index_bits) *index_bits = f.ib;
    return FL_OK;
}

static int hyb_args(const fl_matrix *A, int heavy_cols, int batch_tiles, int band_rows, const char *who) {
    if (heavy_cols < 0 || heavy_cols > std::min(A->n, kHybMaxH32)) {
        set_error("%s: heavy_cols must be in [0, min(n, %d)]", who, kHybMaxH32);
        return FL_ERR_INVALID;
    }
    if (batch_tiles != 64 && batch_tiles != 128 && batch_tiles != 256) {
        set_error("%s: batch_tiles must be 64, 128 or 256", who);
        return FL_ERR_INVALID;
    }
    if (band_rows < 1024 || band_rows > 16384) {
        set_error("%s: band_rows must be in [1024, 16384]", who);
        return FL_ERR_INVALID;
    }
    return FL_OK;
}

int fl_matrix_prepare_hyb(fl_matrix *A, int heavy_cols, int batch_tiles, int band_rows, void *stream) {
    FL_API_BEGIN
    if (!A) return invalid("fl_matrix_prepare_hyb: NULL");
    const int rc = hyb_args(A, heavy_cols, batch_tiles, band_rows, "fl_matrix_prepare_hyb");
    if (rc) return rc;
    hyb_prepare(A, heavy_cols, batch_tiles, band_rows, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_matrix_hyb_info(const fl_matrix *A, int *heavy_cols, long long *heavy_nnz) {
    if (!A) return invalid("fl_matrix_hyb_info: NULL");
    if (!A->hyb_ready) return invalid("fl_matrix_hyb_info: HYB schedule not prepared");
    if (heavy_cols) *heavy_cols = A->hyb_H;
    if (heavy_nnz) *heavy_nnz = A->hyb_nnz;
    return FL_OK;
}

int fl_matrix_prepare_schedule(fl_matrix *A, int kernel, int batch_tiles, int band_rows, void *stream) {
    FL_API_BEGIN
    if (!A) return invalid("fl_matrix_prepare_schedule: NULL");
    if (kernel == FL_KERNEL_HYB) {
        const int H = std::min(A->n, 32768);
        const int rc = hyb_args(A, H, batch_tiles, band_rows, "fl_matrix_prepare_schedule");
        if (rc) return rc;
        hyb_prepare(A, H, batch_tiles, band_rows, (hipStream_t)stream);
        return FL_OK;
    }
    if (kernel != FL_KERNEL_WIN && kernel != FL_KERNEL_WINR && kernel != FL_KERNEL_PB)
        return invalid("fl_matrix_prepare_schedule: kernel must be WIN, WINR, PB or HYB");
    if (A->W != 64) return invalid("fl_matrix_prepare_schedule: needs a 64-wide tile layout");
    if (batch_tiles != 64 && batch_tiles != 128 && batch_tiles != 256)
        return invalid("fl_matrix_prepare_schedule: batch_tiles must be 64, 128 or 256");
    if (kernel == FL_KERNEL_PB && (band_rows < 1024 || band_rows > 16384))
        return invalid("fl_matrix_prepare_schedule: band_rows must be in [1024, 16384]");
    hipStream_t s = (hipStream_t)stream;
    if (kernel == FL_KERNEL_WIN || kernel == FL_KERNEL_WINR) win_prepare(A, batch_tiles, kernel == FL_KERNEL_WINR, s);
    else pb_prepare(A, batch_tiles, band_rows, s);
    return FL_OK;
    FL_API_END
}

int fl_spmv(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *d_work, void *stream) {
    FL_API_BEGIN
    if (!A || (A->tilenum && (!d_sortx || !d_y))) return invalid("fl_spmv: NULL");
    return spmv(A, kernel, val_kind, d_sortx, d_y, d_work, false, (hipStream_t)stream);
    FL_API_END
}

int fl_spmv_assign(fl_matrix *A, int kernel, int val_kind, const void *d_sortx, void *d_y, void *d_work, void *stream) {
    FL_API_BEGIN
    if (!A || (A->tilenum && !d_sortx) || (A->m && !d_y)) return invalid("fl_spmv_assign: NULL");
    return spmv(A, kernel, val_kind, d_sortx, d_y, d_work, true, (hipStream_t)stream);
    FL_API_END
}

}  // extern "C"
