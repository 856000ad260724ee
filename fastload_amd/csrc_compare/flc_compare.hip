// Comparators for measurement only (libfastload_compare.so; include/fastload_compare.h): the
// reference's naive CSC kernel and rocSPARSE's CSR SpMV, timed beside FastLoad as the reference
// times its naive kernel and cuSPARSE (src/main.cu:289-311). libfastload.so never calls them.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>

#include "../../include/fastload_compare.h"

namespace {

thread_local char g_err[512] = {0};

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

struct Failed {
    int rc;
};

#define FLC_HIP(call)                                                                                         \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) {                                                                               \
            set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));              \
            throw Failed{FL_ERR_HIP};                                                                         \
        }                                                                                                     \
    } while (0)

#define FLC_SP(call)                                                                                          \
    do {                                                                                                      \
        rocsparse_status st_ = (call);                                                                        \
        if (st_ != rocsparse_status_success) {                                                                \
            set_error("%s:%d: %s failed: rocsparse status %d", __FILE__, __LINE__, #call, (int)st_);          \
            throw Failed{FL_ERR_HIP};                                                                         \
        }                                                                                                     \
    } while (0)

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t count) { FLC_HIP(hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1))); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

// src/CSCSpMV.h:4-27: one lane per column (grid-stride), one atomic per entry.
template <typename T>
__global__ void k_csc_naive(int n, const int *__restrict__ colptr, const int *__restrict__ rowidx,
                            const T *__restrict__ val, const T *__restrict__ x, T *__restrict__ y) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const T xc = x[c];
        for (int j = colptr[c]; j < colptr[c + 1]; ++j) unsafeAtomicAdd(&y[rowidx[j]], val[j] * xc);
    }
}

// The reference's launch shape: 32 blocks per multiprocessor (compute unit), 64 threads each.
template <typename T>
void launch_naive(int n, const int *colptr, const int *rowidx, const T *val, const T *x, T *y, hipStream_t s) {
    if (n <= 0) return;
    int dev = 0, cus = 0;
    FLC_HIP(hipGetDevice(&dev));
    FLC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    k_csc_naive<T><<<32 * std::max(cus, 1), 64, 0, s>>>(n, colptr, rowidx, val, x, y);
    FLC_HIP(hipGetLastError());
}

bool ptr_ok(const int *p, int count, int nnz) {
    if (p[0] != 0 || p[count] != nnz) return false;
    for (int i = 0; i < count; ++i)
        if (p[i] > p[i + 1]) return false;
    return true;
}

bool idx_ok(const int *idx, int nnz, int limit) {
    for (int i = 0; i < nnz; ++i)
        if (idx[i] < 0 || idx[i] >= limit) return false;
    return true;
}

}  // namespace

struct flc_rocsparse {
    rocsparse_handle handle = nullptr;
    rocsparse_mat_descr descr = nullptr;
    rocsparse_mat_info info = nullptr;
    int rows = 0, cols = 0, nnz = 0, fp64 = 0;
    const int *rowptr = nullptr, *col = nullptr;
    const void *val = nullptr;
    ~flc_rocsparse() {
        if (info) (void)rocsparse_destroy_mat_info(info);
        if (descr) (void)rocsparse_destroy_mat_descr(descr);
        if (handle) (void)rocsparse_destroy_handle(handle);
    }
};

extern "C" {

const char *flc_last_error(void) { return g_err; }

int flc_csc_spmv(const char *filename, int m, int n, int nnz, const int *csc_ptr, const int *csc_rowidx,
                 const double *csc_val, const double *x, double *y, const double *y_golden, double *time_ms) {
    try {
        if (m < 0 || n < 0 || nnz < 0 || !csc_ptr || !x || !y || (nnz && (!csc_rowidx || !csc_val))) {
            set_error("flc_csc_spmv: NULL pointer or negative size");
            return FL_ERR_INVALID;
        }
        if (!ptr_ok(csc_ptr, n, nnz) || !idx_ok(csc_rowidx, nnz, m)) {
            set_error("flc_csc_spmv: column pointers or row indices invalid");
            return FL_ERR_INVALID;
        }
        hipStream_t s = nullptr;
        Dev<int> dp((size_t)n + 1), dr(nnz);
        Dev<double> dv(nnz), dx(n), dy(m);
        FLC_HIP(hipMemcpy(dp.p, csc_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
        if (nnz) {
            FLC_HIP(hipMemcpy(dr.p, csc_rowidx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
            FLC_HIP(hipMemcpy(dv.p, csc_val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
        }
        if (n) FLC_HIP(hipMemcpy(dx.p, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
        double total = 0.0;
        const int reps = 1000;  // src/CSCSpMV.h:80
        for (int r = 0; r < reps; ++r) {
            FLC_HIP(hipMemsetAsync(dy.p, 0, sizeof(double) * std::max(m, 1), s));
            FLC_HIP(hipStreamSynchronize(s));
            const auto t0 = std::chrono::steady_clock::now();
            launch_naive<double>(n, dp.p, dr.p, dv.p, dx.p, dy.p, s);
            FLC_HIP(hipStreamSynchronize(s));
            total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        const double mean = total / reps;
        if (m) FLC_HIP(hipMemcpy(y, dy.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost));
        if (time_ms) *time_ms = mean;
        if (!y_golden) return FL_OK;
        int errors = 0;  // src/CSCSpMV.h:97-100
        for (int i = 0; i < m; ++i)
            if (std::fabs(y_golden[i] - y[i]) > 0.01 * std::fabs(y[i])) ++errors;
        const double gflops = mean > 0 ? 2.0 * (double)nnz * 1.0e-6 / mean : 0.0;
        if (errors == 0) {
            printf("BASIC CSC GPU PASS! (%f ms %4.2f GFlops)%s%s\n", mean, gflops, filename ? " " : "",
                   filename ? filename : "");
        } else {
            printf("BASIC CSC check NO PASS! error_count_cuda = %d \n", errors);
        }
        fflush(stdout);
        return errors ? 1 : FL_OK;
    } catch (const Failed &f) {
        return f.rc;
    } catch (const std::bad_alloc &) {
        set_error("flc_csc_spmv: host allocation failed");
        return FL_ERR_ALLOC;
    }
}

int flc_csc_spmv_device(int m, int n, const int *d_colptr, const int *d_rowidx, const void *d_val, const void *d_x,
                        void *d_y, int fp64, void *stream) {
    try {
        (void)m;
        if (n < 0 || (n && (!d_colptr || !d_rowidx || !d_val || !d_x || !d_y))) {
            set_error("flc_csc_spmv_device: NULL pointer or negative size");
            return FL_ERR_INVALID;
        }
        hipStream_t s = (hipStream_t)stream;
        if (fp64)
            launch_naive<double>(n, d_colptr, d_rowidx, (const double *)d_val, (const double *)d_x, (double *)d_y, s);
        else
            launch_naive<float>(n, d_colptr, d_rowidx, (const float *)d_val, (const float *)d_x, (float *)d_y, s);
        return FL_OK;
    } catch (const Failed &f) {
        return f.rc;
    }
}

int flc_rocsparse_create(flc_rocsparse **out, int rows, int cols, int nnz, const int *d_rowptr, const int *d_col,
                         const void *d_val, int fp64, void *stream) {
    try {
        if (!out || rows < 0 || cols < 0 || nnz < 0 || !d_rowptr || (nnz && (!d_col || !d_val))) {
            set_error("flc_rocsparse_create: NULL pointer or negative size");
            return FL_ERR_INVALID;
        }
        *out = nullptr;
        flc_rocsparse *h = new flc_rocsparse();
        try {
            h->rows = rows;
            h->cols = cols;
            h->nnz = nnz;
            h->fp64 = fp64;
            h->rowptr = d_rowptr;
            h->col = d_col;
            h->val = d_val;
            FLC_SP(rocsparse_create_handle(&h->handle));
            FLC_SP(rocsparse_set_stream(h->handle, (hipStream_t)stream));
            FLC_SP(rocsparse_create_mat_descr(&h->descr));
            FLC_SP(rocsparse_create_mat_info(&h->info));
            if (fp64)
                FLC_SP(rocsparse_dcsrmv_analysis(h->handle, rocsparse_operation_none, rows, cols, nnz, h->descr,
                                                 (const double *)d_val, d_rowptr, d_col, h->info));
            else
                FLC_SP(rocsparse_scsrmv_analysis(h->handle, rocsparse_operation_none, rows, cols, nnz, h->descr,
                                                 (const float *)d_val, d_rowptr, d_col, h->info));
            FLC_HIP(hipStreamSynchronize((hipStream_t)stream));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return FL_OK;
    } catch (const Failed &f) {
        return f.rc;
    } catch (const std::bad_alloc &) {
        set_error("flc_rocsparse_create: host allocation failed");
        return FL_ERR_ALLOC;
    }
}

int flc_rocsparse_run(flc_rocsparse *h, const void *d_x, void *d_y, void *stream) {
    try {
        if (!h || (h->cols && !d_x) || (h->rows && !d_y)) {
            set_error("flc_rocsparse_run: NULL");
            return FL_ERR_INVALID;
        }
        FLC_SP(rocsparse_set_stream(h->handle, (hipStream_t)stream));
        if (h->fp64) {
            const double alpha = 1.0, beta = 0.0;
            FLC_SP(rocsparse_dcsrmv(h->handle, rocsparse_operation_none, h->rows, h->cols, h->nnz, &alpha, h->descr,
                                    (const double *)h->val, h->rowptr, h->col, h->info, (const double *)d_x, &beta,
                                    (double *)d_y));
        } else {
            const float alpha = 1.f, beta = 0.f;
            FLC_SP(rocsparse_scsrmv(h->handle, rocsparse_operation_none, h->rows, h->cols, h->nnz, &alpha, h->descr,
                                    (const float *)h->val, h->rowptr, h->col, h->info, (const float *)d_x, &beta,
                                    (float *)d_y));
        }
        return FL_OK;
    } catch (const Failed &f) {
        return f.rc;
    }
}

int flc_rocsparse_destroy(flc_rocsparse *h) {
    delete h;
    return FL_OK;
}

int flc_rocsparse_spmv(const char *filename, int nnz, int rows, int cols, const int *csr_ptr, const int *csr_col,
                       const double *csr_val, const double *x, double *y, const double *y_result, double *time_ms) {
    try {
        if (rows < 0 || cols < 0 || nnz < 0 || !csr_ptr || !x || !y || (nnz && (!csr_col || !csr_val))) {
            set_error("flc_rocsparse_spmv: NULL pointer or negative size");
            return FL_ERR_INVALID;
        }
        if (!ptr_ok(csr_ptr, rows, nnz) || !idx_ok(csr_col, nnz, cols)) {
            set_error("flc_rocsparse_spmv: row pointers or column indices invalid");
            return FL_ERR_INVALID;
        }
        Dev<int> dp((size_t)rows + 1), dc(nnz);
        Dev<double> dv(nnz), dx(cols), dy(rows);
        FLC_HIP(hipMemcpy(dp.p, csr_ptr, sizeof(int) * ((size_t)rows + 1), hipMemcpyHostToDevice));
        if (nnz) {
            FLC_HIP(hipMemcpy(dc.p, csr_col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
            FLC_HIP(hipMemcpy(dv.p, csr_val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice));
        }
        if (cols) FLC_HIP(hipMemcpy(dx.p, x, sizeof(double) * (size_t)cols, hipMemcpyHostToDevice));
        // src/cusparse_cuda.h:84-101: the timed region holds the library set-up and one SpMV
        const auto t0 = std::chrono::steady_clock::now();
        flc_rocsparse *h = nullptr;
        int rc = flc_rocsparse_create(&h, rows, cols, nnz, dp.p, dc.p, dv.p, 1, nullptr);
        if (rc) return rc;
        rc = flc_rocsparse_run(h, dx.p, dy.p, nullptr);
        FLC_HIP(hipDeviceSynchronize());
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        flc_rocsparse_destroy(h);
        if (rc) return rc;
        if (rows) FLC_HIP(hipMemcpy(y, dy.p, sizeof(double) * (size_t)rows, hipMemcpyDeviceToHost));
        if (time_ms) *time_ms = ms;
        if (!y_result) return FL_OK;
        bool correct = true;  // src/cusparse_cuda.h:121-127: exact comparison
        for (int i = 0; i < rows && correct; ++i) correct = y[i] == y_result[i];
        const double gflops = ms > 0 ? 2.0 * (double)nnz * 1.0e-6 / ms : 0.0;
        if (correct) printf("rocSPARSE test PASSED! (%f ms %4.2f GFlops)%s%s\n", ms, gflops, filename ? " " : "",
                            filename ? filename : "");
        else printf("rocSPARSE test FAILED: wrong result\n");
        fflush(stdout);
        return correct ? FL_OK : 1;
    } catch (const Failed &f) {
        return f.rc;
    } catch (const std::bad_alloc &) {
        set_error("flc_rocsparse_spmv: host allocation failed");
        return FL_ERR_ALLOC;
    }
}

}  // extern "C"
