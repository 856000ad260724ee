// Conjugate gradients on a device-resident FastLoad matrix: the iterative-solver caller of the
// SpMV (no reference counterpart; the reference's repeated-launch loop, src/FastLoad_GPU.h:349-374,
// multiplies the same x a hundred times). One iteration is
//   sp = p[perm]            (the column order of the tiles)
//   q  = A sp               (any prepared schedule, y = A x form)
//   alpha = rr / (p . q);   x += alpha p;   r -= alpha q
//   rr' = r . r;  beta = rr' / rr;  p = r + beta p
// as five kernels whose scalars never leave the device: the dot products are per-block partial
// sums in FP64 that the next kernel's blocks each reduce again (1024 partials, a few microseconds),
// so no host round trip and no atomics are needed. `check_every` iterations are captured once into
// a hipGraph and replayed; between replays the host reads the residual history of those iterations
// and stops at the first one below the tolerance. Vectors stay in the original order; values and
// vectors are FP64 (FL_VAL_F64) or FP32 (FL_VAL_F32, dot products still summed in FP64).
#include <cmath>
#include <vector>

#include "fl_matrix.h"

namespace fl {

void permute_x(const fl_matrix *A, const void *x, void *sx, bool dbl, hipStream_t s);
int spmm_row(const fl_matrix *A, int kind, int k, const void *X, void *Y, int accumulate, hipStream_t s);

namespace {

constexpr int kParts = 1024;  // blocks of the vector kernels = partial sums per dot product
constexpr int kThreads = 256;

// Sum of v over the block, returned to every thread.
__device__ double block_sum(double v, double *lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += lds[i];
    __syncthreads();
    return t;
}

__device__ double sum_parts(const double *__restrict__ part, double *lds) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kParts; i += blockDim.x) v += part[i];
    return block_sum(v, lds);
}

#define VSTRIDE(i, n) for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

// part[block] = sum over this block's elements of a*b
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dot(long long n, const T *__restrict__ a, const T *__restrict__ b,
                                                  double *__restrict__ part) {
    __shared__ double lds[kThreads / 64];
    double v = 0.0;
    VSTRIDE(i, n) v += (double)a[i] * (double)b[i];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// r = b - q, p = r, part = partial sums of r.r
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_init(long long n, const T *__restrict__ b, const T *__restrict__ q,
                                                      T *__restrict__ r, T *__restrict__ p, double *__restrict__ part) {
    __shared__ double lds[kThreads / 64];
    double v = 0.0;
    VSTRIDE(i, n) {
        const T ri = b[i] - q[i];
        r[i] = ri;
        p[i] = ri;
        v += (double)ri * (double)ri;
    }
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// dst = sum of the partials (one block)
__global__ __launch_bounds__(kThreads) void k_finish(const double *__restrict__ part, double *__restrict__ dst) {
    __shared__ double lds[kThreads / 64];
    const double t = sum_parts(part, lds);
    if (threadIdx.x == 0) *dst = t;
}

// alpha = rr / (p.q) from the partials of p.q; x += alpha p; r -= alpha q; part_rr = partials of r.r
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_xr(long long n, const double *__restrict__ part_pq,
                                                    const double *__restrict__ rr, T *__restrict__ x,
                                                    T *__restrict__ r, const T *__restrict__ p,
                                                    const T *__restrict__ q, double *__restrict__ part_rr) {
    __shared__ double lds[kThreads / 64];
    const double pq = sum_parts(part_pq, lds);
    const double alpha = pq > 0.0 ? *rr / pq : 0.0;
    const T a = (T)alpha;
    double v = 0.0;
    VSTRIDE(i, n) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        v += (double)ri * (double)ri;
    }
    v = block_sum(v, lds);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = v;
}

// rr' = sum of part_rr; beta = rr' / rr; p = r + beta p; block 0 records rr' (into the other
// scalar slot, which no block of this kernel reads) and its square root in the history
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_p(long long n, const double *__restrict__ part_rr,
                                                   const double *__restrict__ rr_old, double *__restrict__ rr_new,
                                                   T *__restrict__ p, const T *__restrict__ r,
                                                   double *__restrict__ hist) {
    __shared__ double lds[kThreads / 64];
    const double rn = sum_parts(part_rr, lds);
    const double ro = *rr_old;
    const T beta = (T)(ro > 0.0 ? rn / ro : 0.0);
    VSTRIDE(i, n) p[i] = r[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *rr_new = rn;
        *hist = sqrt(rn);
    }
}

}  // namespace

struct CgState {
    fl_matrix *A = nullptr;
    int kernel = 0, kind = 0;
    hipStream_t s = nullptr;
    DevBuf<double> vec64;  // r, p, q, sp (FP64 mode)
    DevBuf<float> vec32;   // the same in FP32 mode
    DevBuf<double> part1, part2, sc, hist;
    hipGraphExec_t graph = nullptr;
    void *graph_x = nullptr;
    int graph_iters = 0;
    ~CgState() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (s) (void)hipStreamDestroy(s);
    }
};

template <typename T>
static T *vec(CgState *c, int k);

// q = A v: the row kernel reads v in the original column order; the other schedules read the
// permuted copy sp = v[perm].
template <typename T>
static int product(CgState *c, const T *v, T *q) {
    if (c->kernel == FL_KERNEL_ROW) return spmm_row(c->A, c->kind, 1, v, q, 0, c->s);
    T *sp = vec<T>(c, 3);
    permute_x(c->A, v, sp, sizeof(T) == 8, c->s);
    return spmv(c->A, c->kernel, c->kind, sp, q, nullptr, true, c->s);
}

template <typename T>
static T *vec(CgState *c, int k) {
    const size_t n = (size_t)c->A->n;
    if constexpr (sizeof(T) == 8) return c->vec64.p + k * n;
    else return c->vec32.p + k * n;
}

// One iteration: sp = p[perm]; q = A sp; x, r update; p update. `slot` = 0/1 picks the rr scalar
// read (slot) and written (1 - slot); `h` = history entry.
template <typename T>
static int cg_iteration(CgState *c, T *x, int slot, int h) {
    const long long n = c->A->n;
    T *r = vec<T>(c, 0), *p = vec<T>(c, 1), *q = vec<T>(c, 2);
    const int rc = product<T>(c, p, q);
    if (rc) return rc;
    k_dot<T><<<kParts, kThreads, 0, c->s>>>(n, p, q, c->part1.p);
    k_cg_xr<T><<<kParts, kThreads, 0, c->s>>>(n, c->part1.p, c->sc.p + slot, x, r, p, q, c->part2.p);
    k_cg_p<T><<<kParts, kThreads, 0, c->s>>>(n, c->part2.p, c->sc.p + slot, c->sc.p + (1 - slot), p, r,
                                             c->hist.p + h);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

template <typename T>
static int cg_solve(CgState *c, const T *b, T *x, int max_iter, double rtol, int every, int *iters, double *rel) {
    const long long n = c->A->n;
    hipStream_t s = c->s;
    T *r = vec<T>(c, 0), *p = vec<T>(c, 1), *q = vec<T>(c, 2);
    // ||b||, then r = b - A x, p = r, rr = r.r (eager)
    k_dot<T><<<kParts, kThreads, 0, s>>>(n, b, b, c->part1.p);
    k_finish<<<1, kThreads, 0, s>>>(c->part1.p, c->sc.p + 2);
    const int rc = product<T>(c, x, q);
    if (rc) return rc;
    k_cg_init<T><<<kParts, kThreads, 0, s>>>(n, b, q, r, p, c->part1.p);
    k_finish<<<1, kThreads, 0, s>>>(c->part1.p, c->sc.p);
    FL_LAUNCH_CHECK();
    double h0[3];
    FL_HIP(hipMemcpyAsync(h0, c->sc.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    const double bnorm = std::sqrt(h0[2]);
    double res = bnorm > 0.0 ? std::sqrt(h0[0]) / bnorm : 0.0;
    int done = 0;
    if (bnorm == 0.0 || res <= rtol) {
        *iters = 0;
        *rel = res;
        return FL_OK;
    }
    // capture `every` iterations (even, so every replay starts on scalar slot 0)
    if (!c->graph || c->graph_x != (void *)x || c->graph_iters != every) {
        if (c->graph) FL_HIP(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
        FL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        int crc = FL_OK;
        for (int j = 0; j < every && crc == FL_OK; ++j) crc = cg_iteration<T>(c, x, j & 1, j);
        hipGraph_t g = nullptr;
        FL_HIP(hipStreamEndCapture(s, &g));
        if (crc) {
            (void)hipGraphDestroy(g);
            return crc;
        }
        const hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FL_HIP(e);
        c->graph_x = x;
        c->graph_iters = every;
    }
    std::vector<double> hist(every);
    while (done < max_iter) {
        FL_HIP(hipGraphLaunch(c->graph, s));
        FL_HIP(hipMemcpyAsync(hist.data(), c->hist.p, sizeof(double) * every, hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
        int first = -1;
        for (int j = 0; j < every; ++j)
            if (hist[j] / bnorm <= rtol) {
                first = j;
                break;
            }
        done += every;
        res = hist[every - 1] / bnorm;
        if (first >= 0 || !std::isfinite(res)) break;
    }
    *iters = done;
    *rel = res;
    return FL_OK;
}

}  // namespace fl

using namespace fl;

struct fl_cg : CgState {};

extern "C" {

int fl_cg_create(fl_cg **out, fl_matrix *A, int kernel, int val_kind) {
    FL_API_BEGIN
    if (!out || !A) {
        set_error("fl_cg_create: NULL");
        return FL_ERR_INVALID;
    }
    *out = nullptr;
    if (A->m != A->n) {
        set_error("fl_cg_create: the matrix must be square (m = %d, n = %d)", A->m, A->n);
        return FL_ERR_INVALID;
    }
    if (val_kind != FL_VAL_F64 && val_kind != FL_VAL_F32) {
        set_error("fl_cg_create: FL_VAL_F64 or FL_VAL_F32 only");
        return FL_ERR_INVALID;
    }
    fl_cg *c = new fl_cg();
    try {
        c->A = A;
        c->kernel = kernel;
        c->kind = val_kind;
        FL_HIP(hipStreamCreate(&c->s));
        const size_t n = (size_t)std::max(A->n, 1);
        if (val_kind == FL_VAL_F64) c->vec64.alloc(4 * n);
        else c->vec32.alloc(4 * n);
        c->part1.alloc(kParts);
        c->part2.alloc(kParts);
        c->sc.alloc(3);
        // one eager product checks the schedule and makes any first-use allocation now, not
        // during graph capture
        const int rc = val_kind == FL_VAL_F64 ? product<double>(c, vec<double>(c, 0), vec<double>(c, 2))
                                              : product<float>(c, vec<float>(c, 0), vec<float>(c, 2));
        if (rc) {
            delete c;
            return rc;
        }
        FL_HIP(hipStreamSynchronize(c->s));
    } catch (...) {
        delete c;
        throw;
    }
    *out = c;
    return FL_OK;
    FL_API_END
}

int fl_cg_solve(fl_cg *c, const void *d_b, void *d_x, int max_iter, double rtol, int check_every, int *iters,
                double *rel_res) {
    FL_API_BEGIN
    if (!c || !d_b || !d_x || !iters || !rel_res) {
        set_error("fl_cg_solve: NULL");
        return FL_ERR_INVALID;
    }
    if (max_iter < 0 || !(rtol >= 0.0) || check_every < 1 || check_every > 1000) {
        set_error("fl_cg_solve: max_iter >= 0, rtol >= 0 and check_every in [1, 1000] required");
        return FL_ERR_INVALID;
    }
    const int every = (check_every + 1) & ~1;
    if ((int)c->hist.count < every) c->hist.alloc(every);
    if (c->kind == FL_VAL_F64)
        return cg_solve<double>(c, (const double *)d_b, (double *)d_x, max_iter, rtol, every, iters, rel_res);
    return cg_solve<float>(c, (const float *)d_b, (float *)d_x, max_iter, rtol, every, iters, rel_res);
    FL_API_END
}

int fl_cg_destroy(fl_cg *c) {
    FL_API_BEGIN
    delete c;
    return FL_OK;
    FL_API_END
}

}  // extern "C"
