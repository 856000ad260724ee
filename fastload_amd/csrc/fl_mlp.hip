// Learned schedule selector: matrix features and a small MLP trained on the GPU.
// No reference counterpart (the reference has one fixed schedule); its unused column
// statistics (src/auxiliary.h:6-25: mean, standard deviation, ratio) are the closest hook.
//
// Features (40 floats): a 32-bin histogram of log2 column lengths (fractions of columns;
// bin 0 = empty, bin b = lengths in [2^(b-1), 2^b), last bin open) and 8 row/structure
// statistics: log2 m, log2 n, log2 nnz, log2 mean row length, coefficient of variation of row
// lengths, log2 (max/mean row length), fraction of empty rows, mean row span of the columns
// ((last row - first row) / m, which separates local rows from random rows).
//
// MLP: X (B x in) -> ReLU(X W1 + b1) -> ReLU(. W2 + b2) -> . W3 + b3 (logits). GEMMs run on
// MFMA (v_mfma_f32_16x16x16_bf16) with BF16 operands and FP32 accumulation; master weights are
// FP32 (Adam), the forward/backward operands are their BF16 copies. Loss: cross-entropy against
// soft targets (a distribution over schedules derived from measured times).
#include <cmath>
#include <random>
#include <vector>

#include "fl_matrix.h"

namespace fl {

int check_compressed(const int *d_ptr, int count, const int *d_idx, int nnz, int limit, hipStream_t s);

static unsigned sgm(long long count) { return std::max(1u, std::min<unsigned>(grid_for(count), 65536u)); }

#define MSTRIDE(i, count) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += (long long)gridDim.x * blockDim.x)

// --------------------------------------------------------------------------- features
// acc layout (doubles): [0..31] histogram counts, 32 sum row span (normalised), 33 non-empty
// columns; ints: row counts [m]
__global__ void k_feat_cols(const int *__restrict__ colptr, const int *__restrict__ rowidx, int n, int m,
                            double *__restrict__ acc) {
    __shared__ unsigned hist[32];
    __shared__ double span_sum, nonempty;
    if (threadIdx.x < 32) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) { span_sum = 0.0; nonempty = 0.0; }
    __syncthreads();
    double sp = 0.0, ne = 0.0;
    MSTRIDE(c, n) {
        const int a = colptr[c], b = colptr[c + 1], len = b - a;
        int bin = 0;
        if (len > 0) {
            bin = min(31, 32 - __clz(len));  // floor(log2 len) + 1
            sp += (double)(rowidx[b - 1] - rowidx[a]) / (double)max(m, 1);
            ne += 1.0;
        }
        atomicAdd(&hist[bin], 1u);
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_xor(sp, off);
        ne += __shfl_xor(ne, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&span_sum, sp);
        atomicAdd(&nonempty, ne);
    }
    __syncthreads();
    if (threadIdx.x < 32) unsafeAtomicAdd(&acc[threadIdx.x], (double)hist[threadIdx.x]);
    if (threadIdx.x == 0) {
        unsafeAtomicAdd(&acc[32], span_sum);
        unsafeAtomicAdd(&acc[33], nonempty);
    }
}

__global__ void k_feat_rowcount(const int *__restrict__ rowidx, int nnz, int *__restrict__ rc) {
    MSTRIDE(i, nnz) atomicAdd(&rc[rowidx[i]], 1);
}

// acc[34] sum rc, acc[35] sum rc^2, acc[36] empty rows; the largest row count goes to maxrc
__global__ void k_feat_rows(const int *__restrict__ rc, int m, double *__restrict__ acc, int *__restrict__ maxrc) {
    double s = 0.0, s2 = 0.0, e = 0.0;
    int mx = 0;
    MSTRIDE(r, m) {
        const int v = rc[r];
        s += v;
        s2 += (double)v * v;
        e += (v == 0);
        mx = max(mx, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        s2 += __shfl_xor(s2, off);
        e += __shfl_xor(e, off);
        mx = max(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(&acc[34], s);
        unsafeAtomicAdd(&acc[35], s2);
        unsafeAtomicAdd(&acc[36], e);
        atomicMax(maxrc, mx);
    }
}

// ------------------------------------------------------------------------- MFMA GEMM
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// C(m, n) = sum_k A(m, k) B(k, n) with A(m,k) = a[m*sam + k*sak], B(k,n) = b[k*sbk + n*sbn]
// (BF16), C(m,n) = c[m*scm + n*scn] (FP32). One wave per 16x16 tile of C; K in steps of 16.
// Operand layout of v_mfma_f32_16x16x16_bf16 (per lane l): A row l%16, k 4*(l/16)+j; B column
// l%16, k 4*(l/16)+j; D row 4*(l/16)+j, column l%16 (j = 0..3).
__global__ __launch_bounds__(64) void k_gemm_bf16(int M, int N, int K, const uint16_t *__restrict__ a, long long sam,
                                                  long long sak, const uint16_t *__restrict__ b, long long sbk,
                                                  long long sbn, float *__restrict__ c, long long scm, long long scn) {
    const int lane = threadIdx.x;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const int r = lane & 15, q = lane >> 4;
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4_t av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * q + j;
            const int mm = m0 + r, nn = n0 + r;
            av[j] = (short)((mm < M && k < K) ? a[mm * sam + k * sak] : 0);
            bv[j] = (short)((nn < N && k < K) ? b[k * sbk + nn * sbn] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int mm = m0 + 4 * q + j, nn = n0 + r;
        if (mm < M && nn < N) c[mm * scm + nn * scn] = acc[j];
    }
}

__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, long long count) {
    MSTRIDE(i, count) out[i] = to_bf16(in[i]);
}

// Z (B x N) += bias; Hf = relu? ; Hb = bf16(Hf). Z keeps the pre-activation.
__global__ void k_bias_act(float *__restrict__ Z, const float *__restrict__ bias, int B, int N, int relu,
                           float *__restrict__ Hf, uint16_t *__restrict__ Hb) {
    MSTRIDE(i, (long long)B * N) {
        const float z = Z[i] + bias[i % N];
        Z[i] = z;
        const float h = (relu && z < 0.f) ? 0.f : z;
        if (Hf) Hf[i] = h;
        if (Hb) Hb[i] = to_bf16(h);
    }
}

// Softmax cross-entropy against soft targets T (rows sum to 1): dZ = (softmax(Z) - T) / B,
// loss_rows[r] = -sum_c T log softmax(Z).
__global__ void k_softmax_xent(const float *__restrict__ Z, const float *__restrict__ T, int B, int C,
                               float *__restrict__ dZ, uint16_t *__restrict__ dZb, float *__restrict__ loss_rows) {
    MSTRIDE(r, B) {
        const float *z = Z + r * C;
        float mx = -INFINITY;
        for (int c = 0; c < C; ++c) mx = fmaxf(mx, z[c]);
        float s = 0.f;
        for (int c = 0; c < C; ++c) s += expf(z[c] - mx);
        const float ls = logf(s);
        float loss = 0.f;
        for (int c = 0; c < C; ++c) {
            const float lp = z[c] - mx - ls;
            const float t = T[r * C + c];
            loss -= t * lp;
            const float g = (expf(lp) - t) / (float)B;
            dZ[r * C + c] = g;
            dZb[r * C + c] = to_bf16(g);
        }
        loss_rows[r] = loss;
    }
}

// dZ = dH * (Zpre > 0), FP32 and BF16 copies
__global__ void k_relu_back(const float *__restrict__ dH, const float *__restrict__ Zpre, long long count,
                            float *__restrict__ dZ, uint16_t *__restrict__ dZb) {
    MSTRIDE(i, count) {
        const float g = Zpre[i] > 0.f ? dH[i] : 0.f;
        dZ[i] = g;
        dZb[i] = to_bf16(g);
    }
}

__global__ void k_colsum(const float *__restrict__ dZ, int B, int N, float *__restrict__ db) {
    MSTRIDE(c, N) {
        float s = 0.f;
        for (int r = 0; r < B; ++r) s += dZ[r * N + c];
        db[c] = s;
    }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, long long count, float lr, float b1, float b2, float eps, float c1,
                       float c2, uint16_t *__restrict__ wb) {
    MSTRIDE(i, count) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = to_bf16(wi);
    }
}

static void gemm(int M, int N, int K, const uint16_t *a, long long sam, long long sak, const uint16_t *b,
                 long long sbk, long long sbn, float *c, long long scm, long long scn, hipStream_t s) {
    if (M <= 0 || N <= 0) return;
    dim3 grid((M + 15) / 16, (N + 15) / 16);
    k_gemm_bf16<<<grid, 64, 0, s>>>(M, N, K, a, sam, sak, b, sbk, sbn, c, scm, scn);
    FL_LAUNCH_CHECK();
}

}  // namespace fl

using namespace fl;

struct fl_mlp {
    int d[4] = {0, 0, 0, 0};  // in, h1, h2, out
    int t = 0;                // Adam step
    // parameters (FP32 master), BF16 copies of the weights, Adam moments
    DevBuf<float> W[3], bias[3], mW[3], vW[3], mb[3], vb[3], gW[3], gb[3];
    DevBuf<uint16_t> Wb[3];
    // workspace for a batch of `cap` rows
    int cap = 0;
    DevBuf<uint16_t> Xb, Hb[2], dZb[3];
    DevBuf<float> Z[3], dZ[3], dH[2], loss_rows;

    int wcount(int l) const { return d[l] * d[l + 1]; }

    void ensure(int B) {
        if (B <= cap) return;
        cap = B;
        Xb.alloc((size_t)B * d[0]);
        for (int l = 0; l < 2; ++l) {
            Hb[l].alloc((size_t)B * d[l + 1]);
            dH[l].alloc((size_t)B * d[l + 1]);
        }
        for (int l = 0; l < 3; ++l) {
            Z[l].alloc((size_t)B * d[l + 1]);
            dZ[l].alloc((size_t)B * d[l + 1]);
            dZb[l].alloc((size_t)B * d[l + 1]);
        }
        loss_rows.alloc(B);
    }

    // forward: logits in Z[2]
    void forward(const float *X, int B, hipStream_t s) {
        ensure(B);
        k_to_bf16<<<sgm((long long)B * d[0]), kBlock, 0, s>>>(X, Xb.p, (long long)B * d[0]);
        FL_LAUNCH_CHECK();
        const uint16_t *in = Xb.p;
        for (int l = 0; l < 3; ++l) {
            const int K = d[l], N = d[l + 1];
            gemm(B, N, K, in, K, 1, Wb[l].p, N, 1, Z[l].p, N, 1, s);
            const bool hidden = l < 2;
            k_bias_act<<<sgm((long long)B * N), kBlock, 0, s>>>(Z[l].p, bias[l].p, B, N, hidden ? 1 : 0, nullptr,
                                                               hidden ? Hb[l].p : nullptr);
            FL_LAUNCH_CHECK();
            if (hidden) in = Hb[l].p;
        }
    }

    // backward from soft targets T: fills gW/gb, loss_rows
    void backward(const float *T, int B, hipStream_t s) {
        const int C = d[3];
        k_softmax_xent<<<sgm(B), kBlock, 0, s>>>(Z[2].p, T, B, C, dZ[2].p, dZb[2].p, loss_rows.p);
        FL_LAUNCH_CHECK();
        for (int l = 2; l >= 0; --l) {
            const int K = d[l], N = d[l + 1];
            const uint16_t *Hin = (l == 0) ? Xb.p : Hb[l - 1].p;
            // gW[l] (K x N) = Hin^T (K x B) * dZ[l] (B x N)
            gemm(K, N, B, Hin, 1, K, dZb[l].p, N, 1, gW[l].p, N, 1, s);
            k_colsum<<<sgm(N), kBlock, 0, s>>>(dZ[l].p, B, N, gb[l].p);
            FL_LAUNCH_CHECK();
            if (l > 0) {
                // dH (B x K) = dZ[l] (B x N) * W^T (N x K)
                gemm(B, K, N, dZb[l].p, N, 1, Wb[l].p, 1, N, dH[l - 1].p, K, 1, s);
                k_relu_back<<<sgm((long long)B * K), kBlock, 0, s>>>(dH[l - 1].p, Z[l - 1].p, (long long)B * K,
                                                                     dZ[l - 1].p, dZb[l - 1].p);
                FL_LAUNCH_CHECK();
            }
        }
    }

    void refresh_bf16(hipStream_t s) {
        for (int l = 0; l < 3; ++l) {
            k_to_bf16<<<sgm(wcount(l)), kBlock, 0, s>>>(W[l].p, Wb[l].p, wcount(l));
            FL_LAUNCH_CHECK();
        }
    }
};

static int mlp_invalid(const char *m) {
    set_error("%s", m);
    return FL_ERR_INVALID;
}

static float host_loss(fl_mlp *P, int B, hipStream_t s) {
    std::vector<float> lr(B);
    FL_HIP(hipMemcpyAsync(lr.data(), P->loss_rows.p, sizeof(float) * B, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (float v : lr) t += v;
    return (float)(t / B);
}

extern "C" {

int fl_features(int m, int n, int nnz, const int *d_colptr, const int *d_rowidx, float *features, void *stream) {
    FL_API_BEGIN
    if (m < 0 || n < 0 || nnz < 0 || !d_colptr || !features || (nnz && !d_rowidx)) return mlp_invalid("fl_features: bad args");
    hipStream_t s = (hipStream_t)stream;
    // every index the feature kernels use must be inside the matrix
    if (check_compressed(d_colptr, n, d_rowidx, nnz, m, s)) {
        set_error("fl_features: malformed CSC (column pointers not monotone from 0 to nnz, or a row index outside [0, m))");
        return FL_ERR_RANGE;
    }
    DevBuf<double> acc(40);
    DevBuf<int> rc(std::max(m, 1)), maxrc(1);
    FL_HIP(hipMemsetAsync(acc.p, 0, sizeof(double) * 40, s));
    FL_HIP(hipMemsetAsync(rc.p, 0, sizeof(int) * std::max(m, 1), s));
    FL_HIP(hipMemsetAsync(maxrc.p, 0, sizeof(int), s));
    if (n > 0) {
        k_feat_cols<<<std::min(sgm(n), 2048u), kBlock, 0, s>>>(d_colptr, d_rowidx, n, m, acc.p);
        FL_LAUNCH_CHECK();
    }
    if (nnz > 0) {
        k_feat_rowcount<<<sgm(nnz), kBlock, 0, s>>>(d_rowidx, nnz, rc.p);
        FL_LAUNCH_CHECK();
    }
    if (m > 0) {
        k_feat_rows<<<std::min(sgm(m), 2048u), kBlock, 0, s>>>(rc.p, m, acc.p, maxrc.p);
        FL_LAUNCH_CHECK();
    }
    double h[40];
    int hmax = 0;
    FL_HIP(hipMemcpyAsync(h, acc.p, sizeof(double) * 40, hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(&hmax, maxrc.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < 32; ++b) features[b] = n ? (float)(h[b] / n) : 0.f;
    const double mean = m ? h[34] / m : 0.0;
    const double var = m ? std::max(h[35] / m - mean * mean, 0.0) : 0.0;
    features[32] = (float)(std::log2((double)m + 1.0) / 24.0);
    features[33] = (float)(std::log2((double)n + 1.0) / 24.0);
    features[34] = (float)(std::log2((double)nnz + 1.0) / 32.0);
    features[35] = (float)(std::log2(mean + 1.0) / 8.0);
    features[36] = (float)(mean > 0 ? std::sqrt(var) / mean : 0.0);
    features[37] = (float)(mean > 0 ? std::log2((double)hmax / mean + 1.0) / 8.0 : 0.0);
    features[38] = (float)(m ? h[36] / m : 0.0);
    features[39] = (float)(h[33] > 0 ? h[32] / h[33] : 0.0);
    return FL_OK;
    FL_API_END
}

// Schedule space of the selector: tile width x reduction strategy x batch/band parameters.
// Class 0 is the reference schedule (32-wide tiles, segmented sums, atomics per segment).
static const int kSchedules[FL_SCHEDULES][4] = {
    {32, FL_KERNEL_DROPIN, 0, 0},   {64, FL_KERNEL_DROPIN, 0, 0},  {64, FL_KERNEL_WINR, 64, 0},
    {64, FL_KERNEL_WIN, 64, 0},     {64, FL_KERNEL_WIN, 128, 0},   {64, FL_KERNEL_WIN, 256, 0},
    {64, FL_KERNEL_PB, 128, 4096},  {64, FL_KERNEL_PB, 128, 16384}, {64, FL_KERNEL_PB, 256, 4096},
    {64, FL_KERNEL_PB, 256, 8192},  {64, FL_KERNEL_PB, 256, 16384}, {64, FL_KERNEL_ROW, 0, 0},
    {64, FL_KERNEL_HYB, 128, 16384},  // heavy columns min(n, 32768), the rest PB 128 x 16K
};

int fl_schedule_info(int cls, int *width, int *kernel, int *batch_tiles, int *band_rows) {
    if (cls < 0 || cls >= FL_SCHEDULES) return mlp_invalid("fl_schedule_info: class out of range");
    if (width) *width = kSchedules[cls][0];
    if (kernel) *kernel = kSchedules[cls][1];
    if (batch_tiles) *batch_tiles = kSchedules[cls][2];
    if (band_rows) *band_rows = kSchedules[cls][3];
    return FL_OK;
}

int fl_select_schedule(fl_mlp *P, int m, int n, int nnz, const int *d_colptr, const int *d_rowidx, int *cls,
                       float *prob, void *stream) {
    FL_API_BEGIN
    if (!cls) return mlp_invalid("fl_select_schedule: NULL class");
    *cls = 0;
    if (prob) *prob = 0.f;
    if (P && (P->d[0] != 40 || P->d[3] != FL_SCHEDULES))
        return mlp_invalid("fl_select_schedule: network must map 40 features to the FL_SCHEDULES schedules");
    float f[40];
    const int rc = fl_features(m, n, nnz, d_colptr, d_rowidx, f, stream);
    if (rc != FL_OK) return rc;
    if (!P || nnz == 0) return FL_OK;  // fallback: the reference schedule
    for (float v : f)
        if (!std::isfinite(v)) return FL_OK;
    hipStream_t s = (hipStream_t)stream;
    DevBuf<float> df(40);
    FL_HIP(hipMemcpyAsync(df.p, f, sizeof(f), hipMemcpyHostToDevice, s));
    return fl_mlp_predict(P, df.p, 1, cls, prob, stream);
    FL_API_END
}

int fl_mlp_create(fl_mlp **out, int n_in, int n_h1, int n_h2, int n_out, uint64_t seed) {
    FL_API_BEGIN
    if (!out || n_in <= 0 || n_h1 <= 0 || n_h2 <= 0 || n_out <= 0) return mlp_invalid("fl_mlp_create: bad dims");
    fl_mlp *P = new fl_mlp();
    P->d[0] = n_in; P->d[1] = n_h1; P->d[2] = n_h2; P->d[3] = n_out;
    std::mt19937_64 g(seed);
    for (int l = 0; l < 3; ++l) {
        const int K = P->d[l], N = P->d[l + 1];
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / K));  // He initialisation
        std::vector<float> w((size_t)K * N), z((size_t)std::max(K * N, N), 0.f);
        for (auto &v : w) v = nd(g);
        P->W[l].alloc(w.size());
        P->Wb[l].alloc(w.size());
        P->gW[l].alloc(w.size());
        P->mW[l].alloc(w.size());
        P->vW[l].alloc(w.size());
        P->bias[l].alloc(N);
        P->gb[l].alloc(N);
        P->mb[l].alloc(N);
        P->vb[l].alloc(N);
        FL_HIP(hipMemcpy(P->W[l].p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(P->mW[l].p, z.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(P->vW[l].p, z.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(P->bias[l].p, z.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(P->mb[l].p, z.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(P->vb[l].p, z.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    }
    P->refresh_bf16(nullptr);
    FL_HIP(hipDeviceSynchronize());
    *out = P;
    return FL_OK;
    FL_API_END
}

int fl_mlp_destroy(fl_mlp *P) {
    delete P;
    return FL_OK;
}

int fl_mlp_param_count(const fl_mlp *P) {
    if (!P) return 0;
    int c = 0;
    for (int l = 0; l < 3; ++l) c += P->wcount(l) + P->d[l + 1];
    return c;
}

// parameter order: W1 (in x h1, row-major), b1, W2, b2, W3, b3
int fl_mlp_get_params(const fl_mlp *P, float *host) {
    FL_API_BEGIN
    if (!P || !host) return mlp_invalid("fl_mlp_get_params: NULL");
    size_t o = 0;
    for (int l = 0; l < 3; ++l) {
        FL_HIP(hipMemcpy(host + o, P->W[l].p, sizeof(float) * P->wcount(l), hipMemcpyDeviceToHost));
        o += P->wcount(l);
        FL_HIP(hipMemcpy(host + o, P->bias[l].p, sizeof(float) * P->d[l + 1], hipMemcpyDeviceToHost));
        o += P->d[l + 1];
    }
    return FL_OK;
    FL_API_END
}

int fl_mlp_set_params(fl_mlp *P, const float *host) {
    FL_API_BEGIN
    if (!P || !host) return mlp_invalid("fl_mlp_set_params: NULL");
    size_t o = 0;
    for (int l = 0; l < 3; ++l) {
        FL_HIP(hipMemcpy(P->W[l].p, host + o, sizeof(float) * P->wcount(l), hipMemcpyHostToDevice));
        o += P->wcount(l);
        FL_HIP(hipMemcpy(P->bias[l].p, host + o, sizeof(float) * P->d[l + 1], hipMemcpyHostToDevice));
        o += P->d[l + 1];
    }
    P->refresh_bf16(nullptr);
    FL_HIP(hipDeviceSynchronize());
    return FL_OK;
    FL_API_END
}

int fl_mlp_forward(fl_mlp *P, const float *d_X, int batch, float *d_logits, void *stream) {
    FL_API_BEGIN
    if (!P || !d_X || !d_logits || batch <= 0) return mlp_invalid("fl_mlp_forward: bad args");
    hipStream_t s = (hipStream_t)stream;
    P->forward(d_X, batch, s);
    FL_HIP(hipMemcpyAsync(d_logits, P->Z[2].p, sizeof(float) * batch * P->d[3], hipMemcpyDeviceToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
    FL_API_END
}

// Forward pass and per-row argmax; prob[r] = softmax probability of the chosen class.
int fl_mlp_predict(fl_mlp *P, const float *d_X, int batch, int *classes, float *prob, void *stream) {
    FL_API_BEGIN
    if (!P || !d_X || !classes || batch <= 0) return mlp_invalid("fl_mlp_predict: bad args");
    hipStream_t s = (hipStream_t)stream;
    P->forward(d_X, batch, s);
    const int C = P->d[3];
    std::vector<float> z((size_t)batch * C);
    FL_HIP(hipMemcpyAsync(z.data(), P->Z[2].p, sizeof(float) * z.size(), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < batch; ++r) {
        const float *zr = z.data() + (size_t)r * C;
        int best = 0;
        for (int c = 1; c < C; ++c)
            if (zr[c] > zr[best]) best = c;
        double sum = 0.0;
        for (int c = 0; c < C; ++c) sum += std::exp((double)zr[c] - zr[best]);
        classes[r] = best;
        if (prob) prob[r] = (float)(1.0 / sum);
    }
    return FL_OK;
    FL_API_END
}

// Gradients of the mean soft-target cross-entropy at the current parameters (no update), in the
// parameter order of fl_mlp_get_params.
int fl_mlp_gradients(fl_mlp *P, const float *d_X, const float *d_T, int batch, float *host_grads, float *loss,
                     void *stream) {
    FL_API_BEGIN
    if (!P || !d_X || !d_T || batch <= 0 || !host_grads) return mlp_invalid("fl_mlp_gradients: bad args");
    hipStream_t s = (hipStream_t)stream;
    P->forward(d_X, batch, s);
    P->backward(d_T, batch, s);
    size_t o = 0;
    for (int l = 0; l < 3; ++l) {
        FL_HIP(hipMemcpyAsync(host_grads + o, P->gW[l].p, sizeof(float) * P->wcount(l), hipMemcpyDeviceToHost, s));
        o += P->wcount(l);
        FL_HIP(hipMemcpyAsync(host_grads + o, P->gb[l].p, sizeof(float) * P->d[l + 1], hipMemcpyDeviceToHost, s));
        o += P->d[l + 1];
    }
    const float L = host_loss(P, batch, s);
    if (loss) *loss = L;
    return FL_OK;
    FL_API_END
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean soft-target cross-entropy.
int fl_mlp_train_step(fl_mlp *P, const float *d_X, const float *d_T, int batch, float lr, float *loss, void *stream) {
    FL_API_BEGIN
    if (!P || !d_X || !d_T || batch <= 0) return mlp_invalid("fl_mlp_train_step: bad args");
    hipStream_t s = (hipStream_t)stream;
    P->forward(d_X, batch, s);
    P->backward(d_T, batch, s);
    P->t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)P->t), c2 = 1.f - std::pow(b2, (float)P->t);
    for (int l = 0; l < 3; ++l) {
        k_adam<<<sgm(P->wcount(l)), kBlock, 0, s>>>(P->W[l].p, P->gW[l].p, P->mW[l].p, P->vW[l].p, P->wcount(l), lr,
                                                    b1, b2, 1e-8f, c1, c2, P->Wb[l].p);
        k_adam<<<sgm(P->d[l + 1]), kBlock, 0, s>>>(P->bias[l].p, P->gb[l].p, P->mb[l].p, P->vb[l].p, P->d[l + 1], lr,
                                                   b1, b2, 1e-8f, c1, c2, nullptr);
        FL_LAUNCH_CHECK();
    }
    if (loss) *loss = host_loss(P, batch, s);
    return FL_OK;
    FL_API_END
}

}  // extern "C"
