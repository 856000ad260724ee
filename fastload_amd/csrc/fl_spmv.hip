// FastLoad SpMV kernels for gfx950 (wave64).
//
// DROPIN: the reference per-tile algorithm (FastLoad___kernel, src/FastLoad_GPU.h:146-273)
// over the reference layout. A wave holds 64/W tiles (two 32-wide tiles in the reference
// layout). Per tile it reads tile_len/tile_ptr/tile_colidx/tile_format and, for format 2,
// sortrowindex. Format 0 (partial tile): one atomic per lane; format 1 (one row): one atomic
// with the tile sum; format 2: one atomic per segment with the segment sum. The segment sums
// come from a true segmented scan (DPP row_shr / row_bcast) -- the reference subtracts prefix
// sums (FastLoad_GPU.h:263), which cancels in low precision.
//
// FAST: same tiles (W=64), but tile geometry comes from the per-level table (a few scalar
// loads per level instead of 16-20 bytes per tile), segment heads come from comparing each
// lane's row with its left neighbour (DPP wave_shr:1) instead of reading sortrowindex, and a
// wave processes U tiles per step with all loads issued before any reduction.
#include <type_traits>

#include "fl_matrix.h"

namespace fl {

__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <typename V>
struct ValLoad;
template <>
struct ValLoad<double> {
    static __device__ __forceinline__ double get(const double *p, int i) { return p[i]; }
};
template <>
struct ValLoad<float> {
    static __device__ __forceinline__ float get(const float *p, int i) { return p[i]; }
};
template <>
struct ValLoad<uint16_t> {
    static __device__ __forceinline__ float get(const uint16_t *p, int i) { return bf16_to_f32(p[i]); }
};

// ------------------------------------------------------------------------------ DROPIN
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void k_spmv_dropin(int T_, const int *__restrict__ tile_ptr,
                                                        const int *__restrict__ tile_len,
                                                        const int *__restrict__ tile_colidx,
                                                        const int *__restrict__ tile_format,
                                                        const int *__restrict__ sortrowindex,
                                                        const int *__restrict__ sortrowidx,
                                                        const T *__restrict__ sortval, const T *__restrict__ x,
                                                        T *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int sub = lane / W, l = lane % W;
    const long long t = wave * (64 / W) + sub;
    const bool tv = t < T_;
    int len = 0, ptr = 0, col = 0, fmt = 0;
    if (tv) {
        len = tile_len[t];
        ptr = tile_ptr[t];
        col = tile_colidx[t];
        fmt = tile_format[t];
    }
    const bool act = l < len;
    const int g = ptr + l;
    int row = 0;
    T p = T(0);
    bool head = true;  // inactive lanes are singleton segments with p = 0
    if (act) {
        row = sortrowidx[g];
        p = sortval[g] * x[col + l];
        head = (fmt == 0) ? true : (fmt == 1) ? (l == 0) : (sortrowindex[g] != 0);
    }
    const uint64_t hm = ballot64(head);
    const T v = wave_segscan<T>(p, segment_start(hm, lane), lane);
    const bool tail = act && (lane == 63 || ((hm >> (lane + 1)) & 1ull) || l == len - 1);
    if (tail) atomic_add(&y[row], v);
}

// -------------------------------------------------------------------------------- FAST
template <typename V, typename X, typename Y, int U, bool SCAN, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_spmv_fast64(int T_, int groups, LevelTab lt,
                                                        const int *__restrict__ sortrowidx,
                                                        const V *__restrict__ val, const V *__restrict__ val_lo,
                                                        const X *__restrict__ x, Y *__restrict__ y,
                                                        Y *__restrict__ y_lo) {
    const int lane = threadIdx.x & 63;
    const int wave = uni((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    for (int gi = wave; gi < groups; gi += nwaves) {
        const int t0 = gi * 64;
        const int t1 = min(t0 + 64, T_);
        LevelWalk s;
        s.seek(lt, t0);
        for (int t = t0; t < t1; t += U) {
            int ptr[U], cx[U], len[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (t + u < t1) {
                    s.next(lt, ptr[u], cx[u], len[u], t + u + 1 < t1);
                } else {
                    len[u] = 0;
                    ptr[u] = 0;
                    cx[u] = 0;
                }
            }
            int row[U];
            float pv[U], pl[U];
            double pd[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool act = lane < len[u];
                const int g = ptr[u] + lane;
                row[u] = act ? sortrowidx[g] : -1;
                if constexpr (sizeof(Y) == 8 && !SPLIT) {
                    pd[u] = act ? (double)ValLoad<V>::get(val, g) * (double)x[cx[u] + lane] : 0.0;
                } else {
                    const float xv = act ? (float)x[cx[u] + lane] : 0.f;
                    pv[u] = act ? (float)ValLoad<V>::get(val, g) * xv : 0.f;
                    if constexpr (SPLIT) pl[u] = act ? ValLoad<V>::get(val_lo, g) * xv : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool act = lane < len[u];
                if constexpr (SCAN) {
                    const int left = __builtin_amdgcn_update_dpp(-1, row[u], kWaveShr1, 0xf, 0xf, false);
                    const bool head = !act || lane == 0 || row[u] != left;
                    const uint64_t hm = ballot64(head);
                    const int st = segment_start(hm, lane);
                    const bool tail = act && (lane == 63 || ((hm >> (lane + 1)) & 1ull));
                    if constexpr (sizeof(Y) == 8 && !SPLIT) {
                        const double v = wave_segscan<double>(pd[u], st, lane);
                        if (tail) atomic_add((double *)&y[row[u]], v);
                    } else {
                        const float v = wave_segscan<float>(pv[u], st, lane);
                        if constexpr (SPLIT) {
                            const float w = wave_segscan<float>(pl[u], st, lane);
                            if (tail) {
                                atomic_add((float *)&y[row[u]], v);
                                atomic_add((float *)&y_lo[row[u]], w);
                            }
                        } else {
                            if (tail) atomic_add((float *)&y[row[u]], v);
                        }
                    }
                } else {
                    if (act) {
                        if constexpr (sizeof(Y) == 8 && !SPLIT) {
                            atomic_add((double *)&y[row[u]], pd[u]);
                        } else {
                            atomic_add((float *)&y[row[u]], pv[u]);
                            if constexpr (SPLIT) atomic_add((float *)&y_lo[row[u]], pl[u]);
                        }
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------------- conversions
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ void k_to_f32(const double *__restrict__ in, float *__restrict__ out, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

__global__ void k_to_bf16x2(const double *__restrict__ in, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo,
                            int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const float v = (float)in[i];
        const uint16_t h = f32_to_bf16_rne(v);
        hi[i] = h;
        lo[i] = f32_to_bf16_rne(v - bf16_to_f32(h));
    }
}

__global__ void k_pack_pair(const uint16_t *__restrict__ hi, const uint16_t *__restrict__ lo,
                            uint32_t *__restrict__ pair, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        pair[i] = ((uint32_t)hi[i] << 16) | lo[i];
}

__global__ void k_recombine(const float *__restrict__ yh, const float *__restrict__ yl, double *__restrict__ y,
                            int m) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
        y[i] += (double)yh[i] + (double)yl[i];
}

template <typename T>
__global__ void k_permute(const T *__restrict__ x, const int *__restrict__ perm, T *__restrict__ sx, int n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        sx[i] = x[perm[i]];
}

static unsigned sgrid(long long count) { return std::max(1u, std::min<unsigned>(grid_for(count), 65536u)); }

void pack_schedules(fl_matrix *A, hipStream_t s);

void prepare_values(fl_matrix *A, int kind, hipStream_t s) {
    const int nnz = A->nnz;
    if (kind == FL_VAL_PACKED) {  // packed words of the prepared schedules (and of later ones)
        A->pack_wanted = true;
        pack_schedules(A, s);
        if (A->hyb_light) hyb_values(A, kind, s);
        return;
    }
    if ((kind == FL_VAL_F32) && !A->val32.p) {
        A->val32.alloc(std::max(nnz, 1));
        if (nnz) k_to_f32<<<sgrid(nnz), kBlock, 0, s>>>(A->val64.p, A->val32.p, nnz);
        FL_LAUNCH_CHECK();
    }
    if ((kind == FL_VAL_BF16 || kind == FL_VAL_BF16X2) && !A->bf16_hi.p) {
        A->bf16_hi.alloc(std::max(nnz, 1));
        A->bf16_lo.alloc(std::max(nnz, 1));
        if (nnz) k_to_bf16x2<<<sgrid(nnz), kBlock, 0, s>>>(A->val64.p, A->bf16_hi.p, A->bf16_lo.p, nnz);
        FL_LAUNCH_CHECK();
    }
    if (kind == FL_VAL_BF16X2 && !A->bf16_pair.p) {
        A->bf16_pair.alloc(std::max(nnz, 1));
        if (nnz) k_pack_pair<<<sgrid(nnz), kBlock, 0, s>>>(A->bf16_hi.p, A->bf16_lo.p, A->bf16_pair.p, nnz);
        FL_LAUNCH_CHECK();
    }
    if (A->hyb_light) hyb_values(A, kind, s);
    FL_HIP(hipStreamSynchronize(s));
}

void permute_raw(const int *perm, int n, const void *x, void *sx, bool dbl, hipStream_t s) {
    if (n <= 0) return;
    if (dbl)
        k_permute<double><<<sgrid(n), kBlock, 0, s>>>((const double *)x, perm, (double *)sx, n);
    else
        k_permute<float><<<sgrid(n), kBlock, 0, s>>>((const float *)x, perm, (float *)sx, n);
    FL_LAUNCH_CHECK();
}

void permute_x(const fl_matrix *A, const void *x, void *sx, bool dbl, hipStream_t s) {
    permute_raw(A->perm.p, A->n, x, sx, dbl, s);
}

// Launch helpers -------------------------------------------------------------------------
template <typename T, int W>
static void launch_dropin(const fl_matrix *A, const T *x, T *y, hipStream_t s) {
    const long long waves = ((long long)A->tilenum + (64 / W) - 1) / (64 / W);
    k_spmv_dropin<T, W><<<grid_for(waves * 64), kBlock, 0, s>>>(
        A->tilenum, A->tile_ptr.p, A->tile_len.p, A->tile_colidx.p, A->tile_format.p, A->sortrowindex.p,
        A->sortrowidx.p, std::is_same<T, double>::value ? (const T *)A->val64.p : (const T *)A->val32.p, x, y);
    FL_LAUNCH_CHECK();
}

template <typename V, typename X, typename Y, bool SPLIT>
static void launch_fast(const fl_matrix *A, const V *val, const V *val_lo, const X *x, Y *y, Y *y_lo, hipStream_t s) {
    const int groups = (A->tilenum + 63) / 64;
    constexpr int U = 4;
    const unsigned g = grid_for((long long)groups * 64);
    k_spmv_fast64<V, X, Y, U, true, SPLIT><<<g, kBlock, 0, s>>>(A->tilenum, groups, level_tab(A), A->sortrowidx.p, val,
                                                                val_lo, x, y, y_lo);
    FL_LAUNCH_CHECK();
}

void zero_fill(void *d, size_t bytes, hipStream_t s);

static int refuse(const char *msg) {
    set_error("%s", msg);
    return FL_ERR_INVALID;
}

// Every condition under which a call would be refused, checked before anything is enqueued, so
// that an invalid call leaves y as it was.
static int spmv_check(const fl_matrix *A, int kernel, int kind, const void *work) {
    if (kind < FL_VAL_F64 || kind > FL_VAL_PACKED) {
        set_error("fl_spmv: unknown value kind %d", kind);
        return FL_ERR_INVALID;
    }
    if (kernel < FL_KERNEL_DROPIN || kernel > FL_KERNEL_HYB) {
        set_error("fl_spmv: unknown kernel %d", kernel);
        return FL_ERR_INVALID;
    }
    if (A->tilenum == 0) return FL_OK;  // nothing to multiply: y = 0 (assign) or unchanged
    if (kernel == FL_KERNEL_HYB) return hyb_check(A, kind);
    const bool f32_missing = kind == FL_VAL_F32 && !A->val32.p;
    if (kind == FL_VAL_PACKED && (kernel == FL_KERNEL_DROPIN || kernel == FL_KERNEL_FAST))
        return refuse("fl_spmv: FL_VAL_PACKED is a storage of the batched schedules (WIN, WINR, PB)");
    if (kernel == FL_KERNEL_DROPIN) {
        if (kind != FL_VAL_F64 && kind != FL_VAL_F32)
            return refuse("fl_spmv: the drop-in kernel supports FL_VAL_F64 and FL_VAL_F32");
        if (f32_missing) return refuse("fl_spmv: FP32 values not prepared");
        return FL_OK;
    }
    if (A->W != 64) {
        return refuse(kernel == FL_KERNEL_FAST ? "fl_spmv: the FAST kernel needs a 64-wide tile layout"
                                               : "fl_spmv: batched schedules need a 64-wide tile layout");
    }
    if (f32_missing) return refuse("fl_spmv: FP32 values not prepared");
    if (kernel == FL_KERNEL_FAST) {
        if ((kind == FL_VAL_BF16 || kind == FL_VAL_BF16X2) && !A->bf16_hi.p)
            return refuse("fl_spmv: BF16 values not prepared");
        if (kind == FL_VAL_BF16X2 && !work) return refuse("fl_spmv: FL_VAL_BF16X2 needs a work buffer of 2*m floats");
        return FL_OK;
    }
    if (kernel == FL_KERNEL_PB) {
        if (!A->pb_ready) return refuse("fl_spmv: PB schedule not prepared");
    } else if (!A->win_ready || A->ord_kind != (kernel == FL_KERNEL_WINR ? 1 : 0)) {
        return refuse(kernel == FL_KERNEL_WINR ? "fl_spmv: WINR schedule not prepared"
                                               : "fl_spmv: WIN schedule not prepared");
    }
    if ((kind == FL_VAL_BF16 && !A->bf16_hi.p) || (kind == FL_VAL_BF16X2 && !A->bf16_pair.p))
        return refuse("fl_spmv: BF16 values not prepared");
    if (kind == FL_VAL_PACKED && !(kernel == FL_KERNEL_PB ? A->pb_pack.p : A->win_pack.p))
        return refuse("fl_spmv: no packed values for this schedule (fl_matrix_prepare_values(FL_VAL_PACKED) not "
                      "called, or the slot bits and the matrix's exponent range leave fewer than 12 mantissa bits)");
    return FL_OK;
}

// assign: y = A x instead of y += A x. The PB schedule writes every row once and needs no
// clearing; the other kernels accumulate, so y is cleared with a fill kernel first (WIN and
// WINR then store the rows a single batch owns and add the shared ones atomically).
int spmv(fl_matrix *A, int kernel, int kind, const void *x, void *y, void *work, bool assign, hipStream_t s,
         void *pb_bins) {
    const int rc = spmv_check(A, kernel, kind, work);
    if (rc) return rc;
    const size_t ybytes = (size_t)A->m * ((kind == FL_VAL_F64 || kind == FL_VAL_BF16X2) ? 8 : 4);
    // PB and HYB write every row once when assigning; the other kernels accumulate
    if (assign && (A->tilenum == 0 || (kernel != FL_KERNEL_PB && kernel != FL_KERNEL_HYB))) zero_fill(y, ybytes, s);
    if (A->tilenum == 0) return FL_OK;
    if (kernel == FL_KERNEL_HYB) return spmv_hyb(A, kind, x, y, assign, s, pb_bins);
    if (kernel == FL_KERNEL_WIN || kernel == FL_KERNEL_WINR || kernel == FL_KERNEL_PB)
        return spmv_batched(A, kernel, kind, x, y, assign, s, pb_bins);
    if (kernel == FL_KERNEL_DROPIN) {
        if (kind == FL_VAL_F64) {
            if (A->W == 32) launch_dropin<double, 32>(A, (const double *)x, (double *)y, s);
            else launch_dropin<double, 64>(A, (const double *)x, (double *)y, s);
        } else {
            if (A->W == 32) launch_dropin<float, 32>(A, (const float *)x, (float *)y, s);
            else launch_dropin<float, 64>(A, (const float *)x, (float *)y, s);
        }
        return FL_OK;
    }
    switch (kind) {  // FAST
        case FL_VAL_F64:
            launch_fast<double, double, double, false>(A, A->val64.p, nullptr, (const double *)x, (double *)y, nullptr, s);
            return FL_OK;
        case FL_VAL_F32:
            launch_fast<float, float, float, false>(A, A->val32.p, nullptr, (const float *)x, (float *)y, nullptr, s);
            return FL_OK;
        case FL_VAL_BF16:
            launch_fast<uint16_t, float, float, false>(A, A->bf16_hi.p, nullptr, (const float *)x, (float *)y, nullptr, s);
            return FL_OK;
        default: {  // FL_VAL_BF16X2
            float *yh = (float *)work, *yl = yh + A->m;
            zero_fill(yh, sizeof(float) * 2 * (size_t)A->m, s);
            launch_fast<uint16_t, float, float, true>(A, A->bf16_hi.p, A->bf16_lo.p, (const float *)x, yh, yl, s);
            k_recombine<<<sgrid(A->m), kBlock, 0, s>>>(yh, yl, (double *)y, A->m);
            FL_LAUNCH_CHECK();
            return FL_OK;
        }
    }
}

}  // namespace fl
