// HYB schedule: heavy columns from LDS, the rest through propagation blocking.
//
// A power-law matrix keeps most of its entries in few columns: on the 4M-column, 67M-entry
// workload the 32768 longest columns hold 70 % of them. Their x values (128 KB in FP32) fit in one
// workgroup's LDS, so those entries can be summed row by row with x gathered from LDS: each entry
// costs its value and a 16-bit column, and y is written once per row, with no product buffer and
// no row index. The other columns keep the PB schedule (a product and a 16-bit row per entry
// moved twice through the Infinity Cache), run on a sub-matrix of their own.
//
// Layout of the heavy part (all built on the GPU by hyb_prepare):
//   - rows are ordered by heavy entry count, longest first, inside windows of kHybWin rows
//     (hyb_iperm[r] = position of row r inside its window), so the 64 rows of a slice have
//     nearly the same count;
//   - slice s = sorted positions [64 s, 64 s + 64) is an ELL block of hyb_np[s] pair rows: slot
//     (q, lane), at hyb_sb[s] + 64 q + lane, holds entries 2q and 2q + 1 of the lane's row as a
//     value pair and one word of two 16-bit columns; rows shorter than the slice's longest are
//     padded with value 0 and column H, which reads the zero stored after x in LDS. On the
//     power-law workload the padding adds 5 % to the heavy entries. A wave reads one slot per
//     lane per pair row: 8 (16 in FP64) + 4 contiguous bytes, with no per-entry index math;
//   - the heavy kernel writes y by sorted position (hyb_ys), and the sub-matrix's first PB phase 2
//     adds it to y: with bands of kHybWin rows (one band = one window) it reads hyb_ys in order
//     and adds hyb_ys[p] to row hyb_perm[p] of its LDS sums (a gather through hyb_iperm would
//     move a 128-byte line per 4-byte value: 14 of phase 2's 42 us on the power-law workload);
//     with other band sizes it adds hyb_ys[(r & ~(kHybWin - 1)) + hyb_iperm[r]] to row r.
// The column sort is ascending, so the heavy columns are the last H sorted columns: heavy x is
// sortx + (n - H), and the sub-matrix's sorted columns are the first n - H sorted columns in the
// same order (its own column sort of already ascending counts is the identity, so it is not run),
// so the sub-matrix reads sortx directly.
#include <algorithm>

#include "fl_matrix.h"

namespace fl {

void pb_prepare(fl_matrix *A, int bt, int R, hipStream_t s);
void prepare_values(fl_matrix *A, int kind, hipStream_t s);
void zero_fill(void *d, size_t bytes, hipStream_t s);

#define HSTRIDE(i, count) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += (long long)gridDim.x * blockDim.x)

// pair rows per load group of the heavy kernel (two groups in flight per wave)
constexpr int kHybPairs = 2;

static unsigned hg(long long threads) { return std::max(1u, std::min<unsigned>(grid_for(threads), 65536u)); }

static int hread(const int *d, hipStream_t s) {
    int h = 0;
    FL_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return h;
}

// ------------------------------------------------------------------------ preparation
// (sorted column, entry) pairs of every entry in tile order
__global__ void k_hyb_scol(int T, int W, const int *__restrict__ tile_ptr, const int *__restrict__ tile_len,
                           const int *__restrict__ tile_colidx, uint32_t *__restrict__ key, int *__restrict__ order) {
    HSTRIDE(g, (long long)T * W) {
        const long long t = g / W;
        const int l = (int)(g - t * W);
        if (l < tile_len[t]) {
            const int e = tile_ptr[t] + l;
            key[e] = (uint32_t)(tile_colidx[t] + l);
            order[e] = e;
        }
    }
}

__global__ void k_hyb_iota(int *__restrict__ v, int count) {
    HSTRIDE(i, count) v[i] = (int)i;
}

// the light entries (first `count` of the column-sorted order) as a column-sorted CSC
__global__ void k_hyb_light(const int *__restrict__ order, const int *__restrict__ row, const double *__restrict__ val,
                            int count, int *__restrict__ trow, double *__restrict__ tval) {
    HSTRIDE(i, count) {
        const int e = order[i];
        trow[i] = row[e];
        tval[i] = val[e];
    }
}

__global__ void k_hyb_rowkey(const int *__restrict__ order_h, const int *__restrict__ row, int count,
                             uint32_t *__restrict__ key, int *__restrict__ idx) {
    HSTRIDE(i, count) {
        key[i] = (uint32_t)row[order_h[i]];
        idx[i] = (int)i;
    }
}

// ptr[r] = first i with key[i] >= r, r in [0, nb]; key ascending
__global__ void k_hyb_bounds(const uint32_t *__restrict__ key, int count, int nb, int *__restrict__ ptr) {
    HSTRIDE(i, (long long)count + 1) {
        const long long lo = i == 0 ? 0 : (long long)key[i - 1] + 1;
        const long long hi = i == count ? nb : (long long)key[i];
        for (long long r = lo; r <= hi; ++r) ptr[r] = (int)i;
    }
}

__global__ void k_hyb_lenkey(const int *__restrict__ hptr, int m, uint32_t *__restrict__ key, int *__restrict__ rows) {
    HSTRIDE(r, m) {
        key[r] = 65535u - (uint32_t)(hptr[r + 1] - hptr[r]);
        rows[r] = (int)r;
    }
}

__global__ void k_hyb_winkey(const int *__restrict__ rows, int m, uint32_t *__restrict__ key) {
    HSTRIDE(p, m) key[p] = (uint32_t)(rows[p] >> kHybWinLog);
}

__global__ void k_hyb_pos(const int *__restrict__ rows, int m, uint16_t *__restrict__ iperm,
                          uint16_t *__restrict__ perm, int *__restrict__ flag) {
    HSTRIDE(p, m) {
        const int r = rows[p];
        iperm[r] = (uint16_t)(p & (kHybWin - 1));
        perm[p] = (uint16_t)(r & (kHybWin - 1));
        if ((p >> kHybWinLog) != (r >> kHybWinLog)) *flag = 1;  // a row left its window
    }
}

// pair rows of every slice: half the longest row, rounded up (the first lane's: longest first)
__global__ void k_hyb_pairs(const int *__restrict__ rows, const int *__restrict__ hptr, int m, int nslice,
                            int *__restrict__ np, int *__restrict__ slots) {
    HSTRIDE(s, nslice) {
        const int r = rows[s * 64];
        const int p = (hptr[r + 1] - hptr[r] + 1) / 2;
        np[s] = p;
        slots[s] = p;
    }
}

// One lane per sorted position writes its row's slots: the heavy entries in column order, then
// the padding (value 0, column H) up to the slice's pair rows.
__global__ void k_hyb_fill(int m, int H, const int *__restrict__ rows, const int *__restrict__ np,
                           const long long *__restrict__ sb, const int *__restrict__ hptr, const int *__restrict__ ridx,
                           const int *__restrict__ order_h, const uint32_t *__restrict__ scol_h, int first_heavy,
                           const double *__restrict__ val, uint32_t *__restrict__ col2, double *__restrict__ val2) {
    HSTRIDE(p, m) {
        const long long s = p >> 6;
        const int lane = (int)(p & 63), r = rows[p], base = hptr[r], L = hptr[r + 1] - base;
        const int P = np[s];
        for (int q = 0; q < P; ++q) {
            const long long slot = sb[s] + 64LL * q + lane;
            uint32_t c[2];
            double v[2];
            for (int h = 0; h < 2; ++h) {
                const int j = 2 * q + h;
                if (j < L) {
                    const int i = ridx[base + j];
                    c[h] = scol_h[i] - (uint32_t)first_heavy;
                    v[h] = val[order_h[i]];
                } else {
                    c[h] = (uint32_t)H;
                    v[h] = 0.0;
                }
            }
            col2[slot] = c[0] | (c[1] << 16);
            val2[2 * slot] = v[0];
            val2[2 * slot + 1] = v[1];
        }
    }
}

__global__ void k_hyb_fill_u32(uint32_t *__restrict__ d, uint32_t v, long long count) {
    HSTRIDE(i, count) d[i] = v;
}

__global__ void k_hyb_slot_base(const int *__restrict__ pscan, int count, long long *__restrict__ sb) {
    HSTRIDE(i, (long long)count + 1) sb[i] = 64LL * pscan[i];
}

__global__ void k_hyb_f32(const double *__restrict__ in, float *__restrict__ out, long long count) {
    HSTRIDE(i, count) out[i] = (float)in[i];
}

static int cu_count() {
    int dev = 0, cus = 0;
    FL_HIP(hipGetDevice(&dev));
    FL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(cus, 1);
}

void hyb_values(fl_matrix *A, int kind, hipStream_t s) {
    fl_matrix *L = A->hyb_light.get();
    if (!L) return;
    if ((kind == FL_VAL_F32 || kind == FL_VAL_PACKED) && !A->hyb_val32.p) {
        const long long c = 2 * (A->hyb_slots + 2LL * kHybPairs * 64);
        A->hyb_val32.alloc((size_t)c);
        k_hyb_f32<<<hg(c), kBlock, 0, s>>>(A->hyb_val64.p, A->hyb_val32.p, c);
        FL_LAUNCH_CHECK();
    }
    if (kind == FL_VAL_F32 || kind == FL_VAL_PACKED) prepare_values(L, kind, s);
    FL_HIP(hipStreamSynchronize(s));
}

void hyb_prepare(fl_matrix *A, int H, int bt, int R, hipStream_t s) {
    if (A->hyb_ready && A->hyb_H == H && A->hyb_bt == bt && A->hyb_R == R) return;
    A->hyb_ready = false;
    A->hyb_light.reset();
    A->hyb_val32.release();
    A->hyb_ys32.release();
    A->hyb_ys64.release();
    const int m = A->m, n = A->n, nnz = A->nnz, T = A->tilenum, nl = n - H;
    // entries by sorted column (stable: inside a column in tile order, i.e. by depth)
    TmpBuf<uint32_t> key(std::max(nnz, 1), s);
    TmpBuf<int> order(std::max(nnz, 1), s);
    if (nnz > 0) {
        k_hyb_scol<<<hg((long long)T * A->W), kBlock, 0, s>>>(T, A->W, A->tile_ptr.p, A->tile_len.p, A->tile_colidx.p,
                                                             key.p, order.p);
        FL_LAUNCH_CHECK();
        radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(n - 1, 0)), s);
    }
    TmpBuf<int> cps((size_t)n + 1, s);
    exclusive_scan_i32(A->sortnnz.p, cps.p, n, s);
    const int nnz_l = nl > 0 ? hread(cps.p + nl, s) : 0;
    const int nnz_h = nnz - nnz_l;

    // light sub-matrix: sorted columns [0, nl) in the same order, its own tiles and PB schedule
    auto L = std::make_unique<fl_matrix>();
    L->m = m;
    L->n = nl;
    L->nnz = nnz_l;
    L->W = 64;
    L->perm.alloc(std::max(nl, 1));
    L->sortnnz.alloc(std::max(nl, 1));
    if (nl > 0) {
        k_hyb_iota<<<hg(nl), kBlock, 0, s>>>(L->perm.p, nl);
        FL_LAUNCH_CHECK();
        FL_HIP(hipMemcpyAsync(L->sortnnz.p, A->sortnnz.p, sizeof(int) * (size_t)nl, hipMemcpyDeviceToDevice, s));
        L->max_col_nnz = hread(A->sortnnz.p + nl - 1, s);
    }
    {
        TmpBuf<int> trow(std::max(nnz_l, 1), s);
        TmpBuf<double> tval(std::max(nnz_l, 1), s);
        if (nnz_l > 0) {
            k_hyb_light<<<hg(nnz_l), kBlock, 0, s>>>(order.p, A->sortrowidx.p, A->val64.p, nnz_l, trow.p, tval.p);
            FL_LAUNCH_CHECK();
        }
        ElemSource src;
        src.P = cps.p;
        src.rowidx = trow.p;
        src.val = tval.p;
        transform_device(L.get(), L->sortnnz.p, src, s, nullptr, nullptr);
        FL_HIP(hipStreamSynchronize(s));
    }
    pb_prepare(L.get(), bt, R, s);

    // heavy part
    const int nslice = (m + 63) / 64;
    A->hyb_nslice = nslice;
    A->hyb_nnz = nnz_h;
    A->hyb_sb.alloc((size_t)nslice + 1);
    A->hyb_np.alloc(std::max(nslice, 1));
    A->hyb_iperm.alloc(std::max(m, 1));
    A->hyb_perm.alloc(std::max(m, 1));
    long long slots = 0;
    TmpBuf<int> hptr((size_t)m + 1, s), rows(std::max(m, 1), s), flag(1, s);
    TmpBuf<uint32_t> rkey(std::max(nnz_h, 1), s);
    TmpBuf<int> ridx(std::max(nnz_h, 1), s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    FL_HIP(hipMemsetAsync(A->hyb_sb.p, 0, sizeof(long long) * ((size_t)nslice + 1), s));
    if (m > 0) {
        if (nnz_h > 0) {  // heavy entries by row (stable: inside a row by column)
            k_hyb_rowkey<<<hg(nnz_h), kBlock, 0, s>>>(order.p + nnz_l, A->sortrowidx.p, nnz_h, rkey.p, ridx.p);
            FL_LAUNCH_CHECK();
            radix_sort_pairs(rkey.p, ridx.p, nnz_h, bits_for((uint32_t)(m - 1)), s);
        }
        k_hyb_bounds<<<hg((long long)nnz_h + 1), kBlock, 0, s>>>(rkey.p, nnz_h, m, hptr.p);
        FL_LAUNCH_CHECK();
        // rows by (window, heavy count descending, row): two stable passes
        TmpBuf<uint32_t> wkey(m, s);
        k_hyb_lenkey<<<hg(m), kBlock, 0, s>>>(hptr.p, m, wkey.p, rows.p);
        FL_LAUNCH_CHECK();
        radix_sort_pairs(wkey.p, rows.p, m, 16, s);
        k_hyb_winkey<<<hg(m), kBlock, 0, s>>>(rows.p, m, wkey.p);
        FL_LAUNCH_CHECK();
        radix_sort_pairs(wkey.p, rows.p, m, bits_for((uint32_t)((m - 1) >> kHybWinLog)), s);
        k_hyb_pos<<<hg(m), kBlock, 0, s>>>(rows.p, m, A->hyb_iperm.p, A->hyb_perm.p, flag.p);
        TmpBuf<int> pairs(std::max(nslice, 1), s), pscan((size_t)nslice + 1, s);
        k_hyb_pairs<<<hg(nslice), kBlock, 0, s>>>(rows.p, hptr.p, m, nslice, A->hyb_np.p, pairs.p);
        FL_LAUNCH_CHECK();
        exclusive_scan_i32(pairs.p, pscan.p, nslice, s);
        k_hyb_slot_base<<<hg((long long)nslice + 1), kBlock, 0, s>>>(pscan.p, nslice, A->hyb_sb.p);
        FL_LAUNCH_CHECK();
        slots = 64LL * hread(pscan.p + nslice, s);
        if (hread(flag.p, s)) {
            set_error("HYB schedule: preprocessing moved a row out of its window");
            throw HipError{hipErrorInvalidValue};
        }
    }
    // slots past the last slice: the kernel's look-ahead loads past a wave's last slice read
    // kHybPairs pair rows at hyb_slots, and those past a slice's end up to kHybPairs - 1 pair
    // rows beyond the last slice
    const long long padded = slots + 2LL * kHybPairs * 64;
    A->hyb_slots = slots;
    A->hyb_col2.alloc((size_t)padded);
    A->hyb_val64.alloc((size_t)padded * 2);
    k_hyb_fill_u32<<<hg(padded), kBlock, 0, s>>>(A->hyb_col2.p, (uint32_t)H | ((uint32_t)H << 16), padded);
    FL_LAUNCH_CHECK();
    FL_HIP(hipMemsetAsync(A->hyb_val64.p, 0, sizeof(double) * (size_t)padded * 2, s));
    if (m > 0 && nnz_h > 0) {
        k_hyb_fill<<<hg(m), kBlock, 0, s>>>(m, H, rows.p, A->hyb_np.p, A->hyb_sb.p, hptr.p, ridx.p, order.p + nnz_l,
                                            key.p + nnz_l, nl, A->val64.p, A->hyb_col2.p, A->hyb_val64.p);
        FL_LAUNCH_CHECK();
    }
    FL_HIP(hipStreamSynchronize(s));
    A->hyb_light = std::move(L);
    A->hyb_H = H;
    A->hyb_bt = bt;
    A->hyb_R = R;
    A->hyb_grid = cu_count();
    if (A->val32.p) hyb_values(A, FL_VAL_F32, s);
    if (A->pack_wanted) hyb_values(A, FL_VAL_PACKED, s);
    FL_HIP(hipStreamSynchronize(s));
    A->hyb_ready = true;
}

// --------------------------------------------------------------------------- SpMV
// One load group: U pair rows q0 .. q0+U-1 of slice s (P pair rows), and whether it is the
// slice's last group; valid = false past the wave's last slice (its loads read padding).
template <typename V2, int U>
struct HybGroup {
    V2 v[U];
    uint32_t c[U];
    long long s;
    int q0, P;
    bool last, valid;
};

// Per wave: the slices it takes, in snake order (round k: slice k * nw + g0, or k * nw + nw - 1 -
// g0 when k is odd), so that its long and short slices (long first inside every sorting window)
// even out, and the position of the next load group.
struct HybCursor {
    long long s, slot;  // current slice, its first slot + lane
    int k, q, P;        // round, next pair row to load, pair rows of the slice
};

__device__ __forceinline__ long long hyb_slice_of(int k, int nw, int g0) {
    return (long long)k * nw + ((k & 1) ? nw - 1 - g0 : g0);
}

__device__ __forceinline__ void hyb_enter(HybCursor &c, int nslice, const long long *__restrict__ sb,
                                          const int *__restrict__ np, long long pad_slot, int lane) {
    if (c.s < nslice) {
        c.P = np[c.s];
        c.slot = sb[c.s] + lane;
    } else {
        c.P = 0;
        c.slot = pad_slot + lane;
    }
    c.q = 0;
}

// Issues the loads of the cursor's next group (always: past the last slice, or past a slice's pair
// rows, they read padding or the next slice's slots, which hyb_consume skips) and advances the
// cursor, entering the next slice after a slice's last group.
template <typename V2, int U, bool S>
__device__ __forceinline__ void hyb_load(HybGroup<V2, U> &g, HybCursor &c, int nw, int g0, int nslice,
                                         const long long *__restrict__ sb, const int *__restrict__ np,
                                         long long pad_slot, int lane, const V2 *__restrict__ val2,
                                         const uint32_t *__restrict__ col2) {
    g.s = c.s;
    g.q0 = c.q;
    g.P = c.P;
    g.valid = c.s < nslice;
    const long long slot = g.valid ? c.slot + 64LL * c.q : pad_slot + lane;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if constexpr (S) {
            g.v[u] = ld_stream(&val2[slot + 64 * u]);
            g.c[u] = ld_stream(&col2[slot + 64 * u]);
        } else {
            g.v[u] = val2[slot + 64 * u];
            g.c[u] = col2[slot + 64 * u];
        }
    }
    c.q += U;
    g.last = c.q >= c.P;
    if (g.last && g.valid) {
        ++c.k;
        c.s = hyb_slice_of(c.k, nw, g0);
        hyb_enter(c, nslice, sb, np, pad_slot, lane);
    }
}

// Adds the group's pair rows below P (wave-uniform test: the rest belong to the next slice);
// padding entries multiply value 0 with the zero at xs[H]. After a slice's last group, writes its
// 64 sums by sorted position.
template <typename V2, typename X, typename Y, int U>
__device__ __forceinline__ void hyb_consume(double &acc, const HybGroup<V2, U> &g, const X *xs, int m, int lane,
                                            Y *__restrict__ ys) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (g.q0 + u < g.P) {
            const X x0 = xs[g.c[u] & 0xffffu], x1 = xs[g.c[u] >> 16];
            if constexpr (sizeof(X) == 4) acc += (double)__builtin_fmaf(g.v[u].x, x0, g.v[u].y * x1);  // FP32 pair
            else acc = fma(g.v[u].x, x0, fma(g.v[u].y, x1, acc));
        }
    }
    if (g.last) {
        const long long p = g.s * 64 + lane;
        if (p < m) ys[p] = (Y)acc;
        acc = 0.0;
    }
}

// Persistent workgroups of 16 waves: x of the H heavy columns (and a zero after them) is staged
// in LDS once; then every wave runs through its slices as one stream of load groups, the loads
// of the next group (possibly the next slice's first) in flight while the current one is summed.
// Every group's loads are unconditional, since a conditional load would make the compiler wait
// for all loads in flight before each sum.
// S: read the slots with the streaming hint (stream_reads).
template <typename V2, typename X, typename Y, int U, bool S>
__global__ __launch_bounds__(1024) void k_hyb(int m, int H, int nslice, const long long *__restrict__ sb,
                                              const int *__restrict__ np, long long pad_slot,
                                              const V2 *__restrict__ val2, const uint32_t *__restrict__ col2,
                                              const X *__restrict__ xh, Y *__restrict__ ys) {
    extern __shared__ double hyb_lds[];
    X *xs = reinterpret_cast<X *>(hyb_lds);
    // stage x: 8 loads in flight per thread before their LDS stores (one load-store round trip
    // per element would keep every wave waiting on L2 latency about 10 us before its first slice)
    constexpr int kStage = 8;
    for (int base = 0; base <= H; base += 1024 * kStage) {
        X r[kStage];
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int i = base + j * 1024 + (int)threadIdx.x;
            r[j] = i < H ? xh[i] : X(0);
        }
#pragma unroll
        for (int j = 0; j < kStage; ++j) {
            const int i = base + j * 1024 + (int)threadIdx.x;
            if (i <= H) xs[i] = r[j];
        }
    }
    __syncthreads();
    const int wave = uni((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int nw = (int)gridDim.x * 16, g0 = (int)blockIdx.x * 16 + wave;
    HybCursor c;
    c.k = 0;
    c.s = hyb_slice_of(0, nw, g0);
    hyb_enter(c, nslice, sb, np, pad_slot, lane);
    HybGroup<V2, U> a, b;
    double acc = 0.0;
    hyb_load<V2, U, S>(a, c, nw, g0, nslice, sb, np, pad_slot, lane, val2, col2);
    while (a.valid) {
        hyb_load<V2, U, S>(b, c, nw, g0, nslice, sb, np, pad_slot, lane, val2, col2);
        hyb_consume<V2, X, Y, U>(acc, a, xs, m, lane, ys);
        if (!b.valid) break;
        hyb_load<V2, U, S>(a, c, nw, g0, nslice, sb, np, pad_slot, lane, val2, col2);
        hyb_consume<V2, X, Y, U>(acc, b, xs, m, lane, ys);
    }
}

// y = ys in row order (assign) or y += it: the HYB schedule when no light entries exist
template <typename Y, bool ASSIGN>
__global__ void k_hyb_unperm(int m, const uint16_t *__restrict__ iperm, const Y *__restrict__ ys, Y *__restrict__ y) {
    HSTRIDE(r, m) {
        const Y v = ys[(r & ~(long long)(kHybWin - 1)) + iperm[r]];
        if constexpr (ASSIGN) y[r] = v;
        else y[r] += v;
    }
}

int hyb_check(const fl_matrix *A, int kind) {
    if (!A->hyb_ready) {
        set_error("fl_spmv: HYB schedule not prepared");
        return FL_ERR_INVALID;
    }
    if (kind != FL_VAL_F64 && kind != FL_VAL_F32 && kind != FL_VAL_PACKED) {
        set_error("fl_spmv: the HYB schedule supports FL_VAL_F64, FL_VAL_F32 and FL_VAL_PACKED");
        return FL_ERR_INVALID;
    }
    if (kind == FL_VAL_F64 && A->hyb_nnz > 0 && A->hyb_H > kHybMaxH64) {
        set_error("fl_spmv: FL_VAL_F64 with the HYB schedule needs at most %d heavy columns (prepared: %d)",
                  kHybMaxH64, A->hyb_H);
        return FL_ERR_INVALID;
    }
    if (kind != FL_VAL_F64 && A->hyb_nnz > 0 && !A->hyb_val32.p) {
        set_error("fl_spmv: FP32 values not prepared");
        return FL_ERR_INVALID;
    }
    const fl_matrix *L = A->hyb_light.get();
    if (L->tilenum > 0) {
        if (kind == FL_VAL_F32 && !L->val32.p) {
            set_error("fl_spmv: FP32 values not prepared");
            return FL_ERR_INVALID;
        }
        if (kind == FL_VAL_PACKED && !L->pb_pack.p) {
            set_error("fl_spmv: no packed values for the HYB schedule's PB part (fl_matrix_prepare_values("
                      "FL_VAL_PACKED) not called, or fewer than 12 mantissa bits would remain)");
            return FL_ERR_INVALID;
        }
    }
    return FL_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t hyb_scratch_bytes(const fl_matrix *A, int kind) {
    const size_t ysz = kind == FL_VAL_F64 ? sizeof(double) : sizeof(float);
    const fl_matrix *L = A->hyb_light.get();
    const size_t bins = (L && L->tilenum > 0) ? (size_t)L->pb_ring * ysz : 0;
    return align256((size_t)A->m * ysz) + bins;
}

template <typename V2, typename X, typename Y>
static void launch_hyb(const fl_matrix *A, const V2 *val2, const X *xh, Y *ys, bool stream, hipStream_t s) {
    constexpr int U = kHybPairs;
    const size_t lds = ((size_t)A->hyb_H + 1) * sizeof(X);
    auto k = stream ? k_hyb<V2, X, Y, U, true> : k_hyb<V2, X, Y, U, false>;
    FL_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // two workgroups per compute unit (32 waves) when their LDS fits, else one
    const int per_cu = lds <= 80 * 1024 ? 2 : 1;
    const int grid = std::max(1, std::min(A->hyb_grid * per_cu, (A->hyb_nslice + 15) / 16));
    k<<<grid, 1024, lds, s>>>(A->m, A->hyb_H, A->hyb_nslice, A->hyb_sb.p, A->hyb_np.p, A->hyb_slots, val2,
                              A->hyb_col2.p, xh, ys);
    FL_LAUNCH_CHECK();
}

// Called by spmv after hyb_check. x = sortx (n entries).
int spmv_hyb(fl_matrix *A, int kind, const void *x, void *y, bool assign, hipStream_t s, void *scratch) {
    fl_matrix *L = A->hyb_light.get();
    const bool dbl = kind == FL_VAL_F64;
    const int m = A->m, nl = L->n;
    void *ys = nullptr, *bins = nullptr;
    if (scratch) {
        ys = scratch;
        bins = (char *)scratch + align256((size_t)m * (dbl ? 8 : 4));
    } else if (A->hyb_nnz > 0) {
        if (dbl) {
            if (!A->hyb_ys64.p) A->hyb_ys64.alloc(std::max(m, 1));
            ys = A->hyb_ys64.p;
        } else {
            if (!A->hyb_ys32.p) A->hyb_ys32.alloc(std::max(m, 1));
            ys = A->hyb_ys32.p;
        }
    }
    const bool heavy = A->hyb_nnz > 0;
    // one streaming decision for the three kernels, which share the caches
    const long long heavy_bytes = heavy ? A->hyb_slots * (long long)(dbl ? 16 + 4 : 8 + 4) : 0;
    const bool stream =
        stream_reads(heavy_bytes + (L->tilenum > 0 ? batched_stream_bytes(L, FL_KERNEL_PB, kind) : 0));
    if (heavy) {
        if (dbl) launch_hyb<double2, double, double>(A, (const double2 *)A->hyb_val64.p, (const double *)x + nl, (double *)ys, stream, s);
        else launch_hyb<float2, float, float>(A, (const float2 *)A->hyb_val32.p, (const float *)x + nl, (float *)ys, stream, s);
    }
    if (L->tilenum > 0) {
        PbFold fold;
        if (heavy) {
            fold.ys = ys;
            fold.iperm = A->hyb_iperm.p;
            fold.perm = A->hyb_perm.p;
        }
        return spmv_batched(L, FL_KERNEL_PB, kind, x, y, assign, s, bins, &fold, stream ? 1 : 0);
    }
    if (heavy) {
        if (dbl) {
            if (assign) k_hyb_unperm<double, true><<<hg(m), kBlock, 0, s>>>(m, A->hyb_iperm.p, (const double *)ys, (double *)y);
            else k_hyb_unperm<double, false><<<hg(m), kBlock, 0, s>>>(m, A->hyb_iperm.p, (const double *)ys, (double *)y);
        } else {
            if (assign) k_hyb_unperm<float, true><<<hg(m), kBlock, 0, s>>>(m, A->hyb_iperm.p, (const float *)ys, (float *)y);
            else k_hyb_unperm<float, false><<<hg(m), kBlock, 0, s>>>(m, A->hyb_iperm.p, (const float *)ys, (float *)y);
        }
        FL_LAUNCH_CHECK();
    } else if (assign) {
        zero_fill(y, (size_t)m * (dbl ? 8 : 4), s);
    }
    return FL_OK;
}

}  // namespace fl
