// Multi-right-hand-side SpMM (ROW schedule), block-diagonal batching of many matrices, zero
// fill and hipGraph capture. No reference counterpart: the reference multiplies one matrix by
// one vector per call (src/FastLoad_GPU.h:275-409).
//
// ROW: Y (m x k, row-major) = A * X (n x k, row-major, original column order). With k >= 8 a
// FastLoad-style column sweep would need k scattered y updates per non-zero (global atomics run
// at ~20 G cache lines/s on MI355X); instead the entries are re-ordered by row once, a row is
// owned by k lanes, X rows are gathered whole (k*4 bytes, a full 128-byte line at k = 32) and
// each Y row is written once, without atomics.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fl_matrix.h"

namespace fl {

static unsigned sgb(long long count) { return std::max(1u, std::min<unsigned>(grid_for(count), 65536u)); }

#define GSTRIDE(i, count) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += (long long)gridDim.x * blockDim.x)

// original column of every entry (tile order): tile t, lane l -> perm[tile_colidx[t] + l]
__global__ void k_col_of(int T, int W, const int *__restrict__ tile_ptr, const int *__restrict__ tile_len,
                         const int *__restrict__ tile_colidx, const int *__restrict__ perm, int *__restrict__ col_of) {
    GSTRIDE(g, (long long)T * W) {
        const long long t = g / W;
        const int l = (int)(g - t * W);
        if (l < tile_len[t]) col_of[tile_ptr[t] + l] = perm[tile_colidx[t] + l];
    }
}

__global__ void k_rows_key(const int *__restrict__ col, uint32_t *__restrict__ key, int *__restrict__ order, int count) {
    GSTRIDE(i, count) {
        key[i] = (uint32_t)col[i];
        order[i] = (int)i;
    }
}

__global__ void k_rows_key2(const int *__restrict__ row, const int *__restrict__ order, uint32_t *__restrict__ key,
                            int count) {
    GSTRIDE(i, count) key[i] = (uint32_t)row[order[i]];
}

__global__ void k_rows_fill(const int *__restrict__ order, const int *__restrict__ col_of,
                            const double *__restrict__ val, int *__restrict__ rcol, double *__restrict__ rv64,
                            float *__restrict__ rv32, int count) {
    GSTRIDE(i, count) {
        const int g = order[i];
        rcol[i] = col_of[g];
        const double v = val[g];
        rv64[i] = v;
        rv32[i] = (float)v;
    }
}

__global__ void k_row_bounds(const uint32_t *__restrict__ k, int count, int m, int *__restrict__ start) {
    GSTRIDE(r, (long long)m + 1) {
        int lo = 0, hi = count;
        while (lo < hi) {
            const int mid = (int)(((long long)lo + hi) >> 1);
            if (k[mid] < (uint32_t)r) lo = mid + 1; else hi = mid;
        }
        start[r] = lo;
    }
}

// out[k] = first row whose entries start at or after k/8 of nnz (out[0] = 0, out[8] = m)
__global__ void k_row_split(const int *__restrict__ row_ptr, int m, int nnz, int *__restrict__ out) {
    const int k = threadIdx.x;
    if (k > 8) return;
    const long long target = (long long)nnz * k / 8;
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (row_ptr[mid] < target) lo = mid + 1; else hi = mid;
    }
    out[k] = k == 8 ? m : lo;
}

static void compute_row_split(fl_matrix *A, hipStream_t s) {
    DevBuf<int> d(9);
    k_row_split<<<1, 64, 0, s>>>(A->row_ptr.p, A->m, A->nnz, d.p);
    FL_LAUNCH_CHECK();
    FL_HIP(hipMemcpyAsync(A->row_split.r, d.p, sizeof(int) * 9, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
}

void row_prepare(fl_matrix *A, hipStream_t s) {
    if (A->row_ready) return;
    const int nnz = A->nnz, m = A->m, T = A->tilenum;
    A->row_ptr.alloc((size_t)m + 1);
    A->row_col.alloc(std::max(nnz, 1));
    A->row_val32.alloc(std::max(nnz, 1));
    A->row_val64.alloc(std::max(nnz, 1));
    if (nnz == 0) {
        FL_HIP(hipMemsetAsync(A->row_ptr.p, 0, sizeof(int) * ((size_t)m + 1), s));
        compute_row_split(A, s);
        A->row_ready = true;
        return;
    }
    DevBuf<int> col_of(nnz), order(nnz);
    DevBuf<uint32_t> key(nnz);
    k_col_of<<<sgb((long long)T * A->W), kBlock, 0, s>>>(T, A->W, A->tile_ptr.p, A->tile_len.p, A->tile_colidx.p,
                                                         A->perm.p, col_of.p);
    // (row, original column) order: stable by column, then stable by row. The order inside a row
    // then depends only on A, not on the column sort, so a block of a batch sums exactly as the
    // same matrix alone.
    k_rows_key<<<sgb(nnz), kBlock, 0, s>>>(col_of.p, key.p, order.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(A->n - 1, 0)), s);
    k_rows_key2<<<sgb(nnz), kBlock, 0, s>>>(A->sortrowidx.p, order.p, key.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(m - 1, 0)), s);
    k_rows_fill<<<sgb(nnz), kBlock, 0, s>>>(order.p, col_of.p, A->val64.p, A->row_col.p, A->row_val64.p,
                                            A->row_val32.p, nnz);
    k_row_bounds<<<sgb((long long)m + 1), kBlock, 0, s>>>(key.p, nnz, m, A->row_ptr.p);
    FL_LAUNCH_CHECK();
    compute_row_split(A, s);
    A->row_ready = true;
}

__global__ void k_row_val32(const double *__restrict__ v64, float *__restrict__ v32, int count) {
    GSTRIDE(i, count) v32[i] = (float)v64[i];
}

// The row order read back from a file: row_ptr [m+1] must be a pointer array over nnz and every
// column inside [0, n) before the kernels may index with them.
int row_install(fl_matrix *A, const int *row_ptr, const int *row_col, const double *row_val, hipStream_t s) {
    const int m = A->m, n = A->n, nnz = A->nnz;
    bool ok = row_ptr[0] == 0 && row_ptr[m] == nnz;
    for (int r = 0; ok && r < m; ++r) ok = row_ptr[r] <= row_ptr[r + 1];
    for (long long i = 0; ok && i < nnz; ++i) ok = row_col[i] >= 0 && row_col[i] < n;
    if (!ok) {
        set_error("fl_matrix_load: row order invalid (row pointers or columns out of range)");
        return FL_ERR_INVALID;
    }
    A->row_ptr.alloc((size_t)m + 1);
    A->row_col.alloc(std::max(nnz, 1));
    A->row_val64.alloc(std::max(nnz, 1));
    A->row_val32.alloc(std::max(nnz, 1));
    FL_HIP(hipMemcpyAsync(A->row_ptr.p, row_ptr, sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice, s));
    if (nnz) {
        FL_HIP(hipMemcpyAsync(A->row_col.p, row_col, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(A->row_val64.p, row_val, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, s));
        k_row_val32<<<sgb(nnz), kBlock, 0, s>>>(A->row_val64.p, A->row_val32.p, nnz);
        FL_LAUNCH_CHECK();
    }
    compute_row_split(A, s);
    A->row_ready = true;
    return FL_OK;
}

// Vector of VEC consecutive T (VEC = 1, 2 or 4), loaded and stored as one access.
template <typename T, int VEC>
struct RowVec;
template <> struct RowVec<float, 1> { using type = float; };
template <> struct RowVec<float, 2> { using type = float2; };
template <> struct RowVec<float, 4> { using type = float4; };
template <> struct RowVec<double, 1> { using type = double; };
template <> struct RowVec<double, 2> { using type = double2; };

// K / VEC lanes own one row (64 * VEC / K rows per wave); lane j accumulates columns
// [j*VEC, j*VEC + VEC) of the row's Y block, gathering them from X as one VEC-wide load per
// entry. Every Y element sums its row's entries in entry order whatever VEC is. Persistent:
// waves stride over row groups. ACC = accumulator type.
// The waves of XCD x (workgroups with blockIdx % 8 == x; gridDim a multiple of 8) stride over the
// row groups of that XCD's range [sp.r[x], sp.r[x+1]), so the rows in flight on one XCD at any
// time are a contiguous window, and on matrices whose rows have local columns (bands, stencils,
// block-diagonal batches) the X rows they gather stay in that XCD's L2.
struct XcdRows {
    long long wave, nwaves;
    int r0, r1;
};

__device__ __forceinline__ XcdRows xcd_rows(const RowSplit &sp) {
    XcdRows w;
    const int xcd = blockIdx.x & 7;
    w.wave = uni((int)((((blockIdx.x >> 3) * blockDim.x) + threadIdx.x) >> 6));  // wave-uniform loop
    w.nwaves = ((long long)(gridDim.x >> 3) * blockDim.x) >> 6;
    w.r0 = sp.r[xcd];
    w.r1 = sp.r[xcd + 1];
    return w;
}

template <typename T, typename ACC, int K, int VEC, int UNR>
__global__ __launch_bounds__(kBlock) void k_spmm_row(RowSplit sp, const int *__restrict__ row_ptr,
                                                     const int *__restrict__ row_col, const T *__restrict__ row_val,
                                                     const T *__restrict__ X, T *__restrict__ Y, int accumulate) {
    using V = typename RowVec<T, VEC>::type;
    constexpr int L = K / VEC;   // lanes per row
    constexpr int RPW = 64 / L;  // rows per wave
    const int lane = threadIdx.x & 63;
    const int sub = lane / L, j = lane % L;
    const V *__restrict__ Xv = reinterpret_cast<const V *>(X);
    const XcdRows xw = xcd_rows(sp);
    const long long groups = ((long long)(xw.r1 - xw.r0) + RPW - 1) / RPW;
    for (long long gq = xw.wave; gq < groups; gq += xw.nwaves) {
        const long long r = xw.r0 + gq * RPW + sub;
        if (r >= xw.r1) continue;
        const int e0 = row_ptr[r], e1 = row_ptr[r + 1];
        ACC acc[VEC];
#pragma unroll
        for (int q = 0; q < VEC; ++q) acc[q] = 0;
        auto add = [&](T v, const V &xv) {
            const T *xs = reinterpret_cast<const T *>(&xv);
#pragma unroll
            for (int q = 0; q < VEC; ++q) acc[q] += (ACC)v * (ACC)xs[q];
        };
        int e = e0;
        for (; e + UNR <= e1; e += UNR) {
            int c[UNR];
            T v[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                c[u] = row_col[e + u];
                v[u] = row_val[e + u];
            }
            V xv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) xv[u] = Xv[(long long)c[u] * L + j];
#pragma unroll
            for (int u = 0; u < UNR; ++u) add(v[u], xv[u]);
        }
        for (; e < e1; ++e) add(row_val[e], Xv[(long long)row_col[e] * L + j]);
        T *out = Y + r * K + (long long)j * VEC;
#pragma unroll
        for (int q = 0; q < VEC; ++q) out[q] = accumulate ? (T)((ACC)out[q] + acc[q]) : (T)acc[q];
    }
}

// k = 1: LPR = 4 lanes per row; lane j sums entries e0 + j, e0 + j + 4, ... and the four partial
// sums are added by a butterfly over the row's lanes. The order depends only on the row, so a
// block of a stacked batch still equals its matrix run alone bit for bit. Four short gather
// chains per row instead of one long one: the 67 M-entry power-law matrix takes 0.61 ms instead
// of 0.88 ms with one lane per row (0.67 ms with 2 lanes, 0.59 ms with 8, which loses on short
// rows: 9.3 vs 12.2 us on a 2.4 M-entry block-diagonal matrix of 6-entry rows).
template <typename T, typename ACC, int LPR, int UNR>
__global__ __launch_bounds__(kBlock) void k_spmv_rowl(RowSplit sp, const int *__restrict__ row_ptr,
                                                      const int *__restrict__ row_col, const T *__restrict__ row_val,
                                                      const T *__restrict__ X, T *__restrict__ Y, int accumulate) {
    constexpr int RPW = 64 / LPR;
    const int lane = threadIdx.x & 63;
    const int sub = lane / LPR, j = lane % LPR;
    const XcdRows xw = xcd_rows(sp);
    const long long groups = ((long long)(xw.r1 - xw.r0) + RPW - 1) / RPW;
    for (long long gq = xw.wave; gq < groups; gq += xw.nwaves) {  // uniform per wave: all lanes reach the butterfly
        const long long r = xw.r0 + gq * RPW + sub;
        const bool live = r < xw.r1;
        int e0 = 0, e1 = 0;
        if (live) {
            e0 = row_ptr[r];
            e1 = row_ptr[r + 1];
        }
        ACC acc = 0;
        int e = e0 + j;
        for (; e + (UNR - 1) * LPR < e1; e += UNR * LPR) {
            int c[UNR];
            T v[UNR], xv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                c[u] = row_col[e + u * LPR];
                v[u] = row_val[e + u * LPR];
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) xv[u] = X[c[u]];
#pragma unroll
            for (int u = 0; u < UNR; ++u) acc += (ACC)v[u] * (ACC)xv[u];
        }
        for (; e < e1; e += LPR) acc += (ACC)row_val[e] * (ACC)X[row_col[e]];
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (live && j == 0) Y[r] = accumulate ? (T)((ACC)Y[r] + acc) : (T)acc;
    }
}

// Persistent grid: enough waves for the rows, at most 2048 workgroups, a multiple of 8 (one share
// per XCD).
static unsigned row_grid(long long m, int rows_per_wave) {
    const long long groups = (m + rows_per_wave - 1) / rows_per_wave;
    const unsigned g = std::min<unsigned>(grid_for(groups * 64), 256u * 8u);
    return std::max(8u, (g + 7u) & ~7u);
}

template <typename T, typename ACC>
static void launch_row_k1(const fl_matrix *A, const T *val, const T *X, T *Y, int accumulate, hipStream_t s) {
    constexpr int LPR = 4;
    k_spmv_rowl<T, ACC, LPR, 4><<<row_grid(A->m, 64 / LPR), kBlock, 0, s>>>(A->row_split, A->row_ptr.p, A->row_col.p,
                                                                            val, X, Y, accumulate);
    FL_LAUNCH_CHECK();
}

template <typename T, typename ACC, int K, int VEC>
static void launch_row_v(const fl_matrix *A, const T *val, const T *X, T *Y, int accumulate, hipStream_t s) {
    constexpr int RPW = 64 * VEC / K;
    k_spmm_row<T, ACC, K, VEC, 4><<<row_grid(A->m, RPW), kBlock, 0, s>>>(A->row_split, A->row_ptr.p, A->row_col.p, val,
                                                                         X, Y, accumulate);
    FL_LAUNCH_CHECK();
}

// Widest vector that keeps at least 4 lanes per row (k >= 4), is at most 16 bytes and that X
// and Y are aligned for. On the 512-matrix batch: k = 32 took 3.14 ms with 16-byte lanes
// against 5.28 ms with 4-byte lanes and 3.84 ms with 32-byte lanes; k = 8 took 1.72 ms with
// 8-byte lanes (4 per row) against 1.78 ms with 16-byte lanes (2 per row).
template <typename T, typename ACC, int K>
static void launch_row(const fl_matrix *A, const T *val, const T *X, T *Y, int accumulate, hipStream_t s) {
    constexpr int VT = 16 / (int)sizeof(T);
    constexpr int VK = K / 4 > 1 ? K / 4 : 1;
    constexpr int VMAX = VT < VK ? VT : VK;
    const uintptr_t mis = reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y);
    if constexpr (VMAX >= 4) {
        if ((mis & 15) == 0) return launch_row_v<T, ACC, K, 4>(A, val, X, Y, accumulate, s);
    }
    if constexpr (VMAX >= 2) {
        if ((mis & (2 * sizeof(T) - 1)) == 0) return launch_row_v<T, ACC, K, 2>(A, val, X, Y, accumulate, s);
    }
    launch_row_v<T, ACC, K, 1>(A, val, X, Y, accumulate, s);
}

template <typename T, typename ACC>
static int dispatch_row(const fl_matrix *A, const T *val, int k, const T *X, T *Y, int acc, hipStream_t s) {
    switch (k) {
        case 1: launch_row_k1<T, ACC>(A, val, X, Y, acc, s); return FL_OK;
        case 2: launch_row<T, ACC, 2>(A, val, X, Y, acc, s); return FL_OK;
        case 4: launch_row<T, ACC, 4>(A, val, X, Y, acc, s); return FL_OK;
        case 8: launch_row<T, ACC, 8>(A, val, X, Y, acc, s); return FL_OK;
        case 16: launch_row<T, ACC, 16>(A, val, X, Y, acc, s); return FL_OK;
        case 32: launch_row<T, ACC, 32>(A, val, X, Y, acc, s); return FL_OK;
        case 64: launch_row<T, ACC, 64>(A, val, X, Y, acc, s); return FL_OK;
        default:
            set_error("fl_spmm: k must be 1, 2, 4, 8, 16, 32 or 64");
            return FL_ERR_INVALID;
    }
}

int spmm_row(const fl_matrix *A, int kind, int k, const void *X, void *Y, int accumulate, hipStream_t s) {
    if (!A->row_ready) {
        set_error("fl_spmm: row order not prepared (fl_matrix_prepare_rows)");
        return FL_ERR_INVALID;
    }
    if (A->m == 0) return FL_OK;
    if (kind == FL_VAL_F64) return dispatch_row<double, double>(A, A->row_val64.p, k, (const double *)X, (double *)Y, accumulate, s);
    if (kind == FL_VAL_F32) return dispatch_row<float, float>(A, A->row_val32.p, k, (const float *)X, (float *)Y, accumulate, s);
    set_error("fl_spmm: FL_VAL_F64 or FL_VAL_F32 only");
    return FL_ERR_INVALID;
}

// ------------------------------------------------------------------------- zero fill
__global__ void k_zero(uint32_t *__restrict__ p, long long words) {
    GSTRIDE(i, words) p[i] = 0u;
}

__global__ void k_zero16(uint4 *__restrict__ p, long long quads) {
    GSTRIDE(i, quads) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_zero_bytes(unsigned char *__restrict__ p, long long bytes) {
    GSTRIDE(i, bytes) p[i] = 0;
}

// Zero fill with a kernel (graph-capturable; no memset node).
void zero_fill(void *d, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    if ((reinterpret_cast<uintptr_t>(d) & 15) == 0 && bytes >= 16) {  // 16-byte stores for the bulk
        const long long quads = (long long)(bytes / 16);
        k_zero16<<<sgb(quads), kBlock, 0, s>>>((uint4 *)d, quads);
        FL_LAUNCH_CHECK();
        d = (unsigned char *)d + quads * 16;
        bytes -= (size_t)quads * 16;
        if (!bytes) return;
    }
    const long long words = (long long)(bytes / 4);
    if (words) {
        k_zero<<<sgb(words), kBlock, 0, s>>>((uint32_t *)d, words);
        FL_LAUNCH_CHECK();
    }
    const long long tail = (long long)(bytes - (size_t)words * 4);
    if (tail) {
        k_zero_bytes<<<1, 64, 0, s>>>((unsigned char *)d + words * 4, tail);
        FL_LAUNCH_CHECK();
    }
}

}  // namespace fl

using namespace fl;

extern "C" {

int fl_matrix_prepare_rows(fl_matrix *A, void *stream) {
    FL_API_BEGIN
    if (!A) {
        set_error("fl_matrix_prepare_rows: NULL");
        return FL_ERR_INVALID;
    }
    row_prepare(A, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_spmm(fl_matrix *A, int val_kind, int k, const void *d_X, void *d_Y, int accumulate, void *stream) {
    FL_API_BEGIN
    if (!A || (A->m && (!d_X || !d_Y))) {
        set_error("fl_spmm: NULL");
        return FL_ERR_INVALID;
    }
    return spmm_row(A, val_kind, k, d_X, d_Y, accumulate, (hipStream_t)stream);
    FL_API_END
}

int fl_zero_device(void *d, size_t bytes, void *stream) {
    FL_API_BEGIN
    if (bytes && !d) {
        set_error("fl_zero_device: NULL");
        return FL_ERR_INVALID;
    }
    zero_fill(d, bytes, (hipStream_t)stream);
    return FL_OK;
    FL_API_END
}

int fl_batch_assemble(int count, const int *m, const int *n, const int *const *colptr, const int *const *rowidx,
                      const double *const *val, int *M, int *N, int **colptr_out, int **rowidx_out, double **val_out,
                      int *row_off, int *col_off) {
    if (count < 0 || !m || !n || !colptr || !rowidx || !val || !M || !N || !colptr_out || !rowidx_out || !val_out ||
        !row_off || !col_off) {
        set_error("fl_batch_assemble: NULL or negative count");
        return FL_ERR_INVALID;
    }
    long long MM = 0, NN = 0, NZ = 0;
    row_off[0] = col_off[0] = 0;
    for (int i = 0; i < count; ++i) {
        if (m[i] < 0 || n[i] < 0 || !colptr[i]) {
            set_error("fl_batch_assemble: matrix %d invalid", i);
            return FL_ERR_INVALID;
        }
        // colptr[i] must be a column pointer (starts at 0, never decreases) before its last entry
        // sizes the output and its ranges are copied
        const int *cpi = colptr[i];
        bool ok = cpi[0] == 0;
        for (int c = 0; ok && c < n[i]; ++c) ok = cpi[c] <= cpi[c + 1];
        if (!ok) {
            set_error("fl_batch_assemble: matrix %d colptr is not a column pointer (colptr[0] != 0 or decreasing)", i);
            return FL_ERR_INVALID;
        }
        if (cpi[n[i]] > 0 && (!rowidx[i] || !val[i])) {
            set_error("fl_batch_assemble: matrix %d has entries but NULL rowidx/val", i);
            return FL_ERR_INVALID;
        }
        MM += m[i];
        NN += n[i];
        NZ += colptr[i][n[i]];
        if (MM > 0x7fffffffLL || NN > 0x7fffffffLL || NZ > 0x7fffffffLL) {
            set_error("fl_batch_assemble: batch exceeds int32 indexing");
            return FL_ERR_SIZE;
        }
        row_off[i + 1] = (int)MM;
        col_off[i + 1] = (int)NN;
    }
    int *cp = (int *)malloc(sizeof(int) * ((size_t)NN + 1));
    int *ri = (int *)malloc(sizeof(int) * (size_t)std::max<long long>(NZ, 1));
    double *vv = (double *)malloc(sizeof(double) * (size_t)std::max<long long>(NZ, 1));
    if (!cp || !ri || !vv) {
        free(cp); free(ri); free(vv);
        set_error("fl_batch_assemble: out of memory");
        return FL_ERR_ALLOC;
    }
    long long pos = 0;
    for (int i = 0; i < count; ++i) {
        const int ro = row_off[i], co = col_off[i];
        for (int c = 0; c < n[i]; ++c) {
            cp[co + c] = (int)pos;
            for (int e = colptr[i][c]; e < colptr[i][c + 1]; ++e, ++pos) {
                if (rowidx[i][e] < 0 || rowidx[i][e] >= m[i]) {
                    free(cp); free(ri); free(vv);
                    set_error("fl_batch_assemble: matrix %d row index out of range", i);
                    return FL_ERR_RANGE;
                }
                ri[pos] = ro + rowidx[i][e];
                vv[pos] = val[i][e];
            }
        }
    }
    cp[NN] = (int)pos;
    *M = (int)MM;
    *N = (int)NN;
    *colptr_out = cp;
    *rowidx_out = ri;
    *val_out = vv;
    return FL_OK;
}

// ---------------------------------------------------------------------------- hipGraph
int fl_graph_begin_capture(void *stream) {
    FL_API_BEGIN
    if (!stream) {
        set_error("fl_graph_begin_capture: capture needs a created stream, not the default stream");
        return FL_ERR_INVALID;
    }
    FL_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
    return FL_OK;
    FL_API_END
}

int fl_graph_end_capture(void *stream, void **graph_exec) {
    FL_API_BEGIN
    if (!stream || !graph_exec) {
        set_error("fl_graph_end_capture: NULL");
        return FL_ERR_INVALID;
    }
    hipGraph_t g = nullptr;
    FL_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FL_HIP(e);
    *graph_exec = ge;
    return FL_OK;
    FL_API_END
}

int fl_graph_launch(void *graph_exec, void *stream) {
    FL_API_BEGIN
    FL_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return FL_OK;
    FL_API_END
}

int fl_graph_destroy(void *graph_exec) {
    FL_API_BEGIN
    if (graph_exec) FL_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return FL_OK;
    FL_API_END
}

}  // extern "C"
