// FastLoad preprocessing on the GPU: CSR->CSC, stable column sort, level table, tile
// enumeration, tile data layout, segment detection and tile classification.
//
// Reference pipeline (skyler-ruiter/FastLoad): csr_to_csc (src/csr2csc_cuda.h:51-135),
// ColSort (src/ColSort_GPU.h:46-137), formatTransform (src/formatTransform_GPU.h:289-596).
// The reference's quadratic prefix loops (ColSort_GPU.h:30-33, formatTransform_GPU.h:41-47,
// 123-129) and per-tile serial loops are replaced by scans and closed forms:
//   level k   : sorted position col_k where the length steps up; base_k = length before the
//               step; tall_k = step size
//   tiles     : level k has T_k = ceil((n-col_k)/W) chunks x tall_k depths, ordered
//               (level, chunk, depth); tile t in level k: j = local / tall_k, e = local % tall_k
//   tile_ptr  : nnzoff_k + j*tall_k*W + e*len  (only the last chunk of a level is short)
//   element   : (sorted column colidx+l, depth base_k+e)
#include <algorithm>
#include <vector>

#include "fl_matrix.h"

namespace fl {

// ----------------------------------------------------------------------- small kernels
__global__ void k_iota(int *__restrict__ v, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        v[i] = (int)i;
}

__global__ void k_copy_keys(const int *__restrict__ in, uint32_t *__restrict__ out, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        out[i] = (uint32_t)in[i];
}

__global__ void k_keys_to_int(const uint32_t *__restrict__ in, int *__restrict__ out, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        out[i] = (int)in[i];
}

template <typename T>
__global__ void k_gather(const T *__restrict__ src, const int *__restrict__ idx, T *__restrict__ dst, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

static unsigned stride_grid(long long count) { return std::max(1u, std::min<unsigned>(grid_for(count), 65536u)); }

// ---------------------------------------------------------------------- input checks
// flag bit 0: colptr not monotone / wrong ends; bit 1: row index out of range.
__global__ void k_check_ptr(const int *__restrict__ ptr, int count, int nnz, int *__restrict__ flag) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const int a = ptr[i], b = ptr[i + 1];
        if (a > b || a < 0 || b > nnz) atomicOr(flag, 1);
    }
}

__global__ void k_check_idx(const int *__restrict__ idx, int count, int limit, int *__restrict__ flag) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const int r = idx[i];
        if (r < 0 || r >= limit) atomicOr(flag, 2);
    }
}

// Returns 0 if ptr[0]==0, ptr[count]==nnz, ptr monotone and every idx in [0, limit).
int check_compressed(const int *d_ptr, int count, const int *d_idx, int nnz, int limit, hipStream_t s) {
    TmpBuf<int> flag(1, s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    if (count > 0) {
        k_check_ptr<<<stride_grid(count), kBlock, 0, s>>>(d_ptr, count, nnz, flag.p);
        FL_LAUNCH_CHECK();
    }
    if (nnz > 0) {
        k_check_idx<<<stride_grid(nnz), kBlock, 0, s>>>(d_idx, nnz, limit, flag.p);
        FL_LAUNCH_CHECK();
    }
    int ends[2] = {0, 0}, h = 0;
    FL_HIP(hipMemcpyAsync(&ends[0], d_ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(&ends[1], d_ptr + count, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (ends[0] != 0 || ends[1] != nnz) h |= 1;
    return h;
}

// ------------------------------------------------------------------------------ csr2csc
__global__ void k_row_of(const int *__restrict__ rowptr, int m, int *__restrict__ row_of) {
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += (long long)gridDim.x * blockDim.x)
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) row_of[j] = (int)r;
}

__global__ void k_count(const uint32_t *__restrict__ keys, int count, int *__restrict__ counts) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        atomicAdd(&counts[keys[i]], 1);
}

__global__ void k_csc_fill(const int *__restrict__ order, const int *__restrict__ row_of, const double *__restrict__ val,
                           int *__restrict__ cscrow, double *__restrict__ cscval, int count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const int o = order[i];
        cscrow[i] = row_of[o];
        cscval[i] = val[o];
    }
}

// CSR -> CSC: stable LSD radix sort of the column indices with the CSR position as payload;
// stability keeps rows ascending inside each column (CSR positions ascend with the row).
void csr2csc_device(int m, int n, int nnz, const int *d_rowptr, const int *d_colind, const double *d_val,
                    double *d_cscval, int *d_cscrow, int *d_colptr, int *d_nnzpercol, hipStream_t s) {
    FL_HIP(hipMemsetAsync(d_nnzpercol, 0, sizeof(int) * std::max(n, 1), s));
    if (nnz > 0) {
        TmpBuf<uint32_t> keys(nnz, s);
        TmpBuf<int> order(nnz, s), row_of(nnz, s);
        k_copy_keys<<<stride_grid(nnz), kBlock, 0, s>>>(d_colind, keys.p, nnz);
        k_iota<<<stride_grid(nnz), kBlock, 0, s>>>(order.p, nnz);
        k_row_of<<<stride_grid(m), kBlock, 0, s>>>(d_rowptr, m, row_of.p);
        FL_LAUNCH_CHECK();
        radix_sort_pairs(keys.p, order.p, nnz, bits_for((uint32_t)std::max(n - 1, 0)), s);
        k_csc_fill<<<stride_grid(nnz), kBlock, 0, s>>>(order.p, row_of.p, d_val, d_cscrow, d_cscval, nnz);
        k_count<<<stride_grid(nnz), kBlock, 0, s>>>(keys.p, nnz, d_nnzpercol);
        FL_LAUNCH_CHECK();
    }
    exclusive_scan_i32(d_nnzpercol, d_colptr, n, s);
    FL_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------------ column sort
void colsort_device(int n, const int *d_counts, int *d_perm, int *d_sortnnz, int max_key, hipStream_t s) {
    if (n <= 0) return;
    TmpBuf<uint32_t> keys(n, s);
    k_copy_keys<<<stride_grid(n), kBlock, 0, s>>>(d_counts, keys.p, n);
    k_iota<<<stride_grid(n), kBlock, 0, s>>>(d_perm, n);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(keys.p, d_perm, n, bits_for((uint32_t)max_key), s);
    k_keys_to_int<<<stride_grid(n), kBlock, 0, s>>>(keys.p, d_sortnnz, n);
    FL_LAUNCH_CHECK();
}

// Concatenate the sorted columns (Pre4SortIdxVal, ColSort_GPU.h:21-43): one lane per column
// for short columns; columns at or beyond `long_from` (lengths are ascending) get a wave each.
__global__ void k_sorted_copy_short(int n, int long_from, const int *__restrict__ perm, const int *__restrict__ P,
                                    const int *__restrict__ colptr, const int *__restrict__ rowidx,
                                    const double *__restrict__ val, int *__restrict__ tmp_row,
                                    double *__restrict__ tmp_val) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < long_from; i += (long long)gridDim.x * blockDim.x) {
        const int c = perm[i], src = colptr[c], len = colptr[c + 1] - src, dst = P[i];
        for (int d = 0; d < len; ++d) {
            tmp_row[dst + d] = rowidx[src + d];
            tmp_val[dst + d] = val[src + d];
        }
    }
}

__global__ void k_sorted_copy_long(int n, int long_from, const int *__restrict__ perm, const int *__restrict__ P,
                                   const int *__restrict__ colptr, const int *__restrict__ rowidx,
                                   const double *__restrict__ val, int *__restrict__ tmp_row,
                                   double *__restrict__ tmp_val) {
    const int lane = threadIdx.x & 63;
    const long long wave0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long i = long_from + wave0; i < n; i += nwaves) {
        const int c = perm[i], src = colptr[c], len = colptr[c + 1] - src, dst = P[i];
        for (int d = lane; d < len; d += 64) {
            tmp_row[dst + d] = rowidx[src + d];
            tmp_val[dst + d] = val[src + d];
        }
    }
}

__global__ void k_lower_bound(const int *__restrict__ sorted, int n, int key, int *__restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sorted[mid] < key) lo = mid + 1; else hi = mid;
        }
        *out = lo;
    }
}

void sorted_copy(int n, const int *d_perm, const int *d_sortnnz, const int *d_P, const int *d_colptr,
                 const int *d_rowidx, const double *d_val, int *d_tmp_row, double *d_tmp_val, hipStream_t s) {
    if (n <= 0) return;
    TmpBuf<int> lf(1, s);
    k_lower_bound<<<1, 64, 0, s>>>(d_sortnnz, n, 64, lf.p);
    FL_LAUNCH_CHECK();
    int long_from = n;
    FL_HIP(hipMemcpyAsync(&long_from, lf.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (long_from > 0) {
        k_sorted_copy_short<<<stride_grid(long_from), kBlock, 0, s>>>(n, long_from, d_perm, d_P, d_colptr, d_rowidx,
                                                                       d_val, d_tmp_row, d_tmp_val);
        FL_LAUNCH_CHECK();
    }
    if (long_from < n) {
        const long long waves = n - long_from;
        const unsigned g = std::max(1u, std::min<unsigned>(grid_for(waves * 64), 16384u));
        k_sorted_copy_long<<<g, kBlock, 0, s>>>(n, long_from, d_perm, d_P, d_colptr, d_rowidx, d_val, d_tmp_row,
                                                d_tmp_val);
        FL_LAUNCH_CHECK();
    }
}

// --------------------------------------------------------------------- level table
__global__ void k_level_flags(const int *__restrict__ s, int n, int *__restrict__ flags) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        flags[i] = (i == 0) ? (s[0] != 0) : (s[i] != s[i - 1]);
}

__global__ void k_level_fill(const int *__restrict__ s, int n, int W, const int *__restrict__ lidx,
                             int *__restrict__ lvl_col, int *__restrict__ lvl_base, int *__restrict__ lvl_tall,
                             int *__restrict__ lvl_tiles, int *__restrict__ lvl_nnz) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int k = lidx[i];
        if (lidx[i + 1] == k) continue;  // not a level start
        const int base = (i == 0) ? 0 : s[i - 1];
        const int tall = s[i] - base;
        const long long width = n - i;
        lvl_col[k] = (int)i;
        lvl_base[k] = base;
        lvl_tall[k] = tall;
        lvl_tiles[k] = (int)(((width + W - 1) / W) * tall);
        lvl_nnz[k] = (int)(width * tall);
    }
}

// Level of tile t: the k with tileoff[k] <= t < tileoff[k+1].
__device__ __forceinline__ int find_level(const int *__restrict__ tileoff, int L, int t) {
    int lo = 0, hi = L - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tileoff[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_tiles(int T, int L, int n, int nnz, int W, const int *__restrict__ lvl_col,
                        const int *__restrict__ lvl_base, const int *__restrict__ lvl_tall,
                        const int *__restrict__ tileoff, const int *__restrict__ nnzoff, int *__restrict__ tile_ptr,
                        int *__restrict__ tile_len, int *__restrict__ tile_colidx, int *__restrict__ tile_depth) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (long long)gridDim.x * blockDim.x) {
        const int k = find_level(tileoff, L, (int)t);
        const int local = (int)t - tileoff[k];
        const int tall = lvl_tall[k], col = lvl_col[k];
        const int j = local / tall, e = local - j * tall;
        const int width = n - col;
        const int chunks = (width + W - 1) / W;
        const int len = (j == chunks - 1) ? width - (chunks - 1) * W : W;
        tile_len[t] = len;
        tile_colidx[t] = col + W * j;
        tile_depth[t] = lvl_base[k] + e;
        tile_ptr[t] = nnzoff[k] + j * tall * W + e * len;
        if (t == T - 1) tile_ptr[T] = nnz;
    }
}

// Tile data layout (dataTrans, formatTransform_GPU.h:152-171): one thread per (tile, lane).
__global__ void k_tile_data(int T, int W, const int *__restrict__ tile_ptr, const int *__restrict__ tile_len,
                            const int *__restrict__ tile_colidx, const int *__restrict__ tile_depth, ElemSource src,
                            int *__restrict__ sortrowidx, double *__restrict__ sortval) {
    const long long total = (long long)T * W;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (long long)gridDim.x * blockDim.x) {
        const long long t = g / W;
        const int l = (int)(g - t * W);
        if (l >= tile_len[t]) continue;
        const int sc = tile_colidx[t] + l, d = tile_depth[t];
        const long long at = src.P ? (long long)src.P[sc] + d : (long long)src.colptr[src.perm[sc]] + d;
        const int dst = tile_ptr[t] + l;
        sortrowidx[dst] = src.rowidx[at];
        sortval[dst] = src.val[at];
    }
}

// Segments + tile format (Pre4CountRow .. Pre4tileclass, formatTransform_GPU.h:188-285):
// one wave holds 64/W tiles; heads from a neighbour compare; run lengths from the head mask.
template <int W>
__global__ __launch_bounds__(kBlock) void k_classify(int T, const int *__restrict__ tile_ptr,
                                                     const int *__restrict__ tile_len,
                                                     const int *__restrict__ sortrowidx,
                                                     int *__restrict__ sortrowindex, int *__restrict__ seg_count,
                                                     int *__restrict__ tile_format) {
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int sub = lane / W, l = lane % W;
    const long long t = wave * (64 / W) + sub;
    const bool tv = t < T;
    const int len = tv ? tile_len[t] : 0;
    const int ptr = tv ? tile_ptr[t] : 0;
    const bool act = l < len;
    const int row = act ? sortrowidx[ptr + l] : -1;
    const int prev = __shfl_up(row, 1);
    const bool head = act && (l == 0 || row != prev);
    const uint64_t hm = ballot64(head);
    const uint64_t tmask = (W == 64) ? hm : ((hm >> (sub * 32)) & 0xffffffffull);
    if (act) {
        int run = 0;
        if (head) {
            const uint64_t above = (l + 1 >= 64) ? 0ull : (tmask >> (l + 1));
            const int nh = above ? (l + 1 + __ffsll((long long)above) - 1) : len;
            run = min(nh, len) - l;
        }
        sortrowindex[ptr + l] = run;
    }
    if (tv && l == 0) {
        const int cnt = __popcll(tmask);
        seg_count[t] = cnt;
        tile_format[t] = (len == W) ? (cnt == 1 ? 1 : 2) : 0;
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_segoff(int T, const int *__restrict__ tile_ptr,
                                                   const int *__restrict__ tile_len,
                                                   const int *__restrict__ sortrowindex,
                                                   const int *__restrict__ countrow, int *__restrict__ segoff) {
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int sub = lane / W, l = lane % W;
    const long long t = wave * (64 / W) + sub;
    const bool tv = t < T;
    const int len = tv ? tile_len[t] : 0;
    const int ptr = tv ? tile_ptr[t] : 0;
    const bool act = l < len;
    const int run = act ? sortrowindex[ptr + l] : 0;
    const uint64_t hm = ballot64(run != 0);
    const uint64_t tmask = (W == 64) ? hm : ((hm >> (sub * 32)) & 0xffffffffull);
    if (run != 0) {
        const uint64_t below = (l == 0) ? 0ull : (tmask & ((1ull << l) - 1ull));
        segoff[countrow[t] + __popcll(below)] = run;
    }
}

// Level of the first tile of each group of 64 tiles (for the FAST kernel).
__global__ void k_group_level(int T, int L, const int *__restrict__ tileoff, int *__restrict__ group_level) {
    const long long groups = ((long long)T + 63) / 64;
    for (long long gidx = (long long)blockIdx.x * blockDim.x + threadIdx.x; gidx < groups; gidx += (long long)gridDim.x * blockDim.x)
        group_level[gidx] = find_level(tileoff, L, (int)(gidx * 64));
}

static int read_int(const int *d, hipStream_t s) {
    int h = 0;
    FL_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return h;
}

void transform_device(fl_matrix *A, const int *d_sortnnz, const ElemSource &src, hipStream_t s, float *t_tran_ms,
                      float *t_clas_ms) {
    const int n = A->n, nnz = A->nnz, W = A->W;
    hipEvent_t e0, e1, e2;
    FL_HIP(hipEventCreate(&e0));
    FL_HIP(hipEventCreate(&e1));
    FL_HIP(hipEventCreate(&e2));
    FL_HIP(hipEventRecord(e0, s));

    // levels
    TmpBuf<int> flags(std::max(n, 1), s), lidx((size_t)n + 1, s);
    if (n > 0) {
        k_level_flags<<<stride_grid(n), kBlock, 0, s>>>(d_sortnnz, n, flags.p);
        FL_LAUNCH_CHECK();
    }
    exclusive_scan_i32(flags.p, lidx.p, n, s);
    const int L = read_int(lidx.p + n, s);
    A->L = L;
    A->lvl_col.alloc(std::max(L, 1));
    A->lvl_base.alloc(std::max(L, 1));
    A->lvl_tall.alloc(std::max(L, 1));
    A->lvl_tileoff.alloc((size_t)L + 1);
    A->lvl_nnzoff.alloc((size_t)L + 1);
    TmpBuf<int> lvl_tiles(std::max(L, 1), s), lvl_nnz(std::max(L, 1), s);
    if (n > 0) {
        k_level_fill<<<stride_grid(n), kBlock, 0, s>>>(d_sortnnz, n, W, lidx.p, A->lvl_col.p, A->lvl_base.p,
                                                       A->lvl_tall.p, lvl_tiles.p, lvl_nnz.p);
        FL_LAUNCH_CHECK();
    }
    exclusive_scan_i32(lvl_tiles.p, A->lvl_tileoff.p, L, s);
    exclusive_scan_i32(lvl_nnz.p, A->lvl_nnzoff.p, L, s);
    const int T = read_int(A->lvl_tileoff.p + L, s);
    const int total = read_int(A->lvl_nnzoff.p + L, s);
    if (total != nnz) {
        set_error("formatTransform: level table covers %d elements, expected nnz=%d", total, nnz);
        throw HipError{hipErrorInvalidValue};
    }
    A->tilenum = T;

    // tiles + data
    A->tile_ptr.alloc((size_t)T + 1);
    A->tile_len.alloc(std::max(T, 1));
    A->tile_colidx.alloc(std::max(T, 1));
    A->tile_format.alloc(std::max(T, 1));
    TmpBuf<int> tile_depth(std::max(T, 1), s);
    if (!src.tiled) {
        A->sortrowidx.alloc(std::max(nnz, 1));
        A->val64.alloc(std::max(nnz, 1));
    }
    A->sortrowindex.alloc(std::max(nnz, 1));
    if (T > 0) {
        k_tiles<<<stride_grid(T), kBlock, 0, s>>>(T, L, n, nnz, W, A->lvl_col.p, A->lvl_base.p, A->lvl_tall.p,
                                                  A->lvl_tileoff.p, A->lvl_nnzoff.p, A->tile_ptr.p, A->tile_len.p,
                                                  A->tile_colidx.p, tile_depth.p);
        FL_LAUNCH_CHECK();
        if (!src.tiled) {
            k_tile_data<<<stride_grid((long long)T * W), kBlock, 0, s>>>(T, W, A->tile_ptr.p, A->tile_len.p,
                                                                         A->tile_colidx.p, tile_depth.p, src,
                                                                         A->sortrowidx.p, A->val64.p);
            FL_LAUNCH_CHECK();
        }
    } else {
        FL_HIP(hipMemsetAsync(A->tile_ptr.p, 0, sizeof(int), s));
    }
    FL_HIP(hipEventRecord(e1, s));

    // classification
    TmpBuf<int> seg_count(std::max(T, 1), s);
    A->countrow.alloc((size_t)T + 1);
    if (T > 0) {
        const long long waves = ((long long)T + (64 / W) - 1) / (64 / W);
        const unsigned g = grid_for(waves * 64);
        if (W == 32)
            k_classify<32><<<g, kBlock, 0, s>>>(T, A->tile_ptr.p, A->tile_len.p, A->sortrowidx.p, A->sortrowindex.p,
                                                seg_count.p, A->tile_format.p);
        else
            k_classify<64><<<g, kBlock, 0, s>>>(T, A->tile_ptr.p, A->tile_len.p, A->sortrowidx.p, A->sortrowindex.p,
                                                seg_count.p, A->tile_format.p);
        FL_LAUNCH_CHECK();
    }
    exclusive_scan_i32(seg_count.p, A->countrow.p, T, s);
    A->segsum = read_int(A->countrow.p + T, s);
    A->segmentoffset.alloc((size_t)A->segsum + 1);
    FL_HIP(hipMemsetAsync(A->segmentoffset.p + A->segsum, 0, sizeof(int), s));
    if (T > 0) {
        const long long waves = ((long long)T + (64 / W) - 1) / (64 / W);
        const unsigned g = grid_for(waves * 64);
        if (W == 32)
            k_segoff<32><<<g, kBlock, 0, s>>>(T, A->tile_ptr.p, A->tile_len.p, A->sortrowindex.p, A->countrow.p,
                                              A->segmentoffset.p);
        else
            k_segoff<64><<<g, kBlock, 0, s>>>(T, A->tile_ptr.p, A->tile_len.p, A->sortrowindex.p, A->countrow.p,
                                              A->segmentoffset.p);
        FL_LAUNCH_CHECK();
    }
    FL_HIP(hipEventRecord(e2, s));

    // group -> level hints for the FAST kernel
    const long long groups = ((long long)T + 63) / 64;
    A->group_level.alloc((size_t)std::max<long long>(groups, 1));
    if (groups > 0 && L > 0) {
        k_group_level<<<stride_grid(groups), kBlock, 0, s>>>(T, L, A->lvl_tileoff.p, A->group_level.p);
        FL_LAUNCH_CHECK();
    }
    FL_HIP(hipEventSynchronize(e2));
    FL_HIP(hipStreamSynchronize(s));
    float a = 0.f, b = 0.f;
    FL_HIP(hipEventElapsedTime(&a, e0, e1));
    FL_HIP(hipEventElapsedTime(&b, e1, e2));
    if (t_tran_ms) *t_tran_ms = a;
    if (t_clas_ms) *t_clas_ms = b;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
}

}  // namespace fl
