// Batched FastLoad schedules for the 64-wide layout: y partials accumulate in LDS in FP64
// (ds_add_f64 is several times faster than the FP32 LDS atomic on gfx950), so scattered
// global atomics -- one per segment in the reference kernel (src/FastLoad_GPU.h:186,224,267)
// -- are replaced by LDS work.
//
// The tiles are re-ordered (stable) by the 256-column bucket of their first sorted column, so
// tiles of different levels that read the same part of x run in the same batch. A batch is `bt`
// consecutive re-ordered tiles (<= bt*64 entries), handled by one workgroup whose waves take 16
// tiles each (bt*4 threads; WIN caps it at 512), reading 64 consecutive sorted x values per
// tile as in FastLoad, from a window of x that stays in L1/L2 instead of being re-streamed per
// level.
//
// WIN (rows with locality: stencils, bands, blocks): the distinct rows of a batch are
//   compacted to LDS slots at preprocessing time; entries store a 2-byte slot instead of a
//   4-byte row; products are summed per slot with ds_add_f64 and each slot is flushed with
//   one global atomic (consecutive rows -> coalesced atomics).
// PB (rows without locality, e.g. uniform random rows): two-phase propagation blocking.
//   Phase 1 stages the batch's products in LDS in band-major order (bands of R rows) and
//   copies each (batch, band) run contiguously into a band-major product array. Phase 2 gives
//   each band one workgroup that sums its products per row in LDS (ds_add_f64) and updates y
//   once per row with a coalesced read-modify-write -- no global atomics at all.
#include <algorithm>
#include <type_traits>

#include <atomic>

#include "fl_matrix.h"

namespace fl {

void exclusive_scan_i32(const int *in, int *out, int count, hipStream_t s);

static unsigned sg(long long count) { return std::max(1u, std::min<unsigned>(grid_for(count), 65536u)); }

#define GRID_STRIDE(i, count) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += (long long)gridDim.x * blockDim.x)

// ------------------------------------------------------------------------ preprocessing
// re-ordered tile i (original tile order[i]) -> its entries belong to batch i / bt
__global__ void k_batch_of(int T, int bt, int n, const int *__restrict__ ord_ptr, const int *__restrict__ ord_cx,
                           int *__restrict__ batch_of) {
    GRID_STRIDE(g, (long long)T * 64) {
        const long long i = g >> 6;
        const int l = (int)(g & 63);
        if (l < min(64, n - ord_cx[i])) batch_of[ord_ptr[i] + l] = (int)(i / bt);
    }
}

__global__ void k_tile_bucket(const int *__restrict__ tile_colidx, int T, uint32_t *__restrict__ key) {
    GRID_STRIDE(t, T) key[t] = (uint32_t)(tile_colidx[t] >> 8);
}

// row window (64 rows) of a tile's first entry
__global__ void k_tile_rowkey(const int *__restrict__ tile_ptr, const int *__restrict__ row, int T,
                              uint32_t *__restrict__ key) {
    GRID_STRIDE(t, T) key[t] = (uint32_t)(row[tile_ptr[t]] >> 6);
}

// slots whose row other batches touch too are flushed with an atomic; the others own their row
__global__ void k_row_batches(const int *__restrict__ win_rows, int D, int *__restrict__ cnt) {
    GRID_STRIDE(i, D) atomicAdd(&cnt[win_rows[i]], 1);
}
__global__ void k_mark_shared(int *__restrict__ win_rows, int D, const int *__restrict__ cnt) {
    GRID_STRIDE(i, D) {
        const int r = win_rows[i];
        if (cnt[r] > 1) win_rows[i] = (int)((uint32_t)r | 0x80000000u);
    }
}

__global__ void k_tile_reorder(const int *__restrict__ order, const int *__restrict__ tile_ptr,
                               const int *__restrict__ tile_colidx, int T, int *__restrict__ ord_ptr,
                               int *__restrict__ ord_cx) {
    GRID_STRIDE(i, T) {
        const int t = order[i];
        ord_ptr[i] = tile_ptr[t];
        ord_cx[i] = tile_colidx[t];
    }
}

// entry count of each batch (sum of its re-ordered tiles' lengths)
__global__ void k_batch_sizes(int T, int bt, int nb, int n, const int *__restrict__ ord_cx, int *__restrict__ sizes) {
    GRID_STRIDE(b, nb) {
        int s = 0;
        const int i1 = min((int)(b + 1) * bt, T);
        for (int i = (int)b * bt; i < i1; ++i) s += min(64, n - ord_cx[i]);
        sizes[b] = s;
    }
}

__global__ void k_iota2(int *__restrict__ v, int count) {
    GRID_STRIDE(i, count) v[i] = (int)i;
}

__global__ void k_key_gather(const int *__restrict__ src, const int *__restrict__ order, uint32_t *__restrict__ key,
                             int count) {
    GRID_STRIDE(i, count) key[i] = (uint32_t)src[order[i]];
}

__global__ void k_key_copy(const int *__restrict__ src, uint32_t *__restrict__ key, int count, int div) {
    GRID_STRIDE(i, count) key[i] = (uint32_t)(src[i] / div);
}

// Batch b runs on XCD b % 8 (workgroups are dealt to the 8 XCDs round-robin). Ordering a band's
// runs by (b % 8, b / 8) puts the runs one XCD writes next to each other, so a cache line
// shared by two neighbouring runs is usually completed in one L2 instead of two.
__global__ void k_key_batch_xcd(const int *__restrict__ batch_of, int per_xcd, uint32_t *__restrict__ key,
                                int count) {
    GRID_STRIDE(i, count) {
        const int b = batch_of[i];
        key[i] = (uint32_t)((b & 7) * per_xcd + (b >> 3));
    }
}


// Chunk-major band key of the entry at sorted position i: chunk * nbands + band, where the
// chunk is the number of chunk starts cb[1..C-1] at or below the entry's batch.
__global__ void k_key_chunk_band(const int *__restrict__ batch_of, const int *__restrict__ row,
                                 const int *__restrict__ order, int R, int nbands, const int *__restrict__ cb,
                                 int C, uint32_t *__restrict__ key, int count) {
    GRID_STRIDE(i, count) {
        const int g = order[i];
        const int b = batch_of[g];
        int c = 0;
        for (int k = 1; k < C; ++k) c += (b >= cb[k]);
        key[i] = (uint32_t)c * (uint32_t)nbands + (uint32_t)(row[g] / R);
    }
}

__global__ void k_key_batch_band(const int *__restrict__ batch_of, const int *__restrict__ row, int R, int nbands,
                                 uint32_t *__restrict__ key, int count) {
    GRID_STRIDE(i, count) key[i] = (uint32_t)batch_of[i] * (uint32_t)nbands + (uint32_t)(row[i] / R);
}

// WIN: heads of (batch, row) runs in (batch, row, g) order
__global__ void k_win_heads(const uint32_t *__restrict__ kb, const int *__restrict__ order, const int *__restrict__ row,
                            int *__restrict__ head, int count) {
    GRID_STRIDE(i, count) {
        int h = 1;
        if (i > 0) h = (kb[i] != kb[i - 1]) || (row[order[i]] != row[order[i - 1]]);
        head[i] = h;
    }
}

__global__ void k_win_off(const uint32_t *__restrict__ kb, const int *__restrict__ d, int *__restrict__ win_off,
                          int count) {
    GRID_STRIDE(i, count) if (i == 0 || kb[i] != kb[i - 1]) win_off[kb[i]] = d[i];
}

__global__ void k_win_slot(const uint32_t *__restrict__ kb, const int *__restrict__ order, const int *__restrict__ row,
                           const int *__restrict__ head, const int *__restrict__ d, const int *__restrict__ win_off,
                           uint16_t *__restrict__ slot, int *__restrict__ win_rows, int count) {
    GRID_STRIDE(i, count) {
        const int g = order[i];
        const int di = d[i] + head[i] - 1;  // index of this entry's (batch, row) run
        slot[g] = (uint16_t)(di - win_off[kb[i]]);
        if (head[i]) win_rows[d[i]] = row[g];
    }
}

// PB: dest[g] = product-array position (band-major position minus its chunk's offset);
// rowlocal[pos] = row - band*R at the band-major position
__global__ void k_pb_dest(const int *__restrict__ orderA, const int *__restrict__ row, int R,
                          const uint32_t *__restrict__ key, int nbands, const int *__restrict__ chunk_off,
                          int *__restrict__ dest, uint16_t *__restrict__ rowlocal, int count) {
    GRID_STRIDE(i, count) {
        const int g = orderA[i];
        dest[g] = (int)i - chunk_off[key[i] / (uint32_t)nbands];
        rowlocal[i] = (uint16_t)(row[g] % R);
    }
}

// chunk_off[c] = first band-major position of chunk c, rounded down to a multiple of 4 (keeps
// the 16-byte alignment of phase 2's vector loads in the reused product array)
__global__ void k_chunk_off(const int *__restrict__ band_start, int nbands, int C, int *__restrict__ chunk_off) {
    GRID_STRIDE(c, C) chunk_off[c] = band_start[(size_t)c * nbands] & ~3;
}

// start[b] = first position of key b in the sorted keys (lower bound); start[nb] = count.
__global__ void k_key_bounds(const uint32_t *__restrict__ k, int count, int nb, int *__restrict__ start) {
    GRID_STRIDE(b, (long long)nb + 1) {
        int lo = 0, hi = count;
        while (lo < hi) {
            const int mid = (int)(((long long)lo + hi) >> 1);
            if (k[mid] < (uint32_t)b) lo = mid + 1; else hi = mid;
        }
        start[b] = lo;
    }
}

__global__ void k_pb_heads(const uint32_t *__restrict__ kb, int *__restrict__ head, int count) {
    GRID_STRIDE(i, count) head[i] = (i == 0) || (kb[i] != kb[i - 1]);
}

__global__ void k_pb_runs(const uint32_t *__restrict__ kb, const int *__restrict__ orderB, const int *__restrict__ head,
                          const int *__restrict__ runidx, const int *__restrict__ dest,
                          const int *__restrict__ batch_start, int nbands, uint16_t *__restrict__ lslot,
                          int *__restrict__ run_lstart, int *__restrict__ run_dest, int *__restrict__ run_pos,
                          int *__restrict__ run_off, int count) {
    GRID_STRIDE(i, count) {
        const int batch = (int)(kb[i] / (uint32_t)nbands);
        const int bstart = batch_start[batch];
        const int g = orderB[i];
        lslot[g] = (uint16_t)(i - bstart);
        if (head[i]) {
            const int r = runidx[i];
            run_lstart[r] = (int)i - bstart;
            run_dest[r] = dest[g];
            run_pos[r] = (int)i;
            if (i == 0 || kb[i] / (uint32_t)nbands != kb[i - 1] / (uint32_t)nbands) run_off[batch] = r;
        }
    }
}

__global__ void k_run_len(const int *__restrict__ run_pos, int nruns, int total, int *__restrict__ run_len) {
    GRID_STRIDE(r, nruns) run_len[r] = ((r + 1 < nruns) ? run_pos[r + 1] : total) - run_pos[r];
}

__global__ void k_adjdiff(const int *__restrict__ in, int count, int *__restrict__ out) {
    GRID_STRIDE(i, count) out[i] = in[i + 1] - in[i];
}

// Every entry's destination must be its run's destination plus its offset inside the run (the
// phase-1 copy moves whole runs).
__global__ void k_check_runs(const int *__restrict__ orderB, const int *__restrict__ head,
                             const int *__restrict__ runidx, const int *__restrict__ run_pos,
                             const int *__restrict__ run_dest, const int *__restrict__ dest, int count,
                             int *__restrict__ flag) {
    GRID_STRIDE(i, count) {
        const int r = runidx[i] + head[i] - 1;
        if (dest[orderB[i]] != run_dest[r] + ((int)i - run_pos[r])) atomicOr(flag, 16);
    }
}

// Validators: every index the kernels will use must be inside the buffers they assume.
__global__ void k_check_win(const int *__restrict__ batch_of, const uint16_t *__restrict__ slot,
                            const int *__restrict__ win_off, int cap, int nnz, int m,
                            const int *__restrict__ win_rows, int *__restrict__ flag) {
    GRID_STRIDE(g, nnz) {
        const int b = batch_of[g];
        const int ns = win_off[b + 1] - win_off[b];
        if (ns <= 0 || ns > cap || (int)slot[g] >= ns) {
            atomicOr(flag, 1);
            continue;
        }
        const int r = win_rows[win_off[b] + slot[g]] & 0x7fffffff;  // high bit: row shared with other batches
        if (r >= m) atomicOr(flag, 2);
    }
}

__global__ void k_check_pb(const int *__restrict__ batch_of, const uint16_t *__restrict__ lslot,
                           const int *__restrict__ batch_start, int cap, int nnz, int ring,
                           const int *__restrict__ run_off, const int *__restrict__ run_lstart,
                           const int *__restrict__ run_dest, const int *__restrict__ run_len, int nb,
                           int *__restrict__ flag) {
    GRID_STRIDE(g, nnz) {
        const int b = batch_of[g];
        const int bsize = batch_start[b + 1] - batch_start[b];
        if (bsize > cap || (int)lslot[g] >= bsize) atomicOr(flag, 4);
    }
    GRID_STRIDE(b, nb) {
        const int bsize = batch_start[b + 1] - batch_start[b];
        for (int r = run_off[b]; r < run_off[b + 1]; ++r)
            if (run_len[r] <= 0 || run_lstart[r] < 0 || run_lstart[r] + run_len[r] > bsize || run_dest[r] < 0 ||
                run_dest[r] + run_len[r] > ring)
                atomicOr(flag, 8);
    }
}

static int read1(const int *d, hipStream_t s) {
    int h = 0;
    FL_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return h;
}

static void write1(int *d, int v, hipStream_t s) {
    FL_HIP(hipMemcpyAsync(d, &v, sizeof(int), hipMemcpyHostToDevice, s));
    FL_HIP(hipStreamSynchronize(s));
}

static void ensure_batches(fl_matrix *A, int bt, int kind, hipStream_t s) {
    const int T = A->tilenum;
    if (!A->ord_ptr.p || A->ord_kind != kind) {
        // stable sort of the tiles by 256-column bucket of their first sorted column (kind 0) or
        // by the 64-row window of their first entry (kind 1)
        A->ord_ptr.alloc(std::max(T, 1));
        A->ord_cx.alloc(std::max(T, 1));
        A->ord_kind = kind;
        A->bt = 0;  // batches follow the tile order: rebuild them
        if (T > 0) {
            TmpBuf<uint32_t> key(T, s);
            TmpBuf<int> order(T, s);
            if (kind == 1) k_tile_rowkey<<<sg(T), kBlock, 0, s>>>(A->tile_ptr.p, A->sortrowidx.p, T, key.p);
            else k_tile_bucket<<<sg(T), kBlock, 0, s>>>(A->tile_colidx.p, T, key.p);
            k_iota2<<<sg(T), kBlock, 0, s>>>(order.p, T);
            FL_LAUNCH_CHECK();
            const int top = kind == 1 ? (A->m - 1) >> 6 : (A->n - 1) >> 8;
            radix_sort_pairs(key.p, order.p, T, bits_for((uint32_t)std::max(top, 0)), s);
            k_tile_reorder<<<sg(T), kBlock, 0, s>>>(order.p, A->tile_ptr.p, A->tile_colidx.p, T, A->ord_ptr.p,
                                                    A->ord_cx.p);
            FL_LAUNCH_CHECK();
            FL_HIP(hipStreamSynchronize(s));
        }
    }
    if (A->bt == bt && A->batch_of.p) return;
    A->bt = bt;
    A->nbatches = (T + bt - 1) / bt;
    A->win_ready = A->pb_ready = false;
    A->win_pack.release();
    A->pb_pack.release();
    A->batch_of.alloc(std::max(A->nnz, 1));
    A->batch_start.alloc((size_t)A->nbatches + 1);
    TmpBuf<int> sizes(std::max(A->nbatches, 1), s);
    if (T > 0) {
        k_batch_of<<<sg((long long)T * 64), kBlock, 0, s>>>(T, bt, A->n, A->ord_ptr.p, A->ord_cx.p, A->batch_of.p);
        k_batch_sizes<<<sg(A->nbatches), kBlock, 0, s>>>(T, bt, A->nbatches, A->n, A->ord_cx.p, sizes.p);
        FL_LAUNCH_CHECK();
    }
    exclusive_scan_i32(sizes.p, A->batch_start.p, A->nbatches, s);
    FL_HIP(hipStreamSynchronize(s));
}

void pack_schedules(fl_matrix *A, hipStream_t s);

void win_prepare(fl_matrix *A, int bt, bool rows, hipStream_t s) {
    ensure_batches(A, bt, rows ? 1 : 0, s);
    if (A->win_ready) return;
    A->win_pack.release();
    const int nnz = A->nnz, nb = A->nbatches;
    A->win_slot.alloc(std::max(nnz, 1));
    A->win_off.alloc((size_t)nb + 1);
    if (nnz == 0) {
        A->win_rows.alloc(1);
        FL_HIP(hipMemsetAsync(A->win_off.p, 0, sizeof(int) * ((size_t)nb + 1), s));
        A->win_maxslots = 0;
        A->win_ready = true;
        return;
    }
    TmpBuf<uint32_t> key(nnz, s);
    TmpBuf<int> order(nnz, s), head(nnz, s), d((size_t)nnz + 1, s);
    // (row, g) order, then stable by batch -> (batch, row, g)
    k_key_copy<<<sg(nnz), kBlock, 0, s>>>(A->sortrowidx.p, key.p, nnz, 1);
    k_iota2<<<sg(nnz), kBlock, 0, s>>>(order.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(A->m - 1, 0)), s);
    k_key_gather<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, order.p, key.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(nb - 1, 0)), s);
    k_win_heads<<<sg(nnz), kBlock, 0, s>>>(key.p, order.p, A->sortrowidx.p, head.p, nnz);
    FL_LAUNCH_CHECK();
    exclusive_scan_i32(head.p, d.p, nnz, s);
    const int D = read1(d.p + nnz, s);
    A->win_rows.alloc(std::max(D, 1));
    k_win_off<<<sg(nnz), kBlock, 0, s>>>(key.p, d.p, A->win_off.p, nnz);
    FL_LAUNCH_CHECK();
    write1(A->win_off.p + nb, D, s);
    k_win_slot<<<sg(nnz), kBlock, 0, s>>>(key.p, order.p, A->sortrowidx.p, head.p, d.p, A->win_off.p, A->win_slot.p,
                                          A->win_rows.p, nnz);
    FL_LAUNCH_CHECK();
    TmpBuf<int> flag(1, s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    k_check_win<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->win_slot.p, A->win_off.p, bt * 64, nnz, A->m,
                                           A->win_rows.p, flag.p);
    FL_LAUNCH_CHECK();
    if (read1(flag.p, s)) {
        set_error("WIN schedule: preprocessing produced out-of-range slots");
        throw HipError{hipErrorInvalidValue};
    }
    {
        TmpBuf<int> cnt(std::max(A->m, 1), s);
        FL_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * std::max(A->m, 1), s));
        k_row_batches<<<sg(D), kBlock, 0, s>>>(A->win_rows.p, D, cnt.p);
        k_mark_shared<<<sg(D), kBlock, 0, s>>>(A->win_rows.p, D, cnt.p);
        FL_LAUNCH_CHECK();
        FL_HIP(hipStreamSynchronize(s));
    }
    TmpBuf<int> per_batch(nb, s);
    k_adjdiff<<<sg(nb), kBlock, 0, s>>>(A->win_off.p, nb, per_batch.p);
    FL_LAUNCH_CHECK();
    A->win_maxslots = device_max_i32(per_batch.p, nb, s);
    A->win_ready = true;
    if (A->pack_wanted) pack_schedules(A, s);
}

// Chunk boundaries: C consecutive batch ranges of about nnz/C entries each, cut at multiples of
// 8 batches so that batch b still runs on XCD b % 8 when a chunk's phase 1 starts at its first
// batch. Empty ranges are dropped.
static std::vector<int> pb_chunk_batches(const fl_matrix *A, int C, hipStream_t s) {
    const int nb = A->nbatches;
    std::vector<int> cb{0};
    if (C > 1) {
        std::vector<int> bs((size_t)nb + 1);
        FL_HIP(hipMemcpyAsync(bs.data(), A->batch_start.p, sizeof(int) * ((size_t)nb + 1), hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
        for (int c = 1; c < C; ++c) {
            const long long target = (long long)bs[nb] * c / C;
            int b = (int)(std::lower_bound(bs.begin(), bs.end(), (int)target) - bs.begin());
            b &= ~7;
            if (b > cb.back() && b < nb) cb.push_back(b);
        }
    }
    cb.push_back(nb);
    return cb;
}

void pb_prepare(fl_matrix *A, int bt, int R, hipStream_t s) {
    ensure_batches(A, bt, 0, s);
    if (A->pb_ready && A->pb_R == R) return;
    A->pb_ready = false;
    A->pb_pack.release();
    const int nnz = A->nnz, nb = A->nbatches, m = A->m;
    const int nbands = std::max((m + R - 1) / R, 1);
    if ((long long)nb * nbands > 0x7fffffffLL) {
        set_error("PB schedule: %d batches x %d bands exceed 31-bit keys", nb, nbands);
        throw HipError{hipErrorInvalidValue};
    }
    A->pb_R = R;
    A->nbands = nbands;
    A->pb_lslot.alloc(std::max(nnz, 1));
    A->pb_rowlocal.alloc(std::max(nnz, 1));
    A->pb_run_off.alloc((size_t)nb + 1);
    if (nnz == 0) {
        A->pb_chunks = 1;
        A->pb_chunk_batch = {0, nb};
        A->pb_chunk_off = {0};
        A->pb_band_start.alloc((size_t)nbands + 1);
        FL_HIP(hipMemsetAsync(A->pb_band_start.p, 0, sizeof(int) * ((size_t)nbands + 1), s));
        FL_HIP(hipMemsetAsync(A->pb_run_off.p, 0, sizeof(int) * ((size_t)nb + 1), s));
        A->nruns = 0;
        A->pb_ring = 0;
        A->pb_ready = true;
        return;
    }
    // Two chunks when the FP32 products (4 B each) exceed half of the 256 MB Infinity Cache and y
    // is small next to them (every chunk after the first re-reads and re-writes y). On the
    // 67 M-entry power-law matrix two chunks are 7 % faster than one; three or more are slower
    // (each chunk adds a pass over y and two launch tails).
    const int want = (nnz >= (1 << 25) && (long long)m * 8 <= nnz) ? 2 : 1;
    const std::vector<int> cb = pb_chunk_batches(A, want, s);
    const int C = (int)cb.size() - 1;
    if ((long long)C * nbands > 0x7fffffffLL) {
        set_error("PB schedule: %d chunks x %d bands exceed 31-bit keys", C, nbands);
        throw HipError{hipErrorInvalidValue};
    }
    A->pb_band_start.alloc((size_t)C * nbands + 1);
    TmpBuf<int> d_cb(cb.size(), s), d_off(C, s);
    FL_HIP(hipMemcpyAsync(d_cb.p, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice, s));
    TmpBuf<uint32_t> key(nnz, s);
    TmpBuf<int> order(nnz, s), dest(nnz, s), head(nnz, s), runidx((size_t)nnz + 1, s);
    // A: chunk-major, then band-major order (chunk, band, batch, g): stable by batch (batches in
    // XCD order), then stable by (chunk, band), so each (batch, band) run occupies one contiguous
    // range of the product array
    const int per_xcd = (nb + 7) / 8;
    k_key_batch_xcd<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, per_xcd, key.p, nnz);
    k_iota2<<<sg(nnz), kBlock, 0, s>>>(order.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)std::max(8 * per_xcd - 1, 0)), s);
    k_key_chunk_band<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->sortrowidx.p, order.p, R, nbands, d_cb.p, C, key.p,
                                                nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)(C * nbands - 1)), s);
    k_key_bounds<<<sg((long long)C * nbands + 1), kBlock, 0, s>>>(key.p, nnz, C * nbands, A->pb_band_start.p);
    k_chunk_off<<<sg(C), kBlock, 0, s>>>(A->pb_band_start.p, nbands, C, d_off.p);
    k_pb_dest<<<sg(nnz), kBlock, 0, s>>>(order.p, A->sortrowidx.p, R, key.p, nbands, d_off.p, dest.p,
                                         A->pb_rowlocal.p, nnz);
    FL_LAUNCH_CHECK();
    std::vector<int> off(C), cstart((size_t)C + 1);
    FL_HIP(hipMemcpyAsync(off.data(), d_off.p, sizeof(int) * C, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    int ring = 0;
    for (int c = 0; c <= C; ++c) cstart[c] = read1(A->pb_band_start.p + (size_t)c * nbands, s);
    for (int c = 0; c < C; ++c) ring = std::max(ring, cstart[c + 1] - off[c]);
    // B: (batch, band, g) order -> LDS staging slots and runs
    k_key_batch_band<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->sortrowidx.p, R, nbands, key.p, nnz);
    k_iota2<<<sg(nnz), kBlock, 0, s>>>(order.p, nnz);
    FL_LAUNCH_CHECK();
    radix_sort_pairs(key.p, order.p, nnz, bits_for((uint32_t)(nb * nbands - 1)), s);
    k_pb_heads<<<sg(nnz), kBlock, 0, s>>>(key.p, head.p, nnz);
    FL_LAUNCH_CHECK();
    exclusive_scan_i32(head.p, runidx.p, nnz, s);
    const int nruns = read1(runidx.p + nnz, s);
    A->nruns = nruns;
    A->pb_run_lstart.alloc(nruns);
    A->pb_run_dest.alloc(nruns);
    A->pb_run_len.alloc(nruns);
    TmpBuf<int> run_pos(nruns, s);
    k_pb_runs<<<sg(nnz), kBlock, 0, s>>>(key.p, order.p, head.p, runidx.p, dest.p, A->batch_start.p, nbands,
                                         A->pb_lslot.p, A->pb_run_lstart.p, A->pb_run_dest.p, run_pos.p,
                                         A->pb_run_off.p, nnz);
    k_run_len<<<sg(nruns), kBlock, 0, s>>>(run_pos.p, nruns, nnz, A->pb_run_len.p);
    FL_LAUNCH_CHECK();
    write1(A->pb_run_off.p + nb, nruns, s);
    TmpBuf<int> flag(1, s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    k_check_pb<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->pb_lslot.p, A->batch_start.p, A->bt * 64, nnz, ring,
                                          A->pb_run_off.p, A->pb_run_lstart.p, A->pb_run_dest.p, A->pb_run_len.p, nb,
                                          flag.p);
    k_check_runs<<<sg(nnz), kBlock, 0, s>>>(order.p, head.p, runidx.p, run_pos.p, A->pb_run_dest.p, dest.p, nnz,
                                            flag.p);
    FL_LAUNCH_CHECK();
    if (read1(flag.p, s) || cstart[C] != nnz || ring > nnz) {
        set_error("PB schedule: preprocessing produced out-of-range runs or slots");
        throw HipError{hipErrorInvalidValue};
    }
    A->pb_chunks = C;
    A->pb_ring = ring;
    A->pb_chunk_batch = cb;
    A->pb_chunk_off = off;
    A->pb_ready = true;
    if (A->pack_wanted) pack_schedules(A, s);
}

// ------------------------------------------------------------- packed value + slot words
// Exponent range of the non-zero values: r[0] = min, r[1] = max unbiased exponent among values
// of FP32-normal magnitude (smaller ones are stored as zero), r[2] != 0 when a value is not
// finite or too large to pack (exponent >= 127).
// The wave reduces its lanes' ranges first and one lane per wave updates r (one atomic per
// thread on three words took 6 ms on 67 M values).
__global__ void k_exp_range(const double *__restrict__ val, int count, int *__restrict__ r) {
    int lo = 0x7fffffff, hi = -0x7fffffff, bad = 0;
    GRID_STRIDE(g, count) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(val[g]);
        const int ed = (int)((b >> 52) & 0x7ff) - 1023;
        if (ed >= 127) bad = 1;
        else if (ed >= -126) {
            lo = min(lo, ed);
            hi = max(hi, ed);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
        bad |= __shfl_xor(bad, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&r[0], lo);
        atomicMax(&r[1], hi);
        if (bad) atomicOr(&r[2], 1);
    }
}

// word = sign | exponent code | mantissa rounded to nearest even (from the FP64 value) | index
__global__ void k_pack(const double *__restrict__ val, const uint16_t *__restrict__ idx, int count, PackFmt f,
                       int emin, uint32_t *__restrict__ out) {
    GRID_STRIDE(g, count) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(val[g]);
        const int ed = (int)((b >> 52) & 0x7ff) - 1023;
        uint32_t word = 0;
        if (ed >= -126) {
            const int sh = 52 - f.m;
            const unsigned long long frac = b & ((1ull << 52) - 1ull);
            unsigned long long q = frac >> sh;
            const unsigned long long rem = frac & ((1ull << sh) - 1ull), half = 1ull << (sh - 1);
            int e = ed;
            if (rem > half || (rem == half && (q & 1ull))) ++q;
            if (q >> f.m) {  // rounding carried into the exponent
                q = 0;
                ++e;
            }
            const uint32_t sign = (uint32_t)(b >> 63), ec = (uint32_t)(e - emin + 1);
            word = (sign << (f.ib + f.m + f.e)) | (ec << (f.ib + f.m)) | ((uint32_t)q << f.ib);
        }
        out[g] = word | (uint32_t)idx[g];
    }
}

// sign at bit 31, then e exponent-code bits, m mantissa bits and ib slot bits (1 + e + m + ib = 32):
// e covers the exponent range (one code more for a rounding carry; code 0 = zero), m takes what
// is left up to 23, and e absorbs any bits beyond that.
static bool pack_format(const fl_matrix *A, int ib, PackFmt &f) {
    const int e_need = bits_for((uint32_t)(A->pack_emax - A->pack_emin + 2));
    const int m = std::min(23, 31 - ib - e_need);
    if (!A->pack_ok || m < 12) return false;
    f.ib = ib;
    f.m = m;
    f.e = 31 - ib - m;
    f.ebase = A->pack_emin + 126;
    f.imask = (1u << ib) - 1u;
    f.fmask = (f.e + m >= 32) ? ~0u : (1u << (f.e + m)) - 1u;
    f.mmask = (1u << m) - 1u;
    f.ebias = (uint32_t)f.ebase << 23;
    return true;
}

// Builds the packed words of every prepared batched schedule that has none yet. A schedule whose
// slot bits leave fewer than 12 mantissa bits (or a matrix with non-finite or huge values) keeps
// no packed words, and FL_VAL_PACKED calls on it are refused.
void pack_schedules(fl_matrix *A, hipStream_t s) {
    const int nnz = A->nnz;
    if (!A->pack_range_known) {
        TmpBuf<int> r(3, s);
        const int init[3] = {0x7fffffff, -0x7fffffff, 0};
        FL_HIP(hipMemcpyAsync(r.p, init, sizeof init, hipMemcpyHostToDevice, s));
        if (nnz) k_exp_range<<<std::min(sg(nnz), 4096u), kBlock, 0, s>>>(A->val64.p, nnz, r.p);
        FL_LAUNCH_CHECK();
        int h[3];
        FL_HIP(hipMemcpyAsync(h, r.p, sizeof h, hipMemcpyDeviceToHost, s));
        FL_HIP(hipStreamSynchronize(s));
        A->pack_ok = h[2] == 0;
        A->pack_emin = h[0] <= h[1] ? h[0] : 0;
        A->pack_emax = h[0] <= h[1] ? h[1] : 0;
        A->pack_range_known = true;
    }
    if (A->win_ready && !A->win_pack.p) {
        PackFmt f;
        if (pack_format(A, std::max(1, bits_for((uint32_t)std::max(A->win_maxslots - 1, 0))), f)) {
            A->win_pack.alloc(std::max(nnz, 1));
            if (nnz) k_pack<<<sg(nnz), kBlock, 0, s>>>(A->val64.p, A->win_slot.p, nnz, f, A->pack_emin, A->win_pack.p);
            FL_LAUNCH_CHECK();
            A->win_pf = f;
        }
    }
    if (A->pb_ready && !A->pb_pack.p) {
        PackFmt f;
        if (pack_format(A, bits_for((uint32_t)(A->bt * 64 - 1)), f)) {
            A->pb_pack.alloc(std::max(nnz, 1));
            if (nnz) k_pack<<<sg(nnz), kBlock, 0, s>>>(A->val64.p, A->pb_lslot.p, nnz, f, A->pack_emin, A->pb_pack.p);
            FL_LAUNCH_CHECK();
            A->pb_pf = f;
        }
    }
    FL_HIP(hipStreamSynchronize(s));
}

// ----------------------------------------------------------- schedules read back from a file
// The arrays come from host memory (a file); every index a kernel will use is checked before
// the schedule is marked ready, on the host where the check is a scan and on the device with
// the validators above, so a damaged file is refused instead of indexing out of bounds.
template <typename T>
static void upload_to(DevBuf<T> &d, const T *h, size_t count, hipStream_t s) {
    d.alloc(std::max<size_t>(count, 1));
    if (count) FL_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

static bool ptr_ok(const int *p, size_t count, int first, int last) {
    if (p[0] != first || p[count - 1] != last) return false;
    for (size_t i = 0; i + 1 < count; ++i)
        if (p[i] > p[i + 1]) return false;
    return true;
}

static int bad_file(const char *what) {
    set_error("fl_matrix_load: %s", what);
    return FL_ERR_INVALID;
}

int win_install(fl_matrix *A, int bt, bool rows, int nb_file, const uint16_t *slot, const int *win_rows, long long D,
                const int *off, hipStream_t s) {
    ensure_batches(A, bt, rows ? 1 : 0, s);
    const int nb = A->nbatches, nnz = A->nnz;
    if (nb != nb_file) return bad_file("WIN schedule: batch count differs from the layout's");
    if (D < 0 || D > (long long)nnz || !ptr_ok(off, (size_t)nb + 1, 0, (int)D))
        return bad_file("WIN schedule: slot offsets are not a pointer array over the distinct rows");
    int maxslots = 0;
    for (int b = 0; b < nb; ++b) maxslots = std::max(maxslots, off[b + 1] - off[b]);
    if (maxslots > bt * 64) return bad_file("WIN schedule: a batch has more slots than entries");
    upload_to(A->win_slot, slot, nnz, s);
    upload_to(A->win_rows, win_rows, (size_t)D, s);
    upload_to(A->win_off, off, (size_t)nb + 1, s);
    TmpBuf<int> flag(1, s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    if (nnz)
        k_check_win<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->win_slot.p, A->win_off.p, bt * 64, nnz, A->m,
                                               A->win_rows.p, flag.p);
    FL_LAUNCH_CHECK();
    if (read1(flag.p, s)) {
        A->win_slot.release();
        A->win_rows.release();
        A->win_off.release();
        return bad_file("WIN schedule: slots or rows out of range");
    }
    A->win_maxslots = maxslots;
    A->win_ready = true;
    A->win_pack.release();
    return FL_OK;
}

int pb_install(fl_matrix *A, int bt, int R, int nb_file, int C, const int *chunk_batch, const int *chunk_off,
               long long ring, const uint16_t *lslot, const uint16_t *rowlocal, const int *band_start,
               const int *run_off, long long nruns, const int *run_lstart, const int *run_dest, const int *run_len,
               hipStream_t s) {
    ensure_batches(A, bt, 0, s);
    const int nb = A->nbatches, nnz = A->nnz, m = A->m;
    if (nb != nb_file) return bad_file("PB schedule: batch count differs from the layout's");
    if (R < 1024 || R > 16384) return bad_file("PB schedule: band rows outside [1024, 16384]");
    const int nbands = std::max((m + R - 1) / R, 1);
    if (C < 1 || (long long)C * nbands > 0x7fffffffLL) return bad_file("PB schedule: bad chunk count");
    if (ring < 0 || ring > nnz || nruns < 0 || nruns > (long long)nnz)
        return bad_file("PB schedule: product ring or run count out of range");
    if (chunk_batch[0] != 0 || chunk_batch[C] != nb) return bad_file("PB schedule: chunks do not cover the batches");
    for (int c = 0; c < C; ++c)
        if (chunk_batch[c] >= chunk_batch[c + 1] || (c > 0 && chunk_batch[c] % 8))
            return bad_file("PB schedule: chunk boundaries not increasing multiples of 8");
    const size_t nbs = (size_t)C * nbands + 1;
    if (!ptr_ok(band_start, nbs, 0, nnz)) return bad_file("PB schedule: band starts are not a pointer array over nnz");
    for (int c = 0; c < C; ++c) {
        const long long lo = band_start[(size_t)c * nbands], hi = band_start[(size_t)(c + 1) * nbands];
        if (chunk_off[c] < 0 || chunk_off[c] % 4 || chunk_off[c] > lo || hi - chunk_off[c] > ring)
            return bad_file("PB schedule: a chunk's products fall outside the product ring");
    }
    if (!ptr_ok(run_off, (size_t)nb + 1, 0, (int)nruns)) return bad_file("PB schedule: run offsets invalid");
    for (long long i = 0; i < nnz; ++i)
        if (rowlocal[i] >= R) return bad_file("PB schedule: a band-local row exceeds the band");
    upload_to(A->pb_lslot, lslot, nnz, s);
    upload_to(A->pb_rowlocal, rowlocal, nnz, s);
    upload_to(A->pb_band_start, band_start, nbs, s);
    upload_to(A->pb_run_off, run_off, (size_t)nb + 1, s);
    upload_to(A->pb_run_lstart, run_lstart, (size_t)nruns, s);
    upload_to(A->pb_run_dest, run_dest, (size_t)nruns, s);
    upload_to(A->pb_run_len, run_len, (size_t)nruns, s);
    TmpBuf<int> flag(1, s);
    FL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    if (nnz)
        k_check_pb<<<sg(nnz), kBlock, 0, s>>>(A->batch_of.p, A->pb_lslot.p, A->batch_start.p, bt * 64, nnz, (int)ring,
                                              A->pb_run_off.p, A->pb_run_lstart.p, A->pb_run_dest.p, A->pb_run_len.p,
                                              nb, flag.p);
    FL_LAUNCH_CHECK();
    if (read1(flag.p, s)) {
        A->pb_ready = false;
        return bad_file("PB schedule: staging slots or runs out of range");
    }
    A->pb_R = R;
    A->nbands = nbands;
    A->nruns = (int)nruns;
    A->pb_chunks = C;
    A->pb_ring = ring;
    A->pb_chunk_batch.assign(chunk_batch, chunk_batch + C + 1);
    A->pb_chunk_off.assign(chunk_off, chunk_off + C);
    A->pb_ready = true;
    A->pb_pack.release();
    return FL_OK;
}

// ------------------------------------------------------------------------------ kernels
constexpr int kBigBlock = 1024;

__device__ __forceinline__ void gatomic(float *p, double v) { unsafeAtomicAdd(p, (float)v); }
__device__ __forceinline__ void gatomic(double *p, double v) { unsafeAtomicAdd(p, v); }

// Re-ordered tiles of the batched schedules: tile i starts at entry ptr[i] and covers sorted
// columns [cx[i], cx[i] + min(64, n - cx[i])).
struct OrdTiles {
    const int *__restrict__ ptr;
    const int *__restrict__ cx;
    int n, T;
};

// A wave's tiles of one batch: t0..t1 of the re-ordered tiles, with their entry offset (mp)
// and first column (mc) held by lane t - t0.
struct WaveTiles {
    int t0, t1, mp, mc;
};

template <int NT>
__device__ __forceinline__ WaveTiles wave_tiles(const OrdTiles &ot, int bt, int batch) {
    const int lane = threadIdx.x & 63, wave = uni((int)(threadIdx.x >> 6));  // wave-uniform: scalar tile bounds
    const int tpw = bt / (NT / 64);
    WaveTiles w;
    w.t0 = batch * bt + wave * tpw;
    w.t1 = min(w.t0 + tpw, ot.T);
    const bool have = lane < w.t1 - w.t0;
    w.mp = have ? ot.ptr[w.t0 + lane] : 0;
    w.mc = have ? ot.cx[w.t0 + lane] : 0;
    return w;
}

// Value modes of the batched kernels: kPlain -- a value array (double, float or BF16) next to the
// 16-bit slot array; kSplit -- hi+lo BF16 pair in one 32-bit word (hi << 16 | lo) next to the slot
// array; kPacked -- value and slot in one 32-bit word (PackFmt), no slot array. kStream, or'ed in,
// reads the value and slot streams with the streaming hint (ld_stream; see stream_reads).
constexpr int kPlain = 0, kSplit = 1, kPacked = 2, kStream = 4;
constexpr int mode_kind(int mode) { return mode & 3; }

// Raw operands of U tiles of one wave (value, x, LDS slot; slot -1 = no entry), loaded one group
// ahead of the group being accumulated. Packed words are decoded to float as they are loaded.
template <typename V, typename X, int MODE, int U>
struct TileGroup {
    std::conditional_t<mode_kind(MODE) == kPacked, float, V> v[U];
    X xv[U];
    int sl[U];
};

template <typename V>
__device__ __forceinline__ auto widen(V v) {
    if constexpr (sizeof(V) == 2) return __uint_as_float((uint32_t)v << 16);
    else return v;
}

// base[l] for a wave-uniform base and a small lane index, addressed as base + 32-bit byte offset.
template <typename T>
__device__ __forceinline__ T lane_load(const T *__restrict__ base, uint32_t l) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + l * (uint32_t)sizeof(T));
}

// the same for a stream read once (values, slots)
template <typename T>
__device__ __forceinline__ T lane_load_stream(const T *__restrict__ base, uint32_t l) {
    return ld_stream(reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + l * (uint32_t)sizeof(T)));
}

// Tile u of the group: its entries are lanes [0, len) at entry offset ptr and sorted column cx (all
// uniform). Lanes past the tile read its first entry instead (every array has one when the matrix
// has tiles) and are given the slot `dummy + lane`, a per-lane spare LDS element past the batch's
// slots, so that loads, products and LDS updates need no branches and no two spare lanes collide.
template <typename V, typename X, int MODE, int U>
__device__ __forceinline__ void load_group(const OrdTiles &ot, const WaveTiles &w, int t,
                                           const uint16_t *__restrict__ slot, const V *__restrict__ val,
                                           const X *__restrict__ x, const PackFmt &pf, int dummy,
                                           TileGroup<V, X, MODE, U> &G) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int len = 0, ptr = 0, cx = 0;
        if (t + u < w.t1) {
            ptr = __builtin_amdgcn_readlane(w.mp, t + u - w.t0);
            cx = __builtin_amdgcn_readlane(w.mc, t + u - w.t0);
            len = min(64, ot.n - cx);
        }
        const bool act = lane < len;
        const uint32_t l = act ? (uint32_t)lane : 0u;
        // uniform base (scalar registers) + 32-bit lane byte offset: no 64-bit address arithmetic
        constexpr bool S = (MODE & kStream) != 0;
        if constexpr (mode_kind(MODE) == kPacked) {
            const uint32_t wd = S ? lane_load_stream(val + ptr, l) : lane_load(val + ptr, l);
            G.sl[u] = act ? (int)(wd & pf.imask) : dummy + lane;
            G.v[u] = pack_decode(wd, pf);
        } else {
            const uint16_t sv = S ? lane_load_stream(slot + ptr, l) : lane_load(slot + ptr, l);
            G.sl[u] = act ? (int)sv : dummy + lane;
            G.v[u] = S ? lane_load_stream(val + ptr, l) : lane_load(val + ptr, l);
        }
        G.xv[u] = lane_load(x + cx, l);
    }
}

// Product of entry u in FP64 (FP64 mode) or as the FP64 sum of FP32 products (FP32 / BF16 /
// hi+lo BF16).
template <typename V, typename X, int MODE, int U>
__device__ __forceinline__ double group_product(const TileGroup<V, X, MODE, U> &G, int u) {
    if constexpr (sizeof(X) == 8) {
        return (double)widen(G.v[u]) * G.xv[u];
    } else if constexpr (mode_kind(MODE) == kSplit) {  // hi+lo BF16 pair in one word: hi << 16 | lo
        const float hi = __uint_as_float(G.v[u] & 0xffff0000u), lo = __uint_as_float(G.v[u] << 16);
        return (double)(hi * G.xv[u]) + (double)(lo * G.xv[u]);
    } else {
        return (double)(widen(G.v[u]) * G.xv[u]);
    }
}

// WIN: add each product to its row's LDS slot (lanes past a tile add to their spare slot).
template <typename V, typename X, int MODE, int U>
__device__ __forceinline__ void add_group(double *acc, const TileGroup<V, X, MODE, U> &G) {
#pragma unroll
    for (int u = 0; u < U; ++u) atomicAdd(&acc[G.sl[u]], group_product(G, u));
}

// PB phase 1: store each product at its band-major staging slot (or its lane's spare slot).
template <typename V, typename X, int MODE, int U, typename P>
__device__ __forceinline__ void stage_group(P *stage, const TileGroup<V, X, MODE, U> &G) {
#pragma unroll
    for (int u = 0; u < U; ++u) stage[G.sl[u]] = (P)group_product(G, u);
}

// One batch per workgroup: LDS slots of the batch's distinct rows are cleared, every entry adds
// its product to its row's slot (FP64 LDS atomics), and each slot is flushed to y once.
// The operands of the next U tiles are loaded before the current U tiles are accumulated, and
// the first group before the barrier that ends the slot clearing.
// ASSIGN: y was cleared before the launch and this is its only accumulation, so a row owned by
// this batch is stored without reading y first (shared rows still take an atomic).
template <typename V, typename X, typename Y, int MODE, int U, int BT, int NT, bool ASSIGN>
__global__ __launch_bounds__(NT) void k_win(OrdTiles ot, int bt, const uint16_t *__restrict__ slot,
                                            const V *__restrict__ val, PackFmt pf, int spare,
                                            const X *__restrict__ x, const int *__restrict__ win_rows,
                                            const int *__restrict__ win_off, Y *__restrict__ y) {
    extern __shared__ double acc[];
    const int b = blockIdx.x;
    const WaveTiles w = wave_tiles<NT>(ot, bt, b);
    const int s0 = win_off[b], ns = win_off[b + 1] - s0;
    // prefetch the first rows this thread will flush (ns <= BT*64 = K * NT; most batches have
    // far fewer distinct rows than entries); the rest are loaded at the flush
    constexpr int K = BT * 64 / NT;
    constexpr int KP = K < 2 ? K : 2;
    int frow[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int i = threadIdx.x + k * NT;
        frow[k] = (i < ns) ? win_rows[s0 + i] : 0;
    }
    TileGroup<V, X, MODE, U> cur, nxt;
    if (w.t0 < w.t1) load_group<V, X, MODE, U>(ot, w, w.t0, slot, val, x, pf, spare, cur);
    for (int i = threadIdx.x; i < ns; i += NT) acc[i] = 0.0;
    __syncthreads();
    for (int t = w.t0; t < w.t1; t += U) {
        const bool more = t + U < w.t1;
        if (more) load_group<V, X, MODE, U>(ot, w, t + U, slot, val, x, pf, spare, nxt);
        add_group<V, X, MODE, U>(acc, cur);
        if (more) cur = nxt;
    }
    __syncthreads();
    auto flush = [&](int i, int r) {
        if (r >= 0) {  // only this batch touches row r
            if constexpr (ASSIGN) y[r] = (Y)acc[i];
            else y[r] += (Y)acc[i];
        } else {
            gatomic(&y[r & 0x7fffffff], acc[i]);
        }
    };
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < ns) flush(i, frow[k]);
    }
    for (int i = threadIdx.x + KP * NT; i < ns; i += NT) flush(i, win_rows[s0 + i]);
}


// Phase 1 over batches [boff, boff + gridDim.x) (one chunk); boff is a multiple of 8, so batch b
// runs on XCD b % 8 as the run order of the product array assumes.
template <typename V, typename X, typename P, int MODE, int U, int NT = kBigBlock>
__global__ __launch_bounds__(NT) void k_pb1(OrdTiles ot, int bt, int boff, const uint16_t *__restrict__ lslot,
                                                   const V *__restrict__ val, PackFmt pf,
                                                   const X *__restrict__ x, const int *__restrict__ run_off,
                                                   const int *__restrict__ run_lstart, const int *__restrict__ run_dest,
                                                   const int *__restrict__ run_len, P *__restrict__ bins) {
    extern __shared__ double stage_raw[];
    P *stage = reinterpret_cast<P *>(stage_raw);
    const int b = boff + blockIdx.x;
    const int wave = uni((int)(threadIdx.x >> 6));
    // Wave w copies runs r0 + w + (NT/64)*j. Prefetch the metadata of its first 64 runs (lane j holds
    // run j) before the tile loop so the copy loop only reads registers.
    const int r0 = run_off[b], r1 = run_off[b + 1];
    int mlen = 0, mls = 0, mdst = 0;
    {
        const int r = r0 + wave + (NT / 64) * (int)(threadIdx.x & 63);
        if (r < r1) {
            mlen = run_len[r];
            mls = run_lstart[r];
            mdst = run_dest[r];
        }
    }
    // tile loop, operands loaded one group ahead of the group being staged
    const WaveTiles w = wave_tiles<NT>(ot, bt, b);
    TileGroup<V, X, MODE, U> cur, nxt;
    if (w.t0 < w.t1) load_group<V, X, MODE, U>(ot, w, w.t0, lslot, val, x, pf, bt * 64, cur);
    for (int t = w.t0; t < w.t1; t += U) {
        const bool more = t + U < w.t1;
        if (more) load_group<V, X, MODE, U>(ot, w, t + U, lslot, val, x, pf, bt * 64, nxt);
        stage_group<V, X, MODE, U>(stage, cur);
        if (more) cur = nxt;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int base = 0;; base += 64) {
        for (int j = 0; j < 64; ++j) {
            const int len = __builtin_amdgcn_readlane(mlen, j);
            if (len == 0) break;  // past the last run of this wave (every real run has len >= 1)
            const int ls = __builtin_amdgcn_readlane(mls, j), d = __builtin_amdgcn_readlane(mdst, j);
            for (int i = lane; i < len; i += 64) bins[d + i] = stage[ls + i];
        }
        if (r0 + wave + (NT / 64) * (base + 64) >= r1) break;
        const int r = r0 + wave + (NT / 64) * (base + 64 + lane);
        mlen = mls = mdst = 0;
        if (r < r1) {
            mlen = run_len[r];
            mls = run_lstart[r];
            mdst = run_dest[r];
        }
    }
}

// Phase-2 vector loads: 4 products (16 B) with their four 16-bit rows (8 B) in FP32 mode, 2
// products with two rows in FP64 mode.
template <typename P>
struct PbVec;
template <>
struct PbVec<float> {
    static constexpr int N = 4;
    using V = float4;
    using R = uint2;
    static __device__ __forceinline__ void add(double *acc, const V &v, const R &r) {
        atomicAdd(&acc[r.x & 0xffffu], (double)v.x);
        atomicAdd(&acc[r.x >> 16], (double)v.y);
        atomicAdd(&acc[r.y & 0xffffu], (double)v.z);
        atomicAdd(&acc[r.y >> 16], (double)v.w);
    }
};
template <>
struct PbVec<double> {
    static constexpr int N = 2;
    using V = double2;
    using R = uint32_t;
    static __device__ __forceinline__ void add(double *acc, const V &v, const R &r) {
        atomicAdd(&acc[r & 0xffffu], v.x);
        atomicAdd(&acc[r >> 16], v.y);
    }
};

// ASSIGN: y = A x (every row of the band is written, so y needs no clearing); else y += A x.
// band_start points at one chunk's band table; the product at band-major position i is
// bins[i - off] (off a multiple of 4), its row rowlocal[i]. FOLD adds a part of y kept in another
// row order (PbFold), in FP64: kFoldGather adds fys[(r & ~(kHybWin - 1)) + fmap[r]] to row r as it
// is written; kFoldScan (bands of kHybWin rows, so band b is window b) reads the band's fys in
// order after the products and adds fys[p] to the accumulator of row fmap[p] of the band.
constexpr int kFoldNone = 0, kFoldGather = 1, kFoldScan = 2;
template <typename P, typename Y, int U, bool ASSIGN, int FOLD, bool S>
__global__ __launch_bounds__(kBigBlock) void k_pb2(int R, int m, const int *__restrict__ band_start,
                                                   const P *__restrict__ bins, int off,
                                                   const uint16_t *__restrict__ rowlocal, Y *__restrict__ y,
                                                   const Y *__restrict__ fys, const uint16_t *__restrict__ fmap) {
    extern __shared__ double acc[];
    using T = PbVec<P>;
    constexpr int N = T::N;
    const int b = blockIdx.x;
    for (int i = threadIdx.x; i < R; i += kBigBlock) acc[i] = 0.0;
    __syncthreads();
    const int i0 = band_start[b], i1 = band_start[b + 1];
    // [i0, a0) head and [t0, i1) tail hold fewer than N entries each; [a0, t0) is N-aligned
    const int a0 = min((i0 + N - 1) / N * N, i1);
    const int nv = (i1 - a0) / N;
    const int t0 = a0 + nv * N;
    if ((int)threadIdx.x < a0 - i0)
        atomicAdd(&acc[rowlocal[i0 + threadIdx.x]], (double)bins[i0 - off + threadIdx.x]);
    if ((int)threadIdx.x < i1 - t0)
        atomicAdd(&acc[rowlocal[t0 + threadIdx.x]], (double)bins[t0 - off + threadIdx.x]);
    const typename T::V *__restrict__ bv = reinterpret_cast<const typename T::V *>(bins + (a0 - off));
    const typename T::R *__restrict__ rv = reinterpret_cast<const typename T::R *>(rowlocal + a0);
    // groups of U vectors per thread, each loaded while the previous group is accumulated
    // (two register sets used in turn)
    typename T::V va[U], vb[U];
    typename T::R ra[U], rb[U];
    auto full = [&](int k) { return k + (U - 1) * kBigBlock < nv; };
    auto load = [&](int k, typename T::V *v, typename T::R *r) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (S) {
                v[u] = ld_stream(&bv[k + u * kBigBlock]);
                r[u] = ld_stream(&rv[k + u * kBigBlock]);
            } else {
                v[u] = bv[k + u * kBigBlock];
                r[u] = rv[k + u * kBigBlock];
            }
        }
    };
    auto add = [&](const typename T::V *v, const typename T::R *r) {
#pragma unroll
        for (int u = 0; u < U; ++u) T::add(acc, v[u], r[u]);
    };
    int k = threadIdx.x;
    if (full(k)) {
        load(k, va, ra);
        for (;;) {
            bool more = full(k + U * kBigBlock);
            if (more) load(k + U * kBigBlock, vb, rb);
            add(va, ra);
            k += U * kBigBlock;
            if (!more) break;
            more = full(k + U * kBigBlock);
            if (more) load(k + U * kBigBlock, va, ra);
            add(vb, rb);
            k += U * kBigBlock;
            if (!more) break;
        }
    }
    for (; k < nv; k += kBigBlock) T::add(acc, bv[k], rv[k]);
    __syncthreads();
    const int r0 = b * R, rows = min(R, m - r0);
    if constexpr (FOLD == kFoldScan) {
        // positions [r0, r0 + rows) hold the band's rows in another order: each row once, so
        // these adds do not collide. A whole window: each thread loads its 16 consecutive
        // positions with 16-byte loads, all in flight before the adds.
        static_assert(kHybWin == 16 * kBigBlock, "scan fold: 16 positions per thread");
        if (rows == kHybWin) {
            const long long p0 = (long long)r0 + 16 * (long long)threadIdx.x;
            constexpr int NV = 16 * (int)sizeof(Y) / 16;
            float4 yv[NV];
            uint4 mv[2];
#pragma unroll
            for (int j = 0; j < NV; ++j) yv[j] = reinterpret_cast<const float4 *>(fys + p0)[j];
            mv[0] = reinterpret_cast<const uint4 *>(fmap + p0)[0];
            mv[1] = reinterpret_cast<const uint4 *>(fmap + p0)[1];
            const Y *yy = reinterpret_cast<const Y *>(yv);
            const uint16_t *mm = reinterpret_cast<const uint16_t *>(mv);
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[mm[j]] += (double)yy[j];
        } else {
            for (int p = threadIdx.x; p < rows; p += kBigBlock)
                acc[fmap[(long long)r0 + p]] += (double)fys[(long long)r0 + p];
        }
        __syncthreads();
    }
    for (int r = threadIdx.x; r < rows; r += kBigBlock) {
        double v = acc[r];
        if constexpr (FOLD == kFoldGather) {
            const long long g = (long long)r0 + r;
            v += (double)fys[(g & ~(long long)(kHybWin - 1)) + fmap[g]];
        }
        if constexpr (ASSIGN) y[r0 + r] = (Y)v;
        else y[r0 + r] += (Y)v;
    }
}

// ------------------------------------------------------------------------------ launch
static OrdTiles ord_tiles(const fl_matrix *A) {
    OrdTiles ot;
    ot.ptr = A->ord_ptr.p;
    ot.cx = A->ord_cx.p;
    ot.n = A->n;
    ot.T = A->tilenum;
    return ot;
}

template <typename K>
static void allow_lds(K kern, size_t bytes) {
    FL_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

template <typename V, typename X, typename Y, int MODE, int NT, bool ASSIGN>
static auto win_kernel(int bt) {
    constexpr int U = sizeof(X) == 4 ? 8 : 4;  // tiles per group: 8 keeps the 4-byte-x kinds within 64 VGPRs
    return (bt == 256) ? k_win<V, X, Y, MODE, U, 256, NT, ASSIGN>
           : (bt == 128) ? k_win<V, X, Y, MODE, U, 128, NT, ASSIGN> : k_win<V, X, Y, MODE, U, 64, NT, ASSIGN>;
}

template <typename V, typename X, typename Y, int MODE>
static void launch_win(const fl_matrix *A, const V *val, const X *x, Y *y, bool assign,
                       hipStream_t s) {
    const PackFmt pf = A->win_pf;
    const uint16_t *slot = mode_kind(MODE) == kPacked ? nullptr : A->win_slot.p;
    // LDS for the most distinct rows any batch has (at most bt*64). Workgroups of 4 waves per 64
    // tiles (16 tiles per wave; 512 threads at most): more, smaller workgroups per CU overlap
    // the three phases of a batch (slot set-up, accumulation, flush) better than 1024 threads.
    // + 64 spare slots, one per lane, for the lanes past a tile's end
    const size_t lds = ((size_t)std::max(A->win_maxslots, 1) + 64) * sizeof(double);
    const int nt = std::min(A->bt * 4, 512);
    auto kern = (nt == 256) ? (assign ? win_kernel<V, X, Y, MODE, 256, true>(A->bt)
                                      : win_kernel<V, X, Y, MODE, 256, false>(A->bt))
                            : (assign ? win_kernel<V, X, Y, MODE, 512, true>(A->bt)
                                      : win_kernel<V, X, Y, MODE, 512, false>(A->bt));
    allow_lds(kern, lds);
    kern<<<A->nbatches, nt, lds, s>>>(ord_tiles(A), A->bt, slot, val, pf, std::max(A->win_maxslots, 1), x,
                                      A->win_rows.p, A->win_off.p, y);
    FL_LAUNCH_CHECK();
}

template <typename V, typename X, typename P, typename Y, int MODE>
static void launch_pb(const fl_matrix *A, const V *val, const X *x, P *bins, Y *y, bool assign,
                      hipStream_t s, const PbFold *fold) {
    constexpr int U = sizeof(X) == 4 ? 8 : 4;  // tiles per group, as in WIN
    const PackFmt pf = A->pb_pf;
    const uint16_t *lslot = mode_kind(MODE) == kPacked ? nullptr : A->pb_lslot.p;
    constexpr bool S = (MODE & kStream) != 0;
    const size_t lds1 = ((size_t)A->bt * 64 + 64) * sizeof(P);  // + one spare slot per lane
    // 4 waves per 64 tiles of the batch (16 tiles per wave)
    const int nt = A->bt * 4;
    auto k1 = (nt == 256) ? k_pb1<V, X, P, MODE, U, 256>
              : (nt == 512) ? k_pb1<V, X, P, MODE, U, 512> : k_pb1<V, X, P, MODE, U, 1024>;
    allow_lds(k1, std::max<size_t>(lds1, 8));
    const size_t lds2 = (size_t)A->pb_R * sizeof(double);
    const Y *fys = fold ? (const Y *)fold->ys : nullptr;
    const bool scan = fys && fold->perm && A->pb_R == kHybWin;
    const uint16_t *fmap = !fys ? nullptr : scan ? fold->perm : fold->iperm;
    allow_lds(k_pb2<P, Y, 8, true, kFoldNone, S>, lds2);
    allow_lds(k_pb2<P, Y, 8, false, kFoldNone, S>, lds2);
    if (scan) {
        allow_lds(k_pb2<P, Y, 8, true, kFoldScan, S>, lds2);
        allow_lds(k_pb2<P, Y, 8, false, kFoldScan, S>, lds2);
    } else if (fys) {
        allow_lds(k_pb2<P, Y, 8, true, kFoldGather, S>, lds2);
        allow_lds(k_pb2<P, Y, 8, false, kFoldGather, S>, lds2);
    }
    // chunk by chunk: phase 1 writes the chunk's products to the front of the product array,
    // phase 2 adds them into y (the first chunk assigns y when asked to)
    for (int c = 0; c < A->pb_chunks; ++c) {
        const int b0 = A->pb_chunk_batch[c], nbc = A->pb_chunk_batch[c + 1] - b0;
        k1<<<nbc, nt, lds1, s>>>(ord_tiles(A), A->bt, b0, lslot, val, pf, x, A->pb_run_off.p,
                                 A->pb_run_lstart.p, A->pb_run_dest.p, A->pb_run_len.p, bins);
        FL_LAUNCH_CHECK();
        // the first chunk's phase 2 writes every row once: it adds the folded part
        const int fm = (fys && c == 0) ? (scan ? kFoldScan : kFoldGather) : kFoldNone;
        auto k2 = (assign && c == 0)
                      ? (fm == kFoldScan ? k_pb2<P, Y, 8, true, kFoldScan, S>
                         : fm == kFoldGather ? k_pb2<P, Y, 8, true, kFoldGather, S> : k_pb2<P, Y, 8, true, kFoldNone, S>)
                      : (fm == kFoldScan ? k_pb2<P, Y, 8, false, kFoldScan, S>
                         : fm == kFoldGather ? k_pb2<P, Y, 8, false, kFoldGather, S> : k_pb2<P, Y, 8, false, kFoldNone, S>);
        k2<<<A->nbands, kBigBlock, lds2, s>>>(A->pb_R, A->m, A->pb_band_start.p + (size_t)c * A->nbands, bins,
                                              A->pb_chunk_off[c], A->pb_rowlocal.p, y, fys, fmap);
        FL_LAUNCH_CHECK();
    }
}

// The caller (spmv) has checked that the schedule and the value copy are prepared. PB writes
// its products to pb_bins when given (the caller guarantees pb_ring entries), else to the
// matrix's own product array, allocated on first use.
static std::atomic<int> g_stream_policy{FL_STREAM_AUTO};
int stream_policy() { return g_stream_policy.load(std::memory_order_relaxed); }

long long batched_stream_bytes(const fl_matrix *A, int kernel, int kind) {
    const int v = kind == FL_VAL_F64 ? 8 : kind == FL_VAL_BF16 ? 2 : 4;  // value (packed: with its slot)
    long long per = v + (kind == FL_VAL_PACKED ? 0 : 2);
    if (kernel == FL_KERNEL_PB) per += (kind == FL_VAL_F64 ? 8 : 4) + 2;  // products and rows read back
    return (long long)A->nnz * per;
}

int spmv_batched(fl_matrix *A, int kernel, int kind, const void *x, void *y, bool assign, hipStream_t s,
                 void *pb_bins, const PbFold *fold, int stream) {
    const bool win = (kernel == FL_KERNEL_WIN || kernel == FL_KERNEL_WINR);
    double *bins64 = (double *)pb_bins;
    float *bins32 = (float *)pb_bins;
    if (!win && !pb_bins) {
        if (kind == FL_VAL_F64) { if (!A->pb_bins64.p) A->pb_bins64.alloc(std::max(A->nnz, 1)); }
        else if (!A->pb_bins32.p) A->pb_bins32.alloc(std::max(A->nnz, 1));
        bins64 = A->pb_bins64.p;
        bins32 = A->pb_bins32.p;
    }
    const bool st = stream < 0 ? stream_reads(batched_stream_bytes(A, kernel, kind)) : stream != 0;
    switch (kind) {
        case FL_VAL_F64:
            if (win) st ? launch_win<double, double, double, kPlain | kStream>(A, A->val64.p, (const double *)x, (double *)y, assign, s)
                        : launch_win<double, double, double, kPlain>(A, A->val64.p, (const double *)x, (double *)y, assign, s);
            else st ? launch_pb<double, double, double, double, kPlain | kStream>(A, A->val64.p, (const double *)x, bins64, (double *)y, assign, s, fold)
                    : launch_pb<double, double, double, double, kPlain>(A, A->val64.p, (const double *)x, bins64, (double *)y, assign, s, fold);
            return FL_OK;
        case FL_VAL_F32:
            if (win) st ? launch_win<float, float, float, kPlain | kStream>(A, A->val32.p, (const float *)x, (float *)y, assign, s)
                        : launch_win<float, float, float, kPlain>(A, A->val32.p, (const float *)x, (float *)y, assign, s);
            else st ? launch_pb<float, float, float, float, kPlain | kStream>(A, A->val32.p, (const float *)x, bins32, (float *)y, assign, s, fold)
                    : launch_pb<float, float, float, float, kPlain>(A, A->val32.p, (const float *)x, bins32, (float *)y, assign, s, fold);
            return FL_OK;
        case FL_VAL_BF16:
            if (win) st ? launch_win<uint16_t, float, float, kPlain | kStream>(A, A->bf16_hi.p, (const float *)x, (float *)y, assign, s)
                        : launch_win<uint16_t, float, float, kPlain>(A, A->bf16_hi.p, (const float *)x, (float *)y, assign, s);
            else st ? launch_pb<uint16_t, float, float, float, kPlain | kStream>(A, A->bf16_hi.p, (const float *)x, bins32, (float *)y, assign, s, fold)
                    : launch_pb<uint16_t, float, float, float, kPlain>(A, A->bf16_hi.p, (const float *)x, bins32, (float *)y, assign, s, fold);
            return FL_OK;
        case FL_VAL_BF16X2:
            if (win) st ? launch_win<uint32_t, float, double, kSplit | kStream>(A, A->bf16_pair.p, (const float *)x, (double *)y, assign, s)
                        : launch_win<uint32_t, float, double, kSplit>(A, A->bf16_pair.p, (const float *)x, (double *)y, assign, s);
            else st ? launch_pb<uint32_t, float, float, double, kSplit | kStream>(A, A->bf16_pair.p, (const float *)x, bins32, (double *)y, assign, s, fold)
                    : launch_pb<uint32_t, float, float, double, kSplit>(A, A->bf16_pair.p, (const float *)x, bins32, (double *)y, assign, s, fold);
            return FL_OK;
        case FL_VAL_PACKED:
            if (win) st ? launch_win<uint32_t, float, float, kPacked | kStream>(A, A->win_pack.p, (const float *)x, (float *)y, assign, s)
                        : launch_win<uint32_t, float, float, kPacked>(A, A->win_pack.p, (const float *)x, (float *)y, assign, s);
            else st ? launch_pb<uint32_t, float, float, float, kPacked | kStream>(A, A->pb_pack.p, (const float *)x, bins32, (float *)y, assign, s, fold)
                    : launch_pb<uint32_t, float, float, float, kPacked>(A, A->pb_pack.p, (const float *)x, bins32, (float *)y, assign, s, fold);
            return FL_OK;
        default:
            set_error("fl_spmv: unknown value kind %d", kind);
            return FL_ERR_INVALID;
    }
}

}  // namespace fl

extern "C" int fl_set_stream_policy(int policy) {
    if (policy != FL_STREAM_AUTO && policy != FL_STREAM_NEVER && policy != FL_STREAM_ALWAYS) {
        fl::set_error("fl_set_stream_policy: policy must be FL_STREAM_AUTO, FL_STREAM_NEVER or FL_STREAM_ALWAYS");
        return FL_ERR_INVALID;
    }
    fl::g_stream_policy.store(policy, std::memory_order_relaxed);
    return FL_OK;
}

