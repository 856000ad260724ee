// Device-resident FastLoad matrix (internal layout; the public handle is opaque).
#pragma once

#include <memory>
#include <vector>

#include "fl_common.h"
#include "fl_walk.h"

// Packed entries of the batched schedules (FL_VAL_PACKED): one 32-bit word per entry holding the
// value rounded to `m` mantissa bits and the schedule's LDS index (WIN slot / PB staging slot):
//   bit 31 sign | [30 .. ib+m] exponent code (e bits) | [ib+m-1 .. ib] mantissa | [ib-1 .. 0] index
// exponent code 0 = zero; code c > 0 = the FP32 biased exponent c + ebase.
struct PackFmt {
    uint32_t imask = 0, fmask = 0, mmask = 0, ebias = 0;
    int ib = 0, e = 0, m = 0, ebase = 0;
};

// f = exponent code and mantissa; FP32 bits = (f << (23 - m)) + (ebase << 23) with the sign bit
// copied; code 0 (f <= mmask) is zero. Five VALU operations.
__device__ __forceinline__ float pack_decode(uint32_t w, const PackFmt &p) {
    const uint32_t f = (w >> p.ib) & p.fmask;
    const uint32_t bits = ((f << (23 - p.m)) + p.ebias) | (w & 0x80000000u);
    return f > p.mmask ? __uint_as_float(bits) : 0.0f;
}

// Contiguous row ranges of the row-ordered kernels, one per XCD, of about equal entry counts:
// XCD x (the workgroups with blockIdx % 8 == x) owns rows [r[x], r[x+1]).
struct RowSplit {
    int r[9];
};

struct fl_matrix {
    int m = 0, n = 0, nnz = 0, W = 32;
    int L = 0;        // levels (h_count)
    int tilenum = 0;
    int segsum = 0;
    int max_col_nnz = 0;

    // column sort (ColSort): perm[i] = original column of sorted position i; sortnnz[i] its length
    fl::DevBuf<int> perm, sortnnz;
    // per-level table (prepare/countTileNum/auxiliaryTrans): col, base depth, tall,
    // tile offset (L+1), nnz offset (L+1)
    fl::DevBuf<int> lvl_col, lvl_base, lvl_tall, lvl_tileoff, lvl_nnzoff;
    // reference tile metadata (formatTransform outputs)
    fl::DevBuf<int> tile_ptr, tile_len, tile_colidx, tile_format, countrow, segmentoffset;
    // per-nnz arrays in tile order
    fl::DevBuf<int> sortrowidx, sortrowindex;
    fl::DevBuf<double> val64;
    // precision variants (built on demand from val64)
    fl::DevBuf<float> val32;
    fl::DevBuf<uint16_t> bf16_hi, bf16_lo;
    fl::DevBuf<uint32_t> bf16_pair;  // hi << 16 | lo: one load per entry in the batched schedules
    // FAST kernel: level index of the first tile of every 64-tile group
    fl::DevBuf<int> group_level;

    // Batched schedules (W = 64): the tiles are re-ordered by 256-column bucket of their first
    // sorted column (stable), so a batch of `bt` consecutive re-ordered tiles reads a narrow
    // window of x; ord_ptr/ord_cx give each re-ordered tile's entry offset and first column
    // (len = min(64, n - cx)).
    int bt = 0, nbatches = 0;
    fl::DevBuf<int> ord_ptr, ord_cx;
    int ord_kind = -1;  // tile order of ord_ptr/ord_cx: 0 column bucket, 1 row window
    fl::DevBuf<int> batch_of;     // [nnz] batch index of every entry (tile order), built on demand
    fl::DevBuf<int> batch_start;  // [nbatches+1] prefix of batch entry counts
    // WIN: rows of each batch compacted to LDS slots.
    //   win_slot[g]  slot of entry g inside its batch; win_rows[win_off[b] + s] = row of slot s
    bool win_ready = false;
    int win_maxslots = 0;  // most distinct rows (LDS slots) of any batch
    fl::DevBuf<uint16_t> win_slot;
    fl::DevBuf<int> win_rows, win_off;
    // PB (two-phase propagation blocking): rows split in bands of pb_R rows.
    //   phase 1 stages products in LDS at pb_lslot[g] (band-major inside the batch) and copies
    //   runs (batch, band) to the band-major product array; phase 2 accumulates each band.
    //   The batches are split into pb_chunks consecutive chunks (boundaries at multiples of 8
    //   batches); chunk c's products occupy band-major positions [band_start[c*nbands],
    //   band_start[(c+1)*nbands]) and are stored at position - pb_chunk_off[c], so every chunk
    //   reuses the front of the product array and a chunk's products can stay in the Infinity
    //   Cache between its two phases.
    int pb_R = 0, nbands = 0, nruns = 0, pb_chunks = 1;
    long long pb_ring = 0;            // product-array entries the chunks need (the largest chunk)
    bool pb_ready = false;
    std::vector<int> pb_chunk_batch;  // [pb_chunks+1] first batch of every chunk
    std::vector<int> pb_chunk_off;    // [pb_chunks] product-array offset (a multiple of 4)
    fl::DevBuf<uint16_t> pb_lslot, pb_rowlocal;
    fl::DevBuf<int> pb_band_start, pb_run_off, pb_run_lstart, pb_run_dest, pb_run_len;
    fl::DevBuf<float> pb_bins32;
    fl::DevBuf<double> pb_bins64;

    // FL_VAL_PACKED: value + index words of the prepared WIN and PB schedules (built by
    // fl_matrix_prepare_values(FL_VAL_PACKED) and for schedules prepared after it), their formats,
    // and the matrix's exponent range (emin/emax of the non-zero values' FP32 exponents; pack_ok
    // false when a value is not finite or too large for FP32)
    bool pack_wanted = false, pack_range_known = false, pack_ok = false;
    int pack_emin = 0, pack_emax = 0;
    fl::DevBuf<uint32_t> win_pack, pb_pack;
    PackFmt win_pf, pb_pf;

    // ROW (multi-right-hand-side): the entries re-ordered by row (CSR of A, columns in the
    // original numbering, so X needs no permutation): row_ptr [m+1], row_col/row_val [nnz].
    bool row_ready = false;
    RowSplit row_split{};
    fl::DevBuf<int> row_ptr, row_col;
    fl::DevBuf<float> row_val32;
    fl::DevBuf<double> row_val64;

    // HYB: the last hyb_H sorted columns (the longest: the column sort is ascending) are
    // multiplied with x held in LDS, row by row; the other columns form a sub-matrix (hyb_light,
    // its sorted column i = sorted column i here) with its own PB schedule, whose first phase 2
    // adds the heavy part's y. The rows are ordered by heavy entry count, longest first, inside
    // windows of kHybWin rows (hyb_iperm); slice s = sorted positions [64 s, 64 s + 64) is stored
    // as hyb_np[s] pair-slots of 64 lanes: slot (q, lane) holds the lane's row's entries 2q and
    // 2q + 1 (value pair + two 16-bit columns in one word), the rows shorter than the slice's
    // longest padded with value 0 and column H (a zero in LDS).
    bool hyb_ready = false;
    int hyb_H = 0, hyb_bt = 0, hyb_R = 0, hyb_nslice = 0, hyb_grid = 0;
    long long hyb_nnz = 0;            // heavy entries
    long long hyb_slots = 0;          // pair-slots of all slices (64 per pair row)
    std::unique_ptr<fl_matrix> hyb_light;
    fl::DevBuf<long long> hyb_sb;     // [nslice + 1] first slot of every slice
    fl::DevBuf<int> hyb_np;           // [nslice] pair rows of every slice (32-bit: scalar loads)
    fl::DevBuf<uint16_t> hyb_iperm;   // [m] sorted position of each row inside its window
    fl::DevBuf<uint16_t> hyb_perm;    // [m] row (inside its window) at each sorted position
    fl::DevBuf<uint32_t> hyb_col2;    // [hyb_slots + pad] two columns per slot
    fl::DevBuf<double> hyb_val64;     // [2 (hyb_slots + pad)] two values per slot
    fl::DevBuf<float> hyb_val32;      // built with the FP32 values
    fl::DevBuf<float> hyb_ys32;       // [m] heavy part of y by sorted position (first use)
    fl::DevBuf<double> hyb_ys64;
};

namespace fl {

inline LevelTab level_tab(const fl_matrix *A) {
    LevelTab lt;
    lt.col = A->lvl_col.p;
    lt.tall = A->lvl_tall.p;
    lt.tileoff = A->lvl_tileoff.p;
    lt.nnzoff = A->lvl_nnzoff.p;
    lt.group_level = A->group_level.p;
    lt.n = A->n;
    lt.L = A->L;
    return lt;
}

// Pipeline stages shared by the host API and the device-resident builder.
// Stable column sort: perm/sortnnz from per-column counts (device arrays of n ints).
void colsort_device(int n, const int *d_counts, int *d_perm, int *d_sortnnz, int max_key, hipStream_t s);

// Level table + tiles + data layout + classification into A (A->n, A->nnz, A->W set).
// Source of element (sorted column i, depth d):
//   tmp mode (d_P != nullptr): tmp_row/tmp_val[P[i] + d]
//   fused mode: rowidx/val[colptr[perm[i]] + d]
//   tiled mode (tiled = true): A->sortrowidx / A->val64 already hold the elements in tile order
//     (a layout read back from a file); only the level table, tile metadata and classification
//     are computed
struct ElemSource {
    bool tiled = false;
    const int *P = nullptr;
    const int *colptr = nullptr;
    const int *perm = nullptr;
    const int *rowidx = nullptr;
    const double *val = nullptr;
};
void transform_device(fl_matrix *A, const int *d_sortnnz, const ElemSource &src, hipStream_t s, float *t_tran_ms,
                      float *t_clas_ms);

// y = A x (assign) or y += A x with one schedule. work: 2*m floats for the FAST kernel's split
// BF16 mode. pb_bins: product array for the PB schedule (at least pb_ring entries of the value
// kind's product type, 16-byte aligned); nullptr = the matrix's own, allocated on first use.
int spmv(fl_matrix *A, int kernel, int kind, const void *x, void *y, void *work, bool assign, hipStream_t s,
         void *pb_bins = nullptr);

// The HYB schedule's heavy part, which the first chunk's phase 2 of its PB sub-matrix adds to y.
// ys holds the heavy sum of row r at (r & ~(kHybWin - 1)) + iperm[r] (its position inside its
// window); perm is the inverse (position -> row inside the window). When the PB bands are the
// windows (band rows == kHybWin), phase 2 reads ys in order and adds ys[p] to row perm[p] of the
// band in LDS; otherwise it gathers ys through iperm as it writes each row.
struct PbFold {
    const void *ys = nullptr;
    const uint16_t *iperm = nullptr;
    const uint16_t *perm = nullptr;
};

// The WIN/WINR/PB launches (fl_sched.hip); fold: PB only, see PbFold. stream: 1 / 0 reads the
// value, slot and product streams with / without the streaming hint, -1 decides by size
// (stream_reads of batched_stream_bytes).
int spmv_batched(fl_matrix *A, int kernel, int kind, const void *x, void *y, bool assign, hipStream_t s,
                 void *pb_bins, const PbFold *fold = nullptr, int stream = -1);
// Bytes a WIN/WINR/PB SpMV reads once: values (+ 16-bit slots unless packed), and for PB the
// products and their 16-bit rows read back by phase 2.
long long batched_stream_bytes(const fl_matrix *A, int kernel, int kind);

// Infinity Cache capacity of the MI355X. A schedule that streams more than this per SpMV loads
// its read-once streams with the streaming hint (ld_stream): they cannot stay cached between
// calls anyway, and the hint keeps what is reused (x, products still to be read, the HYB
// heavy part's y) from being evicted. A smaller matrix stays in the Infinity Cache from one call
// to the next with ordinary loads (the 55.7 M-entry stencil with packed values: 0.064 ms
// against 0.069 ms with the hint).
constexpr long long kInfinityCacheBytes = 256LL << 20;
// fl_set_stream_policy: FL_STREAM_AUTO (by size), FL_STREAM_NEVER or FL_STREAM_ALWAYS
int stream_policy();
inline bool stream_reads(long long bytes) {
    const int p = stream_policy();
    return p == FL_STREAM_AUTO ? bytes > kInfinityCacheBytes : p == FL_STREAM_ALWAYS;
}

// HYB schedule (fl_hyb.hip). Heavy x is staged in LDS: at most kHybMaxH32 FP32 or kHybMaxH64
// FP64 columns.
constexpr int kHybMaxH32 = 38912, kHybMaxH64 = 19456;
constexpr int kHybWinLog = 14, kHybWin = 1 << kHybWinLog;  // rows of a sorting window
void hyb_prepare(fl_matrix *A, int H, int bt, int R, hipStream_t s);
void hyb_values(fl_matrix *A, int kind, hipStream_t s);
int hyb_check(const fl_matrix *A, int kind);
// scratch: nullptr (the matrix's own buffers, allocated on first use) or hyb_scratch_bytes bytes
size_t hyb_scratch_bytes(const fl_matrix *A, int kind);
int spmv_hyb(fl_matrix *A, int kind, const void *x, void *y, bool assign, hipStream_t s, void *scratch);

}  // namespace fl
