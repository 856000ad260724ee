// Device-resident FastLoad matrix (internal layout; the public handle is opaque).
#pragma once

#include "fl_common.h"

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
    // FAST kernel: level index of the first tile of every 64-tile group
    fl::DevBuf<int> group_level;
};

namespace fl {

// Pipeline stages shared by the host API and the device-resident builder.
// Stable column sort: perm/sortnnz from per-column counts (device arrays of n ints).
void colsort_device(int n, const int *d_counts, int *d_perm, int *d_sortnnz, int max_key, hipStream_t s);

// Level table + tiles + data layout + classification into A (A->n, A->nnz, A->W set).
// Source of element (sorted column i, depth d):
//   tmp mode (d_P != nullptr): tmp_row/tmp_val[P[i] + d]
//   fused mode: rowidx/val[colptr[perm[i]] + d]
struct ElemSource {
    const int *P = nullptr;
    const int *colptr = nullptr;
    const int *perm = nullptr;
    const int *rowidx = nullptr;
    const double *val = nullptr;
};
void transform_device(fl_matrix *A, const int *d_sortnnz, const ElemSource &src, hipStream_t s, float *t_tran_ms,
                      float *t_clas_ms);

}  // namespace fl
