// Wave-uniform walk over 64-wide tiles using the per-level table instead of per-tile metadata.
// Level k covers sorted columns [col_k, n) at depths [base_k, base_k + tall_k); its tiles are
// ordered (chunk j, depth e) with tile index tileoff[k] + j*tall_k + e. Tile (k, j, e) holds
// len = (j == chunks-1 ? (n-col_k) - (chunks-1)*64 : 64) elements starting at
// nnzoff[k] + j*tall_k*64 + e*len, for sorted columns col_k + 64*j + lane.
#pragma once

#include "fl_common.h"

namespace fl {

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

struct LevelTab {
    const int *__restrict__ col;
    const int *__restrict__ tall;
    const int *__restrict__ tileoff;  // L+1
    const int *__restrict__ nnzoff;   // L+1
    const int *__restrict__ group_level;  // level of tile 64*q
    int n, L;
};

struct LevelWalk {
    int k, col, tall, chunks, nnzoff, j, e;

    __device__ __forceinline__ void load(const LevelTab &lt, int kk) {
        k = kk;
        col = uni(lt.col[kk]);
        tall = uni(lt.tall[kk]);
        nnzoff = uni(lt.nnzoff[kk]);
        chunks = (lt.n - col + 63) >> 6;
    }

    // Position the walk at tile t (t < T).
    __device__ __forceinline__ void seek(const LevelTab &lt, int t) {
        int kk = uni(lt.group_level[t >> 6]);
        while (kk + 1 < lt.L && uni(lt.tileoff[kk + 1]) <= t) ++kk;
        load(lt, kk);
        const int local = t - uni(lt.tileoff[kk]);
        j = local / tall;
        e = local - j * tall;
    }

    // Geometry of the current tile; then advance. `more` = another tile follows in this walk.
    __device__ __forceinline__ void next(const LevelTab &lt, int &ptr, int &cx, int &len, bool more) {
        const int width = lt.n - col;
        len = (j == chunks - 1) ? width - (chunks - 1) * 64 : 64;
        ptr = nnzoff + j * tall * 64 + e * len;
        cx = col + 64 * j;
        if (++e == tall) {
            e = 0;
            if (++j == chunks && more) {
                load(lt, k + 1);
                j = 0;
            }
        }
    }
};

}  // namespace fl
