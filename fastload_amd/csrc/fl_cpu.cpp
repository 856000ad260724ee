// Host FastLoad SpMV: fl_FastLoad_cpu / fl_FastLoad_cpu_f32 (replace FastLoad_cpu,
// src/FastLoad_CPU.h:2-46).
//
// The reference loop visits the tiles in tile order and, for every element j of tile t, adds
// sortval[j] * x[tile_colidx[t] + (j - tile_ptr[t]) % tile_len[t]] to y[sortrowidx[j]]
// (x = sortx). With one thread this file does exactly that, in the same order and with the same
// roundings (built with -ffp-contract=off: no fused multiply-add, as the reference's host build),
// so its y is bit-identical to the reference's. With several threads the tiles are split into
// contiguous ranges of about equal element counts and y is updated with atomic adds: the sum is
// the same up to the order of additions (bit-identical on integer-valued data).
//
// The layout is validated before use (tile_ptr monotone and ending at nnz, every tile's element
// count equal to its length, columns inside [0, colA)); row indices are checked inside the loop
// and an out-of-range row is skipped and reported as FL_ERR_RANGE (y then holds the other
// elements' contributions). The reference indexes without checks.
#include <sched.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/fastload.h"

namespace fl {
void set_error(const char *fmt, ...);
}

namespace {

int validate(const slide_matrix *M, int nnz, int rowA, int colA, const void *x, const void *y, const void *val) {
    if (!M || !x || !y) {
        fl::set_error("fl_FastLoad_cpu: NULL pointer");
        return FL_ERR_INVALID;
    }
    const int T = M->tilenum;
    if (T < 0 || nnz < 0 || rowA < 0 || colA < 0) {
        fl::set_error("fl_FastLoad_cpu: negative size");
        return FL_ERR_INVALID;
    }
    if (T == 0) {
        if (nnz != 0) {
            fl::set_error("fl_FastLoad_cpu: no tiles but nnz = %d", nnz);
            return FL_ERR_INVALID;
        }
        return FL_OK;
    }
    if (!M->tile_ptr || !M->tile_len || !M->tile_colidx || !M->sortrowidx || !val) {
        fl::set_error("fl_FastLoad_cpu: slide_matrix arrays missing");
        return FL_ERR_INVALID;
    }
    if (M->tile_ptr[0] != 0 || M->tile_ptr[T] != nnz) {
        fl::set_error("fl_FastLoad_cpu: tile_ptr must run from 0 to nnz");
        return FL_ERR_INVALID;
    }
    for (int t = 0; t < T; ++t) {
        const int len = M->tile_len[t], c = M->tile_colidx[t];
        if (M->tile_ptr[t + 1] - M->tile_ptr[t] != len || len < 0 || c < 0 || (long long)c + len > colA) {
            fl::set_error("fl_FastLoad_cpu: tile %d metadata invalid", t);
            return FL_ERR_INVALID;
        }
    }
    return FL_OK;
}

int thread_count(int threads) {
    if (threads > 0) return threads;
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
    return std::max(1u, std::thread::hardware_concurrency());
}

inline void atomic_add(double *p, double v) {
    unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
    unsigned long long old = __atomic_load_n(q, __ATOMIC_RELAXED), nw;
    do {
        double d;
        std::memcpy(&d, &old, 8);
        d += v;
        std::memcpy(&nw, &d, 8);
    } while (!__atomic_compare_exchange_n(q, &old, nw, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

inline void atomic_add(float *p, float v) {
    unsigned *q = reinterpret_cast<unsigned *>(p);
    unsigned old = __atomic_load_n(q, __ATOMIC_RELAXED), nw;
    do {
        float d;
        std::memcpy(&d, &old, 4);
        d += v;
        std::memcpy(&nw, &d, 4);
    } while (!__atomic_compare_exchange_n(q, &old, nw, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

// Tiles [t0, t1): returns the number of elements whose row lies outside [0, rowA).
template <typename T, bool ATOMIC>
long long tile_range(const slide_matrix *M, const T *val, int t0, int t1, int rowA, const T *x, T *y) {
    const int *tile_ptr = M->tile_ptr, *tile_len = M->tile_len, *tile_colidx = M->tile_colidx;
    const int *row = M->sortrowidx;
    long long bad = 0;
    for (int t = t0; t < t1; ++t) {
        const int start = tile_ptr[t], stop = tile_ptr[t + 1];
        const T *xt = x + tile_colidx[t];  // (j - start) % len == j - start: stop - start == len
        (void)tile_len;
        for (int j = start; j < stop; ++j) {
            const int r = row[j];
            if ((unsigned)r >= (unsigned)rowA) {
                ++bad;
                continue;
            }
            const T p = val[j] * xt[j - start];
            if constexpr (ATOMIC) atomic_add(&y[r], p);
            else y[r] += p;
        }
    }
    return bad;
}

template <typename T>
int run(const slide_matrix *M, const T *val, int nnz, int rowA, const T *x, T *y, int threads) {
    const int T_ = M->tilenum;
    if (T_ == 0) return FL_OK;
    const int nt = std::min(thread_count(threads), std::max(1, nnz / 4096));
    long long bad = 0;
    if (nt <= 1) {
        bad = tile_range<T, false>(M, val, 0, T_, rowA, x, y);
    } else {
        // contiguous tile ranges of about nnz / nt elements each
        std::vector<int> cut(nt + 1, T_);
        cut[0] = 0;
        for (int k = 1; k < nt; ++k) {
            const long long target = (long long)nnz * k / nt;
            cut[k] = (int)(std::lower_bound(M->tile_ptr, M->tile_ptr + T_ + 1, (int)target) - M->tile_ptr);
            cut[k] = std::min(std::max(cut[k], cut[k - 1]), T_);
        }
        std::vector<long long> nbad(nt, 0);
        std::vector<std::thread> pool;
        pool.reserve(nt - 1);
        for (int k = 1; k < nt; ++k)
            pool.emplace_back([&, k] { nbad[k] = tile_range<T, true>(M, val, cut[k], cut[k + 1], rowA, x, y); });
        nbad[0] = tile_range<T, true>(M, val, cut[0], cut[1], rowA, x, y);
        for (auto &th : pool) th.join();
        for (long long b : nbad) bad += b;
    }
    if (bad) {
        fl::set_error("fl_FastLoad_cpu: %lld elements have a row index outside [0, %d)", bad, rowA);
        return FL_ERR_RANGE;
    }
    return FL_OK;
}

}  // namespace

extern "C" {

int fl_FastLoad_cpu(const slide_matrix *matrix, int nnz, int rowA, int colA, const double *x, double *y,
                    const double *y_goldencsc, int threads, int *errors) {
    try {
        if (errors) *errors = 0;
        int rc = validate(matrix, nnz, rowA, colA, x, y, matrix ? matrix->sortval : nullptr);
        if (rc) return rc;
        rc = run<double>(matrix, matrix->sortval, nnz, rowA, x, y, threads);
        if (rc) return rc;
        if (!y_goldencsc) return FL_OK;
        // src/FastLoad_CPU.h:33-45: exact comparison (meaningful on integer-valued data)
        int count = 0;
        for (int i = 0; i < rowA; ++i)
            if (y_goldencsc[i] != y[i]) ++count;
        if (errors) *errors = count;
        if (count) {
            printf("Check NO PASS! error_count_check_cpu= %d \n", count);
            fflush(stdout);
            return 1;
        }
        return FL_OK;
    } catch (...) {
        fl::set_error("fl_FastLoad_cpu: thread start or allocation failed");
        return FL_ERR_ALLOC;
    }
}

int fl_FastLoad_cpu_f32(const slide_matrix *matrix, const float *sortval, int nnz, int rowA, int colA,
                        const float *x, float *y, int threads) {
    try {
        const int rc = validate(matrix, nnz, rowA, colA, x, y, sortval);
        if (rc) return rc;
        return run<float>(matrix, sortval, nnz, rowA, x, y, threads);
    } catch (...) {
        fl::set_error("fl_FastLoad_cpu_f32: thread start or allocation failed");
        return FL_ERR_ALLOC;
    }
}

}  // extern "C"
