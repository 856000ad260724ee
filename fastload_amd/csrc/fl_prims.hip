// Device primitives for the preprocessing pipeline: exclusive scan, stable LSD radix sort of
// (key, index) pairs, max reduction. These replace the reference's CUB DeviceScan calls
// (src/formatTransform_GPU.h:402-496) and thrust::sort_by_key (src/ColSort_GPU.h:101).
#include <atomic>
#include <cstdarg>
#include <vector>

#include "fl_common.h"

namespace fl {

// --------------------------------------------------------------------------------- scan
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 elements per block

__device__ __forceinline__ int block_exclusive_scan(int v, int *lds_wave, int &block_total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int incl = wave_inclusive_scan(v);
    if (lane == 63) lds_wave[w] = incl;
    __syncthreads();
    int prefix = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) {
        const int x = lds_wave[i];
        prefix += (i < w) ? x : 0;
        total += x;
    }
    block_total = total;
    return prefix + incl - v;
}

__global__ __launch_bounds__(kBlock) void k_tile_sums(const int *__restrict__ in, int count, int *__restrict__ partial) {
    __shared__ int lds_wave[kBlock / kWave];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + k;
        s += (i < count) ? in[i] : 0;
    }
    int total;
    (void)block_exclusive_scan(s, lds_wave, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const int *__restrict__ in, int *__restrict__ out, int count,
                                                       const int *__restrict__ carry_in) {
    __shared__ int lds_wave[kBlock / kWave];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + k;
        v[k] = (i < count) ? in[i] : 0;
        s += v[k];
    }
    int total;
    const int excl = block_exclusive_scan(s, lds_wave, total);
    const int carry = carry_in ? carry_in[blockIdx.x] : 0;
    int run = carry + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + k;
        if (i < count) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[count] = carry + total;
}

void exclusive_scan_i32(const int *in, int *out, int count, hipStream_t s) {
    if (count <= 0) {
        FL_HIP(hipMemsetAsync(out, 0, sizeof(int), s));
        return;
    }
    const int nb = div_up(count, kScanTile);
    if (nb == 1) {
        k_scan_tiles<<<1, kBlock, 0, s>>>(in, out, count, nullptr);
        FL_LAUNCH_CHECK();
        return;
    }
    TmpBuf<int> partial(nb, s), partial_scan((size_t)nb + 1, s);
    k_tile_sums<<<nb, kBlock, 0, s>>>(in, count, partial.p);
    FL_LAUNCH_CHECK();
    exclusive_scan_i32(partial.p, partial_scan.p, nb, s);
    k_scan_tiles<<<nb, kBlock, 0, s>>>(in, out, count, partial_scan.p);
    FL_LAUNCH_CHECK();
    // partial buffers are freed on return; hipFree synchronizes with outstanding work.
}

// ------------------------------------------------------------------------------ max-reduce
__global__ __launch_bounds__(kBlock) void k_max_i32(const int *__restrict__ in, int count, int *__restrict__ out) {
    int m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        m = max(m, in[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

int device_max_i32(const int *in, int count, hipStream_t s) {
    if (count <= 0) return 0;
    DevBuf<int> d(1);
    FL_HIP(hipMemsetAsync(d.p, 0, sizeof(int), s));
    const unsigned g = std::min<unsigned>(grid_for(count), 2048u);
    k_max_i32<<<g, kBlock, 0, s>>>(in, count, d.p);
    FL_LAUNCH_CHECK();
    int h = 0;
    FL_HIP(hipMemcpyAsync(&h, d.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return h;
}

int bits_for(uint32_t v) {
    int b = 0;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

// ------------------------------------------------------------------------ radix sort
// One 8-bit LSD pass = histogram per block -> scan (digit-major) -> stable scatter.
// Block tile: kSortRounds rounds of 256 keys in index order; inside a round the rank of a key
// among equal digits is (earlier waves' counts) + (lower lanes of its wave with that digit),
// the latter from a ballot-based match over the 8 digit bits. The scatter stages the tile in
// LDS in digit order before writing it out.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kBlock * kSortRounds;  // 4096 keys per block

__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t *__restrict__ keys, int count, int shift,
                                                       int nblocks, int *__restrict__ hist) {
    __shared__ int cnt[kRadix];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const long long i = base + (long long)r * kBlock + threadIdx.x;
        if (i < count) atomicAdd(&cnt[(keys[i] >> shift) & (kRadix - 1)], 1);
    }
    __syncthreads();
    hist[(long long)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// The tile's keys are first placed in LDS in (digit, input order) order, each at its digit's
// local offset plus its stable rank; then they are written out from LDS position i to
// global_base[d] + (i - local_base[d]): consecutive threads write consecutive positions of one
// digit's run, instead of every key going straight to one of 256 scattered buckets. The tile's
// digit counts are the differences of the scanned histogram (offsets).
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t *__restrict__ keys_in,
                                                          const int *__restrict__ vals_in,
                                                          uint32_t *__restrict__ keys_out, int *__restrict__ vals_out,
                                                          int count, int shift, int nblocks,
                                                          const int *__restrict__ offsets) {
    static_assert(kRadix == kBlock, "one digit per thread");
    __shared__ uint32_t skey[kSortTile];
    __shared__ int sval[kSortTile];
    __shared__ int lbase[kRadix], gbase[kRadix], run[kRadix];
    __shared__ int wsum[kBlock / kWave];
    __shared__ int wcnt[kBlock / kWave][kRadix];  // tagged with the round: (round << 8) | count
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    {
        const long long idx = (long long)threadIdx.x * nblocks + blockIdx.x;
        const int g = offsets[idx];
        const int c = offsets[idx + 1] - g;  // keys of digit threadIdx.x in this tile
        gbase[threadIdx.x] = g;
        // exclusive scan of the 256 counts: inside each wave, then across the 4 waves
        int v = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane == 63) wsum[w] = v;
        __syncthreads();
        int before = 0;
        for (int q = 0; q < w; ++q) before += wsum[q];
        lbase[threadIdx.x] = before + v - c;
        run[threadIdx.x] = before + v - c;
    }
    for (int q = 0; q < kBlock / kWave; ++q) wcnt[q][threadIdx.x] = -1;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long long tile = (long long)blockIdx.x * kSortTile;
    const int n_tile = (int)min((long long)kSortTile, (long long)count - tile);
    for (int r = 0; r < kSortRounds; ++r) {
        const long long i = tile + (long long)r * kBlock + threadIdx.x;
        const bool valid = i < count;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const int val = valid ? vals_in[i] : 0;
        const int d = (int)((key >> shift) & (kRadix - 1));
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = ballot64(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt_mask);
        const int cnt = __popcll(peers);
        const bool leader = valid && rank == 0;
        if (leader) wcnt[w][d] = (r << 8) | cnt;
        __syncthreads();
        if (valid) {
            int pos = run[d] + rank;
            for (int q = 0; q < w; ++q) {
                const int e = wcnt[q][d];
                pos += ((e >> 8) == r) ? (e & 0xff) : 0;
            }
            skey[pos] = key;
            sval[pos] = val;
        }
        __syncthreads();
        if (leader) atomicAdd(&run[d], cnt);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n_tile; i += kBlock) {
        const uint32_t key = skey[i];
        const int d = (int)((key >> shift) & (kRadix - 1));
        const int g = gbase[d] + (i - lbase[d]);
        keys_out[g] = key;
        vals_out[g] = sval[i];
    }
}

void radix_sort_pairs(uint32_t *keys, int *vals, int count, int bits, hipStream_t s) {
    if (count <= 1 || bits <= 0) return;
    const int nb = div_up(count, kSortTile);
    const long long hcount = (long long)nb * kRadix;
    if (hcount > 0x7fffffffLL) {
        set_error("radix sort: histogram too large");
        throw HipError{hipErrorInvalidValue};
    }
    TmpBuf<int> hist((size_t)hcount, s), offs((size_t)hcount + 1, s);
    TmpBuf<uint32_t> k2(count, s);
    TmpBuf<int> v2(count, s);
    uint32_t *kin = keys, *kout = k2.p;
    int *vin = vals, *vout = v2.p;
    const int passes = (bits + kRadixBits - 1) / kRadixBits;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        k_radix_hist<<<nb, kBlock, 0, s>>>(kin, count, shift, nb, hist.p);
        FL_LAUNCH_CHECK();
        exclusive_scan_i32(hist.p, offs.p, (int)hcount, s);
        k_radix_scatter<<<nb, kBlock, 0, s>>>(kin, vin, kout, vout, count, shift, nb, offs.p);
        FL_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        FL_HIP(hipMemcpyAsync(keys, kin, sizeof(uint32_t) * count, hipMemcpyDeviceToDevice, s));
        FL_HIP(hipMemcpyAsync(vals, vin, sizeof(int) * count, hipMemcpyDeviceToDevice, s));
    }
    // no wait here: the temporaries are freed in stream order, and the caller's later work on s
    // (or its own readback) comes after the sort
}

// ------------------------------------------------------------------ temporaries' memory pool
void tmp_pool_init() {
    static std::atomic<unsigned long long> done{0};  // one bit per device (up to 64)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    const unsigned long long bit = 1ull << dev;
    if (done.load(std::memory_order_acquire) & bit) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = kTmpPoolKeep;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    done.fetch_or(bit, std::memory_order_acq_rel);
}

// ------------------------------------------------------------------------------- errors
static thread_local char g_err[512] = {0};

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

}  // namespace fl

extern "C" const char *fl_last_error(void) { return fl::g_err; }
