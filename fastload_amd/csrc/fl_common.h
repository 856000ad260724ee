// Internal helpers shared by the FastLoad HIP translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/fastload.h"

// ---------------------------------------------------------------------------------- errors
namespace fl {

void set_error(const char *fmt, ...);

struct HipError {
    hipError_t err;
};

}  // namespace fl

#define FL_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t fl_e_ = (call);                                                                    \
        if (fl_e_ != hipSuccess) {                                                                    \
            fl::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(fl_e_)); \
            throw fl::HipError{fl_e_};                                                                \
        }                                                                                             \
    } while (0)

#define FL_LAUNCH_CHECK() FL_HIP(hipGetLastError())

// Entry-point wrapper: converts thrown HIP errors / bad_alloc into status codes.
#define FL_API_BEGIN try {
#define FL_API_END                                                         \
    }                                                                      \
    catch (const fl::HipError &) { return FL_ERR_HIP; }                   \
    catch (const std::bad_alloc &) {                                       \
        fl::set_error("host allocation failed");                           \
        return FL_ERR_ALLOC;                                               \
    }                                                                      \
    catch (...) {                                                          \
        fl::set_error("unexpected exception");                             \
        return FL_ERR_INVALID;                                             \
    }

namespace fl {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 waves

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// A load with the nontemporal (streaming) cache hint, for data a kernel reads exactly once: the
// matrix's value and index streams, and products written by the previous kernel. Those lines
// are then the first to be evicted, so the L2 and the Infinity Cache keep what is reused (x, the
// products still to be read, the heavy part's y); on the power-law workload the HYB schedule's
// three kernels take 137 instead of 146 us with it. T is a 2-, 4-, 8- or 16-byte type.
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    static_assert(sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "ld_stream: size");
    T out;
    if constexpr (sizeof(T) == 16) {
        const u32x4 r = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
        __builtin_memcpy(&out, &r, 16);
    } else if constexpr (sizeof(T) == 8) {
        const u32x2 r = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(p));
        __builtin_memcpy(&out, &r, 8);
    } else if constexpr (sizeof(T) == 4) {
        const unsigned r = __builtin_nontemporal_load(reinterpret_cast<const unsigned *>(p));
        __builtin_memcpy(&out, &r, 4);
    } else {
        const unsigned short r = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
        __builtin_memcpy(&out, &r, 2);
    }
    return out;
}
inline unsigned grid_for(long long threads, int block = kBlock) { return (unsigned)((threads + block - 1) / block); }

// RAII device buffer (hipMalloc'd, freed on scope exit).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t count = 0;
    DevBuf() = default;
    explicit DevBuf(size_t c) { alloc(c); }
    void alloc(size_t c) {
        release();
        count = c;
        if (c) FL_HIP(hipMalloc((void **)&p, c * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
    T *release_ownership() {
        T *q = p;
        p = nullptr;
        count = 0;
        return q;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Temporary device buffer of a preparation step, allocated and freed in the order of stream s
// (hipMallocAsync / hipFreeAsync from the device's default memory pool): the host neither waits
// for the device to free it nor pays a fresh hipMalloc for the next one, and the memory is
// reused by later temporaries on the same stream after the kernels that used it. The pool keeps
// up to kTmpPoolKeep bytes cached between calls (set once per device on first use).
constexpr unsigned long long kTmpPoolKeep = 8ull << 30;
void tmp_pool_init();
template <typename T>
struct TmpBuf {
    T *p = nullptr;
    size_t count = 0;
    hipStream_t s = nullptr;
    explicit TmpBuf(hipStream_t st) : s(st) {}
    TmpBuf(size_t c, hipStream_t st) : s(st) { alloc(c); }
    void alloc(size_t c) {
        release();
        count = c;
        if (c) {
            tmp_pool_init();
            FL_HIP(hipMallocAsync((void **)&p, c * sizeof(T), s));
        }
    }
    void release() {
        if (p) (void)hipFreeAsync(p, s);
        p = nullptr;
        count = 0;
    }
    ~TmpBuf() { release(); }
    TmpBuf(const TmpBuf &) = delete;
    TmpBuf &operator=(const TmpBuf &) = delete;
};

// ----------------------------------------------------------------- device-side primitives
// Exclusive prefix sum: out[0]=0, out[i]=sum(in[0..i-1]), out[count]=total. in may equal out
// only when in has count+1 entries. Returns nothing; total is on the device at out[count].
void exclusive_scan_i32(const int *in, int *out, int count, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits of key.
// keys/vals are sorted in place (ping-pong through internal scratch).
void radix_sort_pairs(uint32_t *keys, int *vals, int count, int bits, hipStream_t s);

// Device reduction helpers.
int device_max_i32(const int *in, int count, hipStream_t s);
int bits_for(uint32_t v);

}  // namespace fl

// --------------------------------------------------------------- wave64 device helpers
namespace fl {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// DPP move with compile-time control. old=0 for lanes outside row_mask or reading out of range.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(dpp_i32<CTRL, ROW_MASK>(__float_as_int(v)));
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
    int2 a = *reinterpret_cast<int2 *>(&v);
    int2 r;
    r.x = dpp_i32<CTRL, ROW_MASK>(a.x);
    r.y = dpp_i32<CTRL, ROW_MASK>(a.y);
    return *reinterpret_cast<double *>(&r);
}

template <int CTRL, int ROW_MASK, typename T>
struct Dpp;
template <int CTRL, int ROW_MASK>
struct Dpp<CTRL, ROW_MASK, float> {
    static __device__ __forceinline__ float mov(float v) { return dpp_f32<CTRL, ROW_MASK>(v); }
};
template <int CTRL, int ROW_MASK>
struct Dpp<CTRL, ROW_MASK, double> {
    static __device__ __forceinline__ double mov(double v) { return dpp_f64<CTRL, ROW_MASK>(v); }
};
template <int CTRL, int ROW_MASK>
struct Dpp<CTRL, ROW_MASK, int> {
    static __device__ __forceinline__ int mov(int v) { return dpp_i32<CTRL, ROW_MASK>(v); }
};

// DPP control codes (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// Segmented inclusive scan across the 64 lanes of a wave.
// start = index of the first lane of this lane's segment (start <= lane). On return each lane
// holds the sum of v over [start, lane]. All 64 lanes must execute this call (DPP reads other
// lanes' registers); lanes outside any tile should pass v = 0 and start = lane.
// In-row phase: row_shr 1,2,4,8 inside each 16-lane row; cross-row phase: row_bcast:15
// (lane 15 -> row 1, lane 47 -> row 3) then row_bcast:31 (lane 31 -> rows 2,3).
template <typename T>
__device__ __forceinline__ T wave_segscan(T v, int start, int lane) {
    const int lr = lane & 15;
    T t;
    t = Dpp<kRowShr1, 0xf, T>::mov(v);
    if (lr >= 1 && lane - 1 >= start) v += t;
    t = Dpp<kRowShr2, 0xf, T>::mov(v);
    if (lr >= 2 && lane - 2 >= start) v += t;
    t = Dpp<kRowShr4, 0xf, T>::mov(v);
    if (lr >= 4 && lane - 4 >= start) v += t;
    t = Dpp<kRowShr8, 0xf, T>::mov(v);
    if (lr >= 8 && lane - 8 >= start) v += t;
    t = Dpp<kRowBcast15, 0xa, T>::mov(v);
    if ((lane & 16) && start <= (lane & ~15) - 1) v += t;
    t = Dpp<kRowBcast31, 0xc, T>::mov(v);
    if ((lane & 32) && start <= 31) v += t;
    return v;
}

// Plain inclusive wave scan of ints (segment = whole wave).
__device__ __forceinline__ int wave_inclusive_scan(int v) { return wave_segscan<int>(v, 0, (int)__lane_id()); }

// Start lane of the segment containing `lane`, given the 64-bit mask of segment heads.
__device__ __forceinline__ int segment_start(uint64_t heads, int lane) {
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    return 63 - __clzll((long long)(heads & upto));
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return (uint64_t)__ballot(p); }

}  // namespace fl
