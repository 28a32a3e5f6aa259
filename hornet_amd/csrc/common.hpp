// common.hpp -- shared host/device helpers for libhornet_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <cstdio>

#include "hornet_hip.h"

namespace hn {

constexpr int WAVE = 64;          // CDNA wavefront width
constexpr int BLOCK = 256;        // 4 waves per workgroup
constexpr int NWAVES = BLOCK / WAVE;

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);

struct Status {
    hn_status code;
};

#define HN_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::hn::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return HN_E_HIP;                                                    \
        }                                                                       \
    } while (0)

#define HN_CHECK_LAUNCH()                                                       \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) {                                                 \
            ::hn::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return HN_E_HIP;                                                    \
        }                                                                       \
    } while (0)

#define HN_TRY(expr)                   \
    do {                               \
        hn_status s_ = (expr);         \
        if (s_ != HN_OK) return s_;    \
    } while (0)

#define HN_ARG(cond, msg)                          \
    do {                                           \
        if (!(cond)) {                             \
            ::hn::set_error(msg);                  \
            return HN_E_INVALID_ARG;               \
        }                                          \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// device memory helpers (status-returning)
hn_status dmalloc(void** p, size_t bytes);
template <typename T>
inline hn_status dalloc(T** p, size_t count) {
    return dmalloc(reinterpret_cast<void**>(p), count * sizeof(T) + (count == 0 ? 16 : 0));
}
void dfree(void* p);

// grow-only device scratch buffer
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    hn_status reserve(size_t need) {
        if (need <= bytes) return HN_OK;
        dfree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t want = need + need / 4 + 256;
        hn_status s = dmalloc(&ptr, want);
        if (s != HN_OK) return s;
        bytes = want;
        return HN_OK;
    }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(ptr); }
    void release() {
        dfree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

int num_cus();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int ceil_log2_host(int64_t x) {
    int k = 0;
    while ((int64_t(1) << k) < x) k++;
    return k;
}

// --------------------------------------------------------------- device side
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// inclusive wave scan (64 lanes) via DPP-free shuffles
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    int l = __lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T o = __shfl_up(v, d, WAVE);
        if (l >= d) v += o;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

// exclusive block scan for BLOCK threads; returns exclusive prefix, *total = block sum.
// `lds` must hold NWAVES+1 T.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total) {
    const int w = threadIdx.x / WAVE;
    const int l = __lane_id();
    T inc = wave_inclusive_scan(v);
    if (l == WAVE - 1) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
#pragma unroll
        for (int i = 0; i < NWAVES; i++) {
            T t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NWAVES] = run;
    }
    __syncthreads();
    T ex = inc - v + lds[w];
    *total = lds[NWAVES];
    __syncthreads();
    return ex;
}

// largest i in [lo, hi) with a[i] <= x, assuming a[lo] <= x (upper_bound - 1)
template <typename T, typename P>
__device__ __forceinline__ int64_t upper_bound_m1(const P* a, int64_t lo, int64_t hi, T x) {
    while (hi - lo > 1) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if ((T)a[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;
__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed) { return mix64(seed + GOLDEN); }
__host__ __device__ __forceinline__ uint64_t draw(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * GOLDEN);
}
__host__ __device__ __forceinline__ uint32_t uniform_below(uint64_t r, uint32_t n) {
    return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

// ------------------------------------------------------------ primitives API
// Stable LSD radix sort of 64-bit keys over bits [0, key_bits).  The result
// ends in *keys (the two buffers may be swapped).  tmp from sort_temp_bytes().
size_t sort_temp_bytes(int64_t n);
hn_status radix_sort_u64(uint64_t** keys, uint64_t** alt, int64_t n, int key_bits, void* tmp,
                         hipStream_t st);

}  // namespace hn
