// common.hpp -- host-side helpers for libhornet_hip (gfx950, wave64); the
// device building blocks live in include/hornet/detail/wave.hpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <cstdio>

#include "hornet_hip.h"
#include "hornet/detail/wave.hpp"

namespace hn {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);

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

inline int ceil_log2_host(int64_t x) {
    int k = 0;
    while ((int64_t(1) << k) < x) k++;
    return k;
}

// Stable LSD radix sort of 64-bit keys over bits [0, key_bits).  The result
// ends in *keys (the two buffers may be swapped).  tmp from sort_temp_bytes().
size_t sort_temp_bytes(int64_t n);
hn_status radix_sort_u64(uint64_t** keys, uint64_t** alt, int64_t n, int key_bits, void* tmp,
                         hipStream_t st);
// The same sort carrying a 32-bit value per key (vals / valt swap with the
// keys): the values of equal keys keep their input order.
hn_status radix_sort_u64_kv(uint64_t** keys, uint64_t** alt, uint32_t** vals, uint32_t** valt, int64_t n,
                            int key_bits, void* tmp, hipStream_t st);

}  // namespace hn
