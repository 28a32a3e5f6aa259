// gen.hip -- seeded synthetic inputs on the device.
//
// The reference has no R-MAT / Chung-Lu generator and its randomisers are
// time-seeded (xlib/src/Graph/GraphStd.cpp:220-224,
// hornet/src/Util/BatchFunctions.cpp:56-58), so the generators are specified
// here, as counter-based streams that the host restatement reproduces bit for
// bit:  draw(key, c) = SplitMix64 finaliser of key + (c+1)*golden, with
// key = finaliser(seed + golden).
//   R-MAT pair i, level l: r = high32(draw(key, 64 i + l)); quadrant by integer
//   thresholds 0.57, 0.76, 0.95 of 2^32; row bit -> src bit (scale-1-l).
//   Uniform pair i: r0 = draw(3i); if i > 0 and high32(r0) % 100 == 0 the pair
//   repeats pair floor(high32(draw(3i+1)) * i / 2^32); otherwise
//   (u, v) = (uniform(draw(3i+1)), uniform(draw(3i+2))).
#include <cmath>
#include <vector>

#include "graph.hpp"

namespace hn {
namespace {

constexpr uint32_t RMAT_TA = (uint32_t)((57ull << 32) / 100);
constexpr uint32_t RMAT_TAB = (uint32_t)((76ull << 32) / 100);
constexpr uint32_t RMAT_TABC = (uint32_t)((95ull << 32) / 100);

__global__ void rmat_kernel(int scale, int64_t first, int64_t count, uint64_t key, int32_t* __restrict__ src,
                            int32_t* __restrict__ dst) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    uint64_t i = (uint64_t)(first + k);
    uint32_t s = 0, d = 0;
    for (int l = 0; l < scale; l++) {
        uint32_t r = (uint32_t)(draw(key, i * 64 + (uint64_t)l) >> 32);
        uint32_t row = r >= RMAT_TAB;
        uint32_t col = (r >= RMAT_TA && r < RMAT_TAB) || r >= RMAT_TABC;
        s |= row << (scale - 1 - l);
        d |= col << (scale - 1 - l);
    }
    src[k] = (int32_t)s;
    dst[k] = (int32_t)d;
}

// fresh pairs first; repeats resolved in a second pass (a repeat may point to
// another repeat, so resolve by chasing until a fresh pair is found)
__device__ __forceinline__ bool uniform_is_repeat(uint64_t key, uint64_t i) {
    return i > 0 && ((uint32_t)(draw(key, 3 * i) >> 32)) % 100u == 0u;
}
__device__ __forceinline__ uint64_t uniform_repeat_of(uint64_t key, uint64_t i) {
    return ((draw(key, 3 * i + 1) >> 32) * i) >> 32;
}

__global__ void uniform_kernel(int32_t n, int64_t count, uint64_t key, int32_t* __restrict__ src,
                               int32_t* __restrict__ dst) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint64_t i = (uint64_t)t;
    while (uniform_is_repeat(key, i)) i = uniform_repeat_of(key, i);
    src[t] = (int32_t)uniform_below(draw(key, 3 * i + 1), (uint32_t)n);
    dst[t] = (int32_t)uniform_below(draw(key, 3 * i + 2), (uint32_t)n);
}

// Chung-Lu endpoint: largest v with cum[v] <= high64(r * cum[n])
__device__ __forceinline__ int32_t cdf_pick(const uint64_t* __restrict__ cum, int32_t n, uint64_t r) {
    const uint64_t x = __umul64hi(r, cum[n]);
    int32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int32_t mid = lo + (hi - lo) / 2;
        if (cum[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void chung_lu_kernel(int32_t n, const uint64_t* __restrict__ cum, int64_t first, int64_t count,
                                uint64_t key, int32_t* __restrict__ src, int32_t* __restrict__ dst) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint64_t i = (uint64_t)(first + k);
    src[k] = cdf_pick(cum, n, draw(key, 2 * i));
    dst[k] = cdf_pick(cum, n, draw(key, 2 * i + 1));
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_gen_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&rmat_kernel)) != hipSuccess) (void)hipGetLastError();
}

// cum[v] = sum_{i<v} floor(2^24 * (i+1)^(-1/(gamma-1))), computed on the host
static std::vector<uint64_t> chung_lu_table(int32_t n, double gamma) {
    std::vector<uint64_t> cum((size_t)n + 1);
    cum[0] = 0;
    const double e = -1.0 / (gamma - 1.0);
    for (int32_t i = 0; i < n; i++) cum[i + 1] = cum[i] + (uint64_t)(16777216.0 * std::pow((double)i + 1.0, e));
    return cum;
}

}  // namespace hn

using namespace hn;

extern "C" {

hn_status hn_gen_rmat_pairs(int32_t scale, int64_t first, int64_t count, uint64_t seed, int32_t* src,
                            int32_t* dst, void* stream) {
    HN_ARG(scale >= 1 && scale <= 30, "scale outside [1, 30]");
    HN_ARG(count >= 0 && first >= 0, "negative range");
    HN_ARG(count == 0 || (src && dst), "null output");
    if (count == 0) return HN_OK;
    rmat_kernel<<<(unsigned)ceil_div(count, BLOCK), BLOCK, 0, as_stream(stream)>>>(
        scale, first, count, stream_key(seed), src, dst);
    HN_CHECK_LAUNCH();
    return HN_OK;
}

hn_status hn_gen_uniform_pairs(int32_t n, int64_t count, uint64_t seed, int32_t* src, int32_t* dst,
                               void* stream) {
    HN_ARG(n >= 1 && count >= 0, "bad size");
    HN_ARG(count == 0 || (src && dst), "null output");
    if (count == 0) return HN_OK;
    uniform_kernel<<<(unsigned)ceil_div(count, BLOCK), BLOCK, 0, as_stream(stream)>>>(
        n, count, stream_key(seed), src, dst);
    HN_CHECK_LAUNCH();
    return HN_OK;
}

hn_status hn_gen_chung_lu_pairs(int32_t n, double gamma, int64_t first, int64_t count, uint64_t seed,
                                int32_t* src, int32_t* dst, void* stream) {
    HN_ARG(n >= 1 && gamma > 1.0 && count >= 0 && first >= 0, "bad Chung-Lu arguments");
    HN_ARG(count == 0 || (src && dst), "null output");
    if (count == 0) return HN_OK;
    hipStream_t st = as_stream(stream);
    std::vector<uint64_t> cum = chung_lu_table(n, gamma);
    uint64_t* d_cum = nullptr;
    HN_TRY(dalloc(&d_cum, cum.size()));
    hipError_t e = hipMemcpyAsync(d_cum, cum.data(), sizeof(uint64_t) * cum.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        chung_lu_kernel<<<(unsigned)ceil_div(count, BLOCK), BLOCK, 0, st>>>(n, d_cum, first, count, stream_key(seed),
                                                                            src, dst);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(d_cum);
    if (e != hipSuccess) {
        set_error(std::string("Chung-Lu generator: ") + hipGetErrorString(e));
        return HN_E_HIP;
    }
    return HN_OK;
}

hn_status hn_gen_chung_lu_graph(int32_t n, int32_t avg_degree, double gamma, uint64_t seed, double slack,
                                void* stream, hn_graph** out) {
    HN_ARG(out, "out is null");
    *out = nullptr;
    HN_ARG(n >= 1 && avg_degree >= 1 && gamma > 1.0, "bad Chung-Lu arguments");
    const int64_t m = (int64_t)n * avg_degree / 2;
    HN_ARG(2 * m < (int64_t)0xFFFFFFFF, "Chung-Lu graph too large for 32-bit edge counts");
    hipStream_t st = as_stream(stream);
    int32_t* src = nullptr;
    int32_t* dst = nullptr;
    hn_status s;
    if ((s = dalloc(&src, (size_t)m)) != HN_OK) return s;
    if ((s = dalloc(&dst, (size_t)m)) != HN_OK) { dfree(src); return s; }
    s = hn_gen_chung_lu_pairs(n, gamma, 0, m, seed, src, dst, stream);
    if (s == HN_OK)
        s = hn_graph_create_coo(n, m, src, dst, HN_DEVICE_PTRS | HN_SYMMETRIZE | HN_DROP_SELF_LOOPS, slack, stream,
                                out);
    (void)hipStreamSynchronize(st);
    dfree(src);
    dfree(dst);
    return s;
}

hn_status hn_gen_rmat_graph(int32_t scale, int32_t edge_factor, uint64_t seed, double slack, void* stream,
                            hn_graph** out) {
    HN_ARG(out, "out is null");
    *out = nullptr;
    HN_ARG(scale >= 1 && scale <= 30 && edge_factor >= 1, "bad scale / edge factor");
    const int64_t m = (int64_t)edge_factor << scale;
    HN_ARG(2 * m < (int64_t)0xFFFFFFFF, "R-MAT graph too large for 32-bit edge counts");
    hipStream_t st = as_stream(stream);
    int32_t* src = nullptr;
    int32_t* dst = nullptr;
    hn_status s;
    if ((s = dalloc(&src, (size_t)m)) != HN_OK) return s;
    if ((s = dalloc(&dst, (size_t)m)) != HN_OK) { dfree(src); return s; }
    s = hn_gen_rmat_pairs(scale, 0, m, seed, src, dst, stream);
    if (s == HN_OK)
        s = hn_graph_create_coo(1 << scale, m, src, dst, HN_DEVICE_PTRS | HN_SYMMETRIZE | HN_DROP_SELF_LOOPS,
                                slack, stream, out);
    (void)hipStreamSynchronize(st);
    dfree(src);
    dfree(dst);
    return s;
}

}  // extern "C"
