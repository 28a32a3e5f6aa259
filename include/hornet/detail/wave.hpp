// hornet/detail/wave.hpp -- wave64 device building blocks shared by
// libhornet_hip and the header-only operator layer (forAllEdges & co.).
//
// * wave / block scans over 64-lane wavefronts (DPP row operations);
// * exclusive_scan_dev(): reduce-then-scan over an input functor (two
//   launches up to SCAN_SELF_TILES tiles, three above), the replacement for
//   cub::DeviceScan::ExclusiveSum
//   (xlib/src/Device/Primitives/CubWrapper.cu:583-617);
// * lb_tile(): edge-granular load balancing for one workgroup tile, the
//   wave64 redesign of xlib::binarySearchLB
//   (xlib/include/Device/Primitives/impl/BinarySearchLB.i.cuh:227-332): map
//   work item w to (segment s, offset w - P[s]) with s the largest index with
//   P[s] <= w, skipping empty segments.  The tile's segment boundaries and
//   per-segment bases are staged once in LDS, items are striped across lanes
//   so consecutive lanes touch consecutive edges of the same segment, the
//   tile's segment range is found by the whole workgroup (a 256-ary search,
//   one parallel load per round), and a tile whose range overflows LDS falls
//   back to a global binary search.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hn {

constexpr int WAVE = 64;          // CDNA wavefront width
constexpr int BLOCK = 256;        // 4 waves per workgroup
constexpr int NWAVES = BLOCK / WAVE;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Move v across lanes with one DPP row operation: lanes whose source lies
// outside their row (row_shr) or whose row is masked off (row_bcast) get 0.
// 8-byte values move as two 4-byte halves from the same source lane.
template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ T dpp_move(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte integers");
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)v;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
    }
}

// inclusive wave scan (64 lanes): row_shr 1, 2, 4, 8 scan each 16-lane row;
// row_bcast:15 adds row r's total to row r+1 (rows 1 and 3), row_bcast:31
// adds lane 31's running total to rows 2 and 3
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    v += dpp_move<0x111, 0xf>(v);   // row_shr:1
    v += dpp_move<0x112, 0xf>(v);   // row_shr:2
    v += dpp_move<0x114, 0xf>(v);   // row_shr:4
    v += dpp_move<0x118, 0xf>(v);   // row_shr:8
    v += dpp_move<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
    v += dpp_move<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3
    return v;
}

// wave sum, uniform across the wave (lane 63 of the inclusive scan)
template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
    const T inc = wave_inclusive_scan(v);
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)inc;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, WAVE - 1);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), WAVE - 1);
        return (T)(((uint64_t)hi << 32) | lo);
    } else {
        return (T)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
    }
}

// exclusive block scan for BLOCK threads; returns the exclusive prefix and the
// block sum in *total.  `lds` must hold NWAVES+1 T.
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

// ------------------------------------------------------------------ scan
constexpr int SCAN_IPT = 16;                   // values per lane per tile
constexpr int SCAN_TILE = BLOCK * SCAN_IPT;    // 4096
constexpr int SCAN_WAVE_CHUNK = WAVE * SCAN_IPT;

// up to this many tiles a down-sweep tile sums the totals before it itself
// (at most SCAN_SELF_TILES loads each) instead of a separate partials launch
constexpr int64_t SCAN_SELF_TILES = 1024;

inline size_t scan_temp_bytes(int64_t n) {
    int64_t tiles = ceil_div(n > 0 ? n : 1, SCAN_TILE);
    return (size_t)(tiles + 1) * 8 + 256;
}

// Element count of a scan: a value fixed at launch, or a 32-bit word in
// device memory read by every kernel (launched for an upper bound; tiles past
// the count exit at once), so a scan can sit inside a captured hipGraph whose
// sizes change from one replay to the next.
struct HostCount {
    int64_t n;
    __device__ __forceinline__ int64_t get() const { return n; }
};
struct DevCount {
    const uint32_t* p;
    __device__ __forceinline__ int64_t get() const { return (int64_t)*p; }
};

template <typename Out, typename InFn, typename Cnt>
__global__ __launch_bounds__(BLOCK) void scan_reduce_kernel(InFn in, Cnt cnt, Out* __restrict__ partial) {
    __shared__ Out lds[NWAVES];
    const int64_t n = cnt.get();
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    if (base >= n && blockIdx.x > 0) return;
    Out acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        int64_t i = base + (int64_t)j * BLOCK + threadIdx.x;
        if (i < n) acc += (Out)in(i);
    }
    acc = wave_reduce_sum(acc);
    if (__lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Out t = 0;
        for (int w = 0; w < NWAVES; w++) t += lds[w];
        partial[blockIdx.x] = t;
    }
}

// one workgroup: exclusive scan of `count` partials in place, total -> *total
template <typename Out, typename Cnt>
__global__ __launch_bounds__(BLOCK) void scan_partials_kernel(Out* __restrict__ partial, Cnt cnt,
                                                              Out* __restrict__ total) {
    __shared__ Out lds[NWAVES + 1];
    const int64_t n = cnt.get();
    const int64_t count = n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
    Out carry = 0;
    for (int64_t b = 0; b < count; b += BLOCK) {
        int64_t i = b + threadIdx.x;
        Out v = i < count ? partial[i] : (Out)0;
        Out tot;
        Out ex = block_exclusive_scan(v, lds, &tot);
        if (i < count) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// down-sweep that hands (index, exclusive prefix, value) to a functor.  With
// SELF the tile adds up the totals of the tiles before it itself (used when
// there are few tiles, saving the partials launch), and the last tile writes
// the grand total.
template <typename Out, typename InFn, typename OutFn, typename Cnt, bool SELF = false>
__global__ __launch_bounds__(BLOCK) void scan_down_fn_kernel(InFn in, Cnt cnt, const Out* __restrict__ partial,
                                                             OutFn out, Out* __restrict__ total = nullptr) {
    __shared__ Out wtot[NWAVES];
    const int64_t n = cnt.get();
    const int w = threadIdx.x / WAVE;
    const int l = __lane_id();
    Out toff = 0;   // prefix of the tile totals before this tile
    if constexpr (SELF) {
        const int64_t last = n > 0 ? (n - 1) / SCAN_TILE : 0;
        if ((int64_t)blockIdx.x > last) return;
        Out s = 0;
        for (int64_t k = threadIdx.x; k < (int64_t)blockIdx.x; k += BLOCK) s += partial[k];
        s = wave_reduce_sum(s);
        if (l == 0) wtot[w] = s;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NWAVES; q++) toff += wtot[q];
        if ((int64_t)blockIdx.x == last && threadIdx.x == 0) *total = toff + (n > 0 ? partial[last] : (Out)0);
        __syncthreads();   // wtot is reused below
        if (n <= 0) return;
    } else {
        if ((int64_t)blockIdx.x * SCAN_TILE >= n) return;
    }
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)w * SCAN_WAVE_CHUNK;
    Out pre[SCAN_IPT];
    Out val[SCAN_IPT];
    Out carry = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        int64_t i = base + (int64_t)j * WAVE + l;
        Out v = i < n ? (Out)in(i) : (Out)0;
        Out inc = wave_inclusive_scan(v);
        pre[j] = carry + inc - v;
        val[j] = v;
        carry += __shfl(inc, WAVE - 1, WAVE);
    }
    if (l == 0) wtot[w] = carry;
    __syncthreads();
    Out off = SELF ? toff : partial[blockIdx.x];
    for (int k = 0; k < w; k++) off += wtot[k];
#pragma unroll
    for (int j = 0; j < SCAN_IPT; j++) {
        int64_t i = base + (int64_t)j * WAVE + l;
        if (i < n) out(i, (Out)(off + pre[j]), val[j]);
    }
}

template <typename Out>
__global__ void scan_zero_kernel(Out* out) { *out = 0; }

template <typename Out>
struct ArrayOut {
    Out* a;
    __device__ __forceinline__ void operator()(int64_t i, Out ex, Out) const { a[i] = ex; }
};

template <typename T>
struct ArrayIn {
    const T* a;
    __device__ __forceinline__ T operator()(int64_t i) const { return a[i]; }
};

// Exclusive scan of in(0..n) delivered to out(i, prefix, value); *total = sum.
// tmp: scan_temp_bytes(n) bytes of device memory.
template <typename Out, typename InFn, typename OutFn>
hipError_t exclusive_scan_dev(InFn in, OutFn out, Out* total, int64_t n, void* tmp, hipStream_t st) {
    if (n <= 0) {
        scan_zero_kernel<Out><<<1, 1, 0, st>>>(total);
        return hipGetLastError();
    }
    int64_t tiles = ceil_div(n, SCAN_TILE);
    Out* partial = reinterpret_cast<Out*>(tmp);
    scan_reduce_kernel<Out, InFn, HostCount><<<(unsigned)tiles, BLOCK, 0, st>>>(in, HostCount{n}, partial);
    if (tiles <= SCAN_SELF_TILES) {
        scan_down_fn_kernel<Out, InFn, OutFn, HostCount, true><<<(unsigned)tiles, BLOCK, 0, st>>>(
            in, HostCount{n}, partial, out, total);
        return hipGetLastError();
    }
    scan_partials_kernel<Out, HostCount><<<1, BLOCK, 0, st>>>(partial, HostCount{n}, total);
    scan_down_fn_kernel<Out, InFn, OutFn, HostCount><<<(unsigned)tiles, BLOCK, 0, st>>>(in, HostCount{n}, partial,
                                                                                        out);
    return hipGetLastError();
}

// Same, with the count read on the device (*n_dev <= max_n); a zero count
// leaves *total = 0.  No host round trip, so it may be captured.
template <typename Out, typename InFn, typename OutFn>
hipError_t exclusive_scan_dev_count(InFn in, OutFn out, Out* total, const uint32_t* n_dev, int64_t max_n, void* tmp,
                                    hipStream_t st) {
    const int64_t tiles = ceil_div(max_n > 0 ? max_n : 1, SCAN_TILE);
    Out* partial = reinterpret_cast<Out*>(tmp);
    scan_reduce_kernel<Out, InFn, DevCount><<<(unsigned)tiles, BLOCK, 0, st>>>(in, DevCount{n_dev}, partial);
    if (tiles <= SCAN_SELF_TILES) {
        scan_down_fn_kernel<Out, InFn, OutFn, DevCount, true><<<(unsigned)tiles, BLOCK, 0, st>>>(
            in, DevCount{n_dev}, partial, out, total);
        return hipGetLastError();
    }
    scan_partials_kernel<Out, DevCount><<<1, BLOCK, 0, st>>>(partial, DevCount{n_dev}, total);
    scan_down_fn_kernel<Out, InFn, OutFn, DevCount><<<(unsigned)tiles, BLOCK, 0, st>>>(in, DevCount{n_dev}, partial,
                                                                                       out);
    return hipGetLastError();
}

// ------------------------------------------------------------ load balancing
template <int IPT, typename PrefT, typename BaseT = int64_t>
struct LbShared {
    static constexpr int CAP = BLOCK * IPT + 1;   // segment entries staged per tile
    PrefT pref[CAP];
    BaseT base[CAP];
};

// Largest s in [lo, hi) with P[s] <= x (P[lo] <= x), found by the whole
// workgroup: each round samples BLOCK evenly spaced entries with one parallel
// load and counts the samples <= x, shrinking the range BLOCK-fold.
template <typename PrefT>
__device__ __forceinline__ int64_t block_upper_bound_m1(const PrefT* __restrict__ P, int64_t lo, int64_t hi,
                                                        int64_t x) {
    while (hi - lo > 1) {
        const int64_t len = hi - lo;
        const int64_t step = (len + BLOCK - 1) / BLOCK;
        const int64_t idx = lo + (int64_t)threadIdx.x * step;
        const int pred = idx < hi && (int64_t)P[idx] <= x;
        const int c = __syncthreads_count(pred);   // samples <= x, >= 1 since P[lo] <= x
        lo = lo + (int64_t)(c - 1) * step;
        hi = hi < lo + step ? hi : lo + step;
    }
    return lo;
}

// Items w in [w0, min(w0+BLOCK*IPT, total)) with the tile's segment range
// [lo, hi] (inclusive) known: lo holds item w0 and hi is at or past the
// segment holding the tile's last item.  item(s, w, segbase(s)).  The LDS
// stage holds LbShared<CIPT>::CAP segments (a tile with more searches them
// in global memory); a CIPT below IPT keeps the stage small.
template <int IPT, int CIPT, typename PrefT, typename SegFn, typename ItemFn>
__device__ __forceinline__ void lb_tile_range(const PrefT* __restrict__ P, int64_t lo, int64_t hi, int64_t w0,
                                              int64_t total, LbShared<CIPT, PrefT>& sh, SegFn segbase,
                                              ItemFn item) {
    constexpr int TILE = BLOCK * IPT;
    const int64_t wend = (w0 + (int64_t)TILE) < total ? (w0 + (int64_t)TILE) : total;
    const int64_t span = hi - lo + 1;
    const int count = span <= LbShared<CIPT, PrefT>::CAP ? (int)span : -(int)(span < 0x7fffffff ? span : 0x7fffffff);
    if (count > 0) {
        for (int i = threadIdx.x; i < count; i += BLOCK) {
            sh.pref[i] = P[lo + i];
            sh.base[i] = segbase(lo + i);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            if (w < wend) {
                int s = (int)upper_bound_m1<int64_t>(sh.pref, 0, count, w);
                item(lo + s, w, sh.base[s]);
            }
        }
    } else {
        const int64_t hi_excl = lo + (int64_t)(-count);
        for (int j = 0; j < IPT; j++) {
            int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            if (w < wend) {
                int64_t s = upper_bound_m1<int64_t>(P, lo, hi_excl, w);
                item(s, w, segbase(s));
            }
        }
    }
    __syncthreads();
}

// Phase-split form for kernels that want all of a lane's items in flight at
// once: lb_stage() stages [lo, hi] and returns the staged count (> 0), or
// -(count) when the range overflowed LDS; lb_locate() is then a fixed-step,
// branch-free search (independent per item, so the compiler can interleave
// the IPT searches and the loads that follow).
template <int IPT, typename PrefT, typename SegFn>
__device__ __forceinline__ int lb_stage(const PrefT* __restrict__ P, int64_t lo, int64_t hi,
                                        LbShared<IPT, PrefT>& sh, SegFn segbase) {
    const int64_t span = hi - lo + 1;
    const int count = span <= LbShared<IPT, PrefT>::CAP ? (int)span : -(int)(span < 0x7fffffff ? span : 0x7fffffff);
    if (count > 0) {
        for (int i = threadIdx.x; i < count; i += BLOCK) {
            sh.pref[i] = P[lo + i];
            sh.base[i] = segbase(lo + i);
        }
    }
    __syncthreads();
    return count;
}

constexpr int pow2_floor(int x) { return x <= 1 ? 1 : 2 * pow2_floor(x / 2); }

// largest i in [0, count) with pref[i] <= w (pref[0] <= w), log2(CAP) fixed steps
template <int IPT, typename PrefT, typename BaseT>
__device__ __forceinline__ int lb_locate(const LbShared<IPT, PrefT, BaseT>& sh, int count, int64_t w) {
    constexpr int TOP = pow2_floor(LbShared<IPT, PrefT, BaseT>::CAP - 1);
    int pos = 0;
#pragma unroll
    for (int step = TOP; step >= 1; step >>= 1) {
        const int cand = pos + step;
        const int idx = cand < count ? cand : count - 1;   // always a staged slot
        pos = (cand < count && (int64_t)sh.pref[idx] <= w) ? cand : pos;
    }
    return pos;
}

// Items w in [w0, min(w0+BLOCK*IPT, total)): item(s, w, segbase(s)) where s is
// the segment containing w (P exclusive prefix over nseg segments, P[nseg] =
// total).  Items are striped: item w0 + j*BLOCK + tid.
template <int IPT, int CIPT, typename PrefT, typename SegFn, typename ItemFn>
__device__ __forceinline__ void lb_tile(const PrefT* __restrict__ P, int64_t nseg, int64_t w0, int64_t total,
                                        LbShared<CIPT, PrefT>& sh, SegFn segbase, ItemFn item) {
    constexpr int TILE = BLOCK * IPT;
    const int64_t wend = (w0 + (int64_t)TILE) < total ? (w0 + (int64_t)TILE) : total;
    const int64_t lo = block_upper_bound_m1(P, 0, nseg + 1, w0);
    const int64_t hi = block_upper_bound_m1(P, lo, nseg + 1, wend - 1);
    lb_tile_range<IPT>(P, lo, hi, w0, total, sh, segbase, item);
}

// ------------------------------------------------------------ counter RNG
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

}  // namespace hn
