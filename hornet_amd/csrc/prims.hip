// prims.hip -- wave64 LSD radix sort of 64-bit keys (optionally with a 32-bit
// value per key).
//
// Replaces the lexicographic (src,dst) batch sort of the reference, two stable
// thrust::sort_by_key calls (hornet/include/Core/SoA/impl/SoAPtr.i.cuh:560-571),
// with one sort of packed (src << bits | dst) keys.  Each pass: per-tile digit
// histogram in LDS, a scan of the digit-major histogram, then a stable scatter
// that ranks keys with 8 ballots per digit (one per digit bit, 64-lane
// ballots) and stages the tile in LDS in digit order so that global writes
// come out as per-bucket runs.
#include "prims.hpp"

#include <utility>

namespace hn {

// ------------------------------------------------------------------ radix sort
namespace {

constexpr int RS_BITS = 8;
constexpr int RS_BUCKETS = 1 << RS_BITS;
constexpr int RS_IPT = 16;                      // keys per lane
constexpr int RS_TILE = BLOCK * RS_IPT;         // 4096 keys
constexpr int RS_WAVE_KEYS = WAVE * RS_IPT;     // 1024 keys per wave
static_assert(RS_BUCKETS == BLOCK, "one thread per digit in the per-digit steps");

__global__ __launch_bounds__(BLOCK) void rs_hist_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                        int shift, int64_t ntiles,
                                                        uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[RS_BUCKETS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_IPT; j++) {
        int64_t i = base + (int64_t)j * BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (RS_BUCKETS - 1)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Per digit (one workgroup each): exclusive prefix of its per-tile counts
// across the tiles, and the digit's total.  The scatter adds the digit's start
// (prefix of the totals over smaller digits), so a pass needs no global scan.
__global__ __launch_bounds__(BLOCK) void rs_rowscan_kernel(const uint32_t* __restrict__ hist, int64_t ntiles,
                                                           uint32_t* __restrict__ hoff, uint32_t* __restrict__ dtot) {
    __shared__ uint32_t sc[NWAVES + 1];
    const uint32_t* row = hist + (int64_t)blockIdx.x * ntiles;
    uint32_t* out = hoff + (int64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int64_t t0 = 0; t0 < ntiles; t0 += BLOCK) {
        const int64_t t = t0 + threadIdx.x;
        const uint32_t x = t < ntiles ? row[t] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(x, sc, &tot);
        if (t < ntiles) out[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// KV: a 32-bit value rides with each key (staged beside it in LDS)
template <bool KV>
__global__ __launch_bounds__(BLOCK) void rs_scatter_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                           int shift, int64_t ntiles,
                                                           const uint32_t* __restrict__ hoff,
                                                           const uint32_t* __restrict__ dtot,
                                                           uint64_t* __restrict__ out,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ vout) {
    __shared__ uint32_t wcnt[NWAVES][RS_BUCKETS];   // per-wave digit counts
    __shared__ uint32_t dstart[RS_BUCKETS];         // tile-local start of each digit
    __shared__ uint32_t gbase[RS_BUCKETS];          // global start of this tile's digit run
    __shared__ uint64_t stage[RS_TILE];
    __shared__ uint32_t vstage[KV ? RS_TILE : 1];
    __shared__ uint32_t sc[NWAVES + 1];

    const int w = threadIdx.x / WAVE;
    const int l = __lane_id();
    for (int d = threadIdx.x; d < NWAVES * RS_BUCKETS; d += BLOCK) (&wcnt[0][0])[d] = 0;
    __syncthreads();

    const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)w * RS_WAVE_KEYS;
    uint64_t k[RS_IPT];
    uint32_t rank[RS_IPT];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < RS_IPT; j++) {
        int64_t i = base + (int64_t)j * WAVE + l;
        bool valid = i < n;
        k[j] = valid ? keys[i] : 0;
        uint32_t d = (uint32_t)(k[j] >> shift) & (RS_BUCKETS - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RS_BITS; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cntp = (uint32_t)__popcll(peers);
        uint32_t basec = valid ? wcnt[w][d] : 0;
        rank[j] = basec + before;
        // the lowest peer lane publishes the new count after every peer has read it
        if (valid && before == 0) wcnt[w][d] = basec + cntp;
    }
    __syncthreads();
    // per digit: totals over waves, exclusive across waves
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < NWAVES; q++) {
            uint32_t c = wcnt[q][d];
            wcnt[q][d] = run;
            run += c;
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(run, sc, &tot);
        dstart[d] = ex;
        // digit d starts after every key of the smaller digits
        const uint32_t dbase = block_exclusive_scan(dtot[d], sc, &tot);
        gbase[d] = dbase + hoff[(int64_t)d * ntiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_IPT; j++) {
        int64_t i = base + (int64_t)j * WAVE + l;
        if (i < n) {
            uint32_t d = (uint32_t)(k[j] >> shift) & (RS_BUCKETS - 1);
            stage[dstart[d] + wcnt[w][d] + rank[j]] = k[j];
            if constexpr (KV) vstage[dstart[d] + wcnt[w][d] + rank[j]] = vin[i];
        }
    }
    __syncthreads();
    const int64_t tile_n = min((int64_t)RS_TILE, n - (int64_t)blockIdx.x * RS_TILE);
    for (int p = threadIdx.x; p < tile_n; p += BLOCK) {
        uint64_t key = stage[p];
        uint32_t d = (uint32_t)(key >> shift) & (RS_BUCKETS - 1);
        out[(int64_t)gbase[d] + (p - dstart[d])] = key;
        if constexpr (KV) vout[(int64_t)gbase[d] + (p - dstart[d])] = vstage[p];
    }
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_prims_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&rs_hist_kernel)) != hipSuccess) (void)hipGetLastError();
}

size_t sort_temp_bytes(int64_t n) {
    int64_t tiles = ceil_div(n > 0 ? n : 1, RS_TILE);
    int64_t hn_ = tiles * RS_BUCKETS;
    return (size_t)(hn_ + 1) * 4 * 2 + scan_temp_bytes(hn_) + RS_BUCKETS * 4 + 512;
}

static hn_status radix_sort_impl(uint64_t** keys, uint64_t** alt, uint32_t** vals, uint32_t** valt, int64_t n,
                                 int key_bits, void* tmp, hipStream_t st) {
    if (n <= 1) return HN_OK;
    int64_t tiles = ceil_div(n, RS_TILE);
    int64_t hn_ = tiles * RS_BUCKETS;
    HN_ARG(hn_ < (int64_t)0xFFFFFFFF && n < (int64_t)0xFFFFFFFF, "radix_sort_u64: too many keys");
    uint32_t* hist = reinterpret_cast<uint32_t*>(tmp);
    uint32_t* hoff = hist + hn_ + 1;
    void* stmp = reinterpret_cast<char*>(tmp) + (size_t)(hn_ + 1) * 4 * 2 + 256;
    stmp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(stmp) + 255) & ~uintptr_t(255));
    uint32_t* dtot = reinterpret_cast<uint32_t*>(stmp);   // RS_BUCKETS digit totals
    for (int shift = 0; shift < key_bits; shift += RS_BITS) {
        rs_hist_kernel<<<(unsigned)tiles, BLOCK, 0, st>>>(*keys, n, shift, tiles, hist);
        HN_CHECK_LAUNCH();
        rs_rowscan_kernel<<<RS_BUCKETS, BLOCK, 0, st>>>(hist, tiles, hoff, dtot);
        HN_CHECK_LAUNCH();
        if (vals)
            rs_scatter_kernel<true><<<(unsigned)tiles, BLOCK, 0, st>>>(*keys, n, shift, tiles, hoff, dtot, *alt,
                                                                       *vals, *valt);
        else
            rs_scatter_kernel<false><<<(unsigned)tiles, BLOCK, 0, st>>>(*keys, n, shift, tiles, hoff, dtot, *alt,
                                                                        nullptr, nullptr);
        HN_CHECK_LAUNCH();
        std::swap(*keys, *alt);
        if (vals) std::swap(*vals, *valt);
    }
    return HN_OK;
}

hn_status radix_sort_u64(uint64_t** keys, uint64_t** alt, int64_t n, int key_bits, void* tmp, hipStream_t st) {
    return radix_sort_impl(keys, alt, nullptr, nullptr, n, key_bits, tmp, st);
}

hn_status radix_sort_u64_kv(uint64_t** keys, uint64_t** alt, uint32_t** vals, uint32_t** valt, int64_t n,
                            int key_bits, void* tmp, hipStream_t st) {
    return radix_sort_impl(keys, alt, vals, valt, n, key_bits, tmp, st);
}

}  // namespace hn
