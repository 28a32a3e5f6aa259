// graph.hip -- storage: create from CSR/COO, export, degree queries, errors.
//
// Reference counterparts:
//   Hornet(HornetInit&) / initialize  hornet/include/Core/HornetInitialize/HornetInitialize.i.cuh:86-146
//   GraphStd::COOtoCSR                xlib/src/Graph/GraphStd.cpp:173-341
//   Hornet::getCOO + flattenCOOKernel hornet/include/Core/HornetOperations/HornetQuery.i.cuh:46-65,111-151
//   Hornet::max_degree_id             HornetQuery.i.cuh:10-36
// The reference initialises with a serial host loop over every vertex (one
// BlockArrayManager::insert each) and one H2D copy per 2^23-slot BlockArray;
// here block sizes are a device scan over the degrees and the adjacency copy
// is one load-balanced kernel.
#include <algorithm>
#include <cstring>
#include <vector>
#include <mutex>

#include "graph.hpp"
#include "prims.hpp"

namespace hn {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

hn_status dmalloc(void** p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? HN_E_OOM : HN_E_HIP;
    }
    return HN_OK;
}
void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int num_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        cus = prop.multiProcessorCount;
    }
    return cus;
}

// ------------------------------------------------------------------ kernels
namespace {

struct BlockUnitsIn {
    const int64_t* off;
    __device__ __forceinline__ uint64_t operator()(int64_t v) const {
        int64_t d = off[v + 1] - off[v];
        int k = class_of((int32_t)d);
        return k < 0 ? 0ull : (1ull << (k - MIN_CLASS));
    }
};

__global__ void set_records_kernel(int32_t nv, const int64_t* __restrict__ off,
                                   const uint64_t* __restrict__ uoff, hn_vrec* __restrict__ rec) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    int64_t d = off[v + 1] - off[v];
    hn_vrec r;
    r.off = (int64_t)uoff[v] * UNIT_SLOTS;
    r.deg = (int32_t)d;
    r.cap_log2 = class_of((int32_t)d);
    rec[v] = r;
}

constexpr int COPY_IPT = 8;
constexpr int COPY_TILE = BLOCK * COPY_IPT;

// Segmented copies between the blocks and a CSR (build, flatten / export,
// snapshots).  tile_lo[t] is the segment holding item t * COPY_TILE of the
// prefix P (one search per tile up front, so a tile finds its segments without
// block-wide searches); a tile stages up to COPY_CAP segments in LDS, locates
// all of its items, issues all loads (unconditionally: a lane past the end
// reads slot 0), then all stores.
constexpr int COPY_CAP_IPT = 1;
constexpr int COPY_CAP = LbShared<COPY_CAP_IPT, int64_t>::CAP;

__global__ void tile_starts_kernel(int64_t ntiles, const int64_t* __restrict__ P, int64_t nseg,
                                   uint32_t* __restrict__ tile_lo) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles) tile_lo[t] = (uint32_t)upper_bound_m1<int64_t>(P, 0, nseg + 1, t * COPY_TILE);
}

// A tile's items [w0, wend) as (segment, block slot): base[s] + w is item w's
// slot in its vertex block.
struct SegCopy {
    int64_t seg[COPY_IPT];
    int64_t slot[COPY_IPT];
};

template <typename BaseFn>
__device__ __forceinline__ SegCopy seg_locate(const int64_t* __restrict__ P, int64_t nseg, int64_t ntiles,
                                              const uint32_t* __restrict__ tile_lo, int64_t t, int64_t w0,
                                              int64_t wend, LbShared<COPY_CAP_IPT, int64_t>& sh, BaseFn base) {
    const int64_t lo = tile_lo[t];
    // the segment holding the next tile's first item bounds this tile's
    // segments; for the last tile, every segment from lo on
    const int64_t hi = t + 1 < ntiles ? (int64_t)tile_lo[t + 1] : nseg - 1;
    const int64_t span = hi - lo + 1;
    const int count = span <= COPY_CAP ? (int)span : -1;
    if (count > 0)
        for (int i = threadIdx.x; i < count; i += BLOCK) {
            sh.pref[i] = P[lo + i];
            sh.base[i] = base(lo + i);
        }
    __syncthreads();
    SegCopy c;
#pragma unroll
    for (int j = 0; j < COPY_IPT; j++) {
        const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
        c.seg[j] = -1;
        c.slot[j] = 0;
        if (w < wend) {
            if (count > 0) {
                const int s = lb_locate<COPY_CAP_IPT>(sh, count, w);
                c.seg[j] = lo + s;
                c.slot[j] = sh.base[s] + w;
            } else {
                const int64_t s = upper_bound_m1<int64_t>(P, lo, hi + 1, w);
                c.seg[j] = s;
                c.slot[j] = base(s) + w;
            }
        }
    }
    return c;
}

// arena[rec[v].off + j] = adj[off[v] + j] for all edges
__global__ __launch_bounds__(BLOCK) void copy_in_kernel(int32_t nv, int64_t ne, const int64_t* __restrict__ off,
                                                        const uint32_t* __restrict__ tile_lo,
                                                        const int32_t* __restrict__ adj,
                                                        const hn_vrec* __restrict__ rec,
                                                        int32_t* __restrict__ arena, MetaSet am, MetaSet in) {
    __shared__ LbShared<COPY_CAP_IPT, int64_t> sh;
    const int64_t ntiles = (ne + COPY_TILE - 1) / COPY_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t w0 = t * COPY_TILE;
        const int64_t wend = (w0 + (int64_t)COPY_TILE) < ne ? (w0 + (int64_t)COPY_TILE) : ne;
        const SegCopy c = seg_locate(off, (int64_t)nv, ntiles, tile_lo, t, w0, wend, sh,
                                     [&](int64_t s) { return rec[s].off - off[s]; });
        int32_t v[COPY_IPT];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) {
            const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            v[j] = adj[c.seg[j] >= 0 ? w : 0];
        }
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++)
            if (c.seg[j] >= 0) arena[c.slot[j]] = v[j];
        if (am.any())
#pragma unroll
            for (int j = 0; j < COPY_IPT; j++)
                if (c.seg[j] >= 0) meta_put(am, c.slot[j], in, w0 + (int64_t)j * BLOCK + threadIdx.x);
        __syncthreads();
    }
}

// COO/CSR flatten: out positions from the degree prefix eoff
__global__ __launch_bounds__(BLOCK) void flatten_kernel(int32_t nv, int64_t ne, const int64_t* __restrict__ eoff,
                                                        const uint32_t* __restrict__ tile_lo,
                                                        const hn_vrec* __restrict__ rec,
                                                        const int32_t* __restrict__ arena,
                                                        int32_t* __restrict__ src,
                                                        int32_t* __restrict__ dst, MetaSet out, MetaSet am) {
    __shared__ LbShared<COPY_CAP_IPT, int64_t> sh;
    const int64_t ntiles = (ne + COPY_TILE - 1) / COPY_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t w0 = t * COPY_TILE;
        const int64_t wend = (w0 + (int64_t)COPY_TILE) < ne ? (w0 + (int64_t)COPY_TILE) : ne;
        const SegCopy c = seg_locate(eoff, (int64_t)nv, ntiles, tile_lo, t, w0, wend, sh,
                                     [&](int64_t s) { return rec[s].off - eoff[s]; });
        int32_t v[COPY_IPT];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) v[j] = arena[c.slot[j]];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) {
            if (c.seg[j] < 0) continue;
            const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            dst[w] = v[j];
            if (src) src[w] = (int32_t)c.seg[j];
        }
        if (out.any())
#pragma unroll
            for (int j = 0; j < COPY_IPT; j++)
                if (c.seg[j] >= 0) meta_put(out, w0 + (int64_t)j * BLOCK + threadIdx.x, am, c.slot[j]);
        __syncthreads();
    }
}

// does w's sorted adjacency hold u?
__device__ __forceinline__ bool adj_contains(const int32_t* __restrict__ a, int32_t n, int32_t u) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] < u) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == u;
}

// every stored (s, v) must have its reverse (v, s); flag the first that does not
__global__ __launch_bounds__(BLOCK) void symmetry_kernel(int32_t nv, int64_t ne, const int64_t* __restrict__ eoff,
                                                         const hn_vrec* __restrict__ rec,
                                                         const int32_t* __restrict__ arena,
                                                         uint32_t* __restrict__ flag) {
    __shared__ LbShared<COPY_IPT, int64_t> sh;
    for (int64_t w0 = (int64_t)blockIdx.x * COPY_TILE; w0 < ne; w0 += (int64_t)gridDim.x * COPY_TILE) {
        lb_tile<COPY_IPT>(
            eoff, (int64_t)nv, w0, ne, sh,
            [&](int64_t s) { return rec[s].off - eoff[s]; },
            [&](int64_t s, int64_t w, int64_t base) {
                const int32_t v = arena[base + w];
                const hn_vrec r = rec[v];
                if (!adj_contains(arena + r.off, r.deg, (int32_t)s)) *flag = 1u;
            });
    }
}

struct DegIn {
    const hn_vrec* rec;
    __device__ __forceinline__ int64_t operator()(int64_t v) const { return rec[v].deg; }
};

__global__ void degrees_kernel(int32_t nv, const hn_vrec* __restrict__ rec, int32_t* __restrict__ deg) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) deg[v] = rec[v].deg;
}

__global__ void max_degree_kernel(int32_t nv, const hn_vrec* __restrict__ rec,
                                  unsigned long long* __restrict__ best) {
    unsigned long long mine = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long key = ((unsigned long long)(uint32_t)rec[v].deg << 32) |
                                 (0xFFFFFFFFull - (unsigned long long)v);
        mine = key > mine ? key : mine;
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        unsigned long long o = __shfl_xor(mine, d, WAVE);
        mine = o > mine ? o : mine;
    }
    if (__lane_id() == 0) atomicMax(best, mine);
}

__global__ void init_alloc_kernel(AllocState* st, unsigned long long bump, unsigned long long cap) {
    st->bump_units = bump;
    st->cap_units = cap;
    st->saved_bump = bump;
    for (int k = 0; k < NCLASS; k++) {
        st->free_count[k] = 0;
        st->take[k] = 0;
    }
    st->error = 0;
    st->pad = 0;
    st->want_units = 0;
}

// ---- COO -> CSR on the device
struct PackArgs {
    const int32_t* src;
    const int32_t* dst;
    int64_t m;
    int sbits;
    int sym;
    int drop;
    int32_t nv;
};
constexpr uint64_t SENTINEL = ~0ull;

__global__ void pack_keys_kernel(PackArgs a, uint64_t* __restrict__ keys, unsigned* __restrict__ bad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.m) return;
    int32_t s = a.src[i], d = a.dst[i];
    if (s < 0 || s >= a.nv || d < 0 || d >= a.nv) {
        atomicOr(bad, 1u);
        keys[i] = SENTINEL;
        if (a.sym) keys[a.m + i] = SENTINEL;
        return;
    }
    bool self = s == d;
    keys[i] = (a.drop && self) ? SENTINEL : (((uint64_t)s << a.sbits) | (uint64_t)d);
    if (a.sym) keys[a.m + i] = self ? SENTINEL : (((uint64_t)d << a.sbits) | (uint64_t)s);
}

struct UniqueFlagIn {
    const uint64_t* k;
    uint64_t mask;   // sentinel pattern restricted to the sorted bits
    __device__ __forceinline__ uint32_t operator()(int64_t i) const {
        uint64_t x = k[i];
        if (x == SENTINEL || (x & mask) == mask) return 0u;
        return (i == 0 || k[i - 1] != x) ? 1u : 0u;
    }
};

__global__ void unique_write_kernel(const uint64_t* __restrict__ k, int64_t n, uint64_t mask,
                                    const uint32_t* __restrict__ pos, int sbits,
                                    uint64_t* __restrict__ ukeys, int32_t* __restrict__ adj) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t x = k[i];
    if (x == SENTINEL || (x & mask) == mask) return;
    if (i > 0 && k[i - 1] == x) return;
    uint32_t p = pos[i];
    ukeys[p] = x;
    adj[p] = (int32_t)(x & ((1ull << sbits) - 1));
}

__global__ void offsets_from_keys_kernel(const uint64_t* __restrict__ ukeys, int64_t u, int sbits,
                                         int32_t nv, int64_t* __restrict__ off) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > nv) return;
    // first compacted key with src >= v
    uint64_t target = (uint64_t)v << sbits;
    int64_t lo = 0, hi = u;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ukeys[mid] < target) lo = mid + 1; else hi = mid;
    }
    off[v] = lo;
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_graph_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&set_records_kernel)) != hipSuccess) (void)hipGetLastError();
}

hn_status read_back(hn_graph* g, const void* dptr, size_t bytes, void* host, hipStream_t st) {
    if (bytes > 4096) {
        HN_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        return HN_OK;
    }
    if (!g || !g->h_pinned) {
        HN_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        return HN_OK;
    }
    HN_HIP(hipMemcpyAsync(g->h_pinned, dptr, bytes, hipMemcpyDeviceToHost, st));
    HN_HIP(hipStreamSynchronize(st));
    std::memcpy(host, g->h_pinned, bytes);
    return HN_OK;
}

static void destroy_graph(hn_graph* g) {
    if (!g) return;
    if (g->ms && g->ms_release) g->ms_release(g->ms);
    g->ms = nullptr;
    dfree(g->rec);
    dfree(g->arena);
    for (int f = 0; f < g->meta.n; f++) dfree(g->meta.p[f]);
    dfree(g->alloc);
    dfree(g->stacks);
    dfree(g->stack_off);
    g->s_bfs.release();
    g->s_scan.release();
    g->s_upd.release();
    g->s_tc.release();
    g->s_mstage.release();
    g->s_stamp.release();
    g->s_misc.release();
    if (g->h_pinned) (void)hipHostFree(g->h_pinned);
    for (hipEvent_t& e : g->bfs_ev)
        if (e) (void)hipEventDestroy(e);
    delete g;
}

// allocate arena + allocator state for `used_units` occupied units
static hn_status setup_arena(hn_graph* g, uint64_t used_units, double slack, hipStream_t st) {
    if (slack < 1.0) slack = 2.0;
    uint64_t cap = (uint64_t)((double)used_units * slack) + 4096;
    HN_ARG(cap < (1ull << 32), "edge arena would exceed 2^32 units of 128 B");
    g->cap_units = (int64_t)cap;
    HN_TRY(dalloc(&g->arena, (size_t)cap * UNIT_SLOTS));
    for (int f = 0; f < g->meta.n; f++)
        HN_TRY(dmalloc(reinterpret_cast<void**>(&g->meta.p[f]), (size_t)cap * UNIT_SLOTS * (size_t)g->meta.bytes[f]));
    // class k stack holds at most cap >> (k - MIN_CLASS) blocks
    int64_t tot = 0;
    for (int k = 0; k < NCLASS; k++) {
        g->h_stack_off[k] = tot;
        if (k >= MIN_CLASS) tot += (int64_t)(cap >> (k - MIN_CLASS)) + 1;
    }
    HN_TRY(dalloc(&g->stacks, (size_t)tot));
    HN_TRY(dalloc(&g->stack_off, NCLASS));
    HN_HIP(hipMemcpyAsync(g->stack_off, g->h_stack_off, sizeof(g->h_stack_off), hipMemcpyHostToDevice, st));
    HN_TRY(dalloc(&g->alloc, 1));
    init_alloc_kernel<<<1, 1, 0, st>>>(g->alloc, used_units, cap);
    HN_CHECK_LAUNCH();
    return HN_OK;
}

hn_status grow_arena(hn_graph* g, uint64_t need_units, hipStream_t st) {
    AllocState a;
    HN_TRY(read_back(g, g->alloc, sizeof(AllocState), &a, st));
    need_units = std::max<uint64_t>(need_units, a.want_units);
    const uint64_t cap = (uint64_t)g->cap_units;
    uint64_t want = std::max<uint64_t>(2 * cap, a.bump_units + need_units + 4096);
    if (want > 0xFFFFFFFFull) want = 0xFFFFFFFFull;   // unit offsets are 32-bit
    if (want <= cap || want < a.bump_units + need_units) {
        set_error("edge arena at its limit of 2^32 units of 128 B");
        return HN_E_OOM;
    }
    int32_t* arena = nullptr;
    uint32_t* stacks = nullptr;
    char* meta[HN_MAX_EDGE_META] = {};
    auto fail = [&](hn_status s) {
        dfree(arena);
        dfree(stacks);
        for (int f = 0; f < g->meta.n; f++) dfree(meta[f]);
        set_error("edge arena growth to " + std::to_string(want) + " units: " + hn_last_error());
        return s;
    };
    hn_status s;
    if ((s = dalloc(&arena, (size_t)want * UNIT_SLOTS)) != HN_OK) return fail(s);
    for (int f = 0; f < g->meta.n; f++)
        if ((s = dmalloc(reinterpret_cast<void**>(&meta[f]), (size_t)want * UNIT_SLOTS * (size_t)g->meta.bytes[f])) !=
            HN_OK)
            return fail(s);
    int64_t off[NCLASS];
    int64_t tot = 0;
    for (int k = 0; k < NCLASS; k++) {
        off[k] = tot;
        if (k >= MIN_CLASS) tot += (int64_t)(want >> (k - MIN_CLASS)) + 1;
    }
    if ((s = dalloc(&stacks, (size_t)tot)) != HN_OK) return fail(s);
    // blocks below the bump pointer keep their offsets; the free stacks keep
    // their entries, at their classes' new starts
    const size_t used = (size_t)a.bump_units * UNIT_SLOTS;
    if (used) {
        HN_HIP(hipMemcpyAsync(arena, g->arena, used * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        for (int f = 0; f < g->meta.n; f++)
            HN_HIP(hipMemcpyAsync(meta[f], g->meta.p[f], used * (size_t)g->meta.bytes[f], hipMemcpyDeviceToDevice, st));
    }
    for (int k = MIN_CLASS; k < NCLASS; k++)
        if (a.free_count[k])
            HN_HIP(hipMemcpyAsync(stacks + off[k], g->stacks + g->h_stack_off[k], sizeof(uint32_t) * a.free_count[k],
                                  hipMemcpyDeviceToDevice, st));
    HN_HIP(hipMemcpyAsync(g->stack_off, off, sizeof(off), hipMemcpyHostToDevice, st));
    const unsigned long long cap64 = want;
    HN_HIP(hipMemcpyAsync(&g->alloc->cap_units, &cap64, sizeof(cap64), hipMemcpyHostToDevice, st));
    HN_HIP(hipStreamSynchronize(st));
    dfree(g->arena);
    dfree(g->stacks);
    for (int f = 0; f < g->meta.n; f++) {
        dfree(g->meta.p[f]);
        g->meta.p[f] = meta[f];
    }
    g->arena = arena;
    g->stacks = stacks;
    g->cap_units = (int64_t)want;
    std::memcpy(g->h_stack_off, off, sizeof(off));
    return HN_OK;
}

// HIP loads each source file's code object when one of its kernels is first
// used (about 2 ms per file on MI355X: a graph's first insert took 3.9 ms
// against 0.6 ms for the next).  The first graph of a process resolves all of
// them, so that no first batch, traversal or count pays it.
static void preload_code_objects() {
    static std::once_flag once;
    std::call_once(once, [] {
        preload_bfs_code();
        preload_gen_code();
        preload_graph_code();
        preload_msbfs_code();
        preload_prims_code();
        preload_tc_code();
        preload_update_code();
    });
}

static hn_status new_graph(int32_t nv, hn_graph** out) {
    preload_code_objects();
    hn_graph* g = new hn_graph();
    g->nv = nv;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&g->h_pinned), 4096, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g->h_pinned = nullptr;
        delete g;
        set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
        return e == hipErrorNoDevice ? HN_E_NO_DEVICE : HN_E_HIP;
    }
    *out = g;
    return HN_OK;
}

// Key (row << sbits) | dst of CSR entry j, its row found by binary search of
// the offsets; flags a row whose entries are not ascending (one atomic per
// wave that saw one).
__global__ void row_keys_kernel(int32_t nv, int64_t ne, const int64_t* __restrict__ off,
                                const int32_t* __restrict__ adj, int sbits, uint64_t* __restrict__ keys,
                                unsigned* __restrict__ unsorted) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool down = false;
    if (j < ne) {
        int64_t lo = 0, hi = nv;   // off[lo] <= j < off[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (off[mid] <= j) lo = mid;
            else hi = mid;
        }
        const int32_t d = adj[j];
        keys[j] = ((uint64_t)lo << sbits) | (uint64_t)(uint32_t)d;
        down = j > off[lo] && adj[j - 1] > d;
    }
    if (__ballot(down) != 0 && (threadIdx.x & (WAVE - 1)) == 0) atomicOr(unsorted, 1u);
}

// sorted (row, dst) keys back to entries; with `idx` (each key's input
// position) the meta-data fields are gathered along
__global__ void keys_to_adj_kernel(int64_t ne, const uint64_t* __restrict__ keys, uint64_t mask,
                                   int32_t* __restrict__ adj, const uint32_t* __restrict__ idx, MetaSet out,
                                   MetaSet in) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    adj[j] = (int32_t)(keys[j] & mask);
    if (idx) meta_put(out, j, in, idx[j]);
}

__global__ void iota_u32_kernel(int64_t n, uint32_t* __restrict__ a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) a[j] = (uint32_t)j;
}

// HornetInit accepts rows in any order (HornetInit.cuh:70-73) while the
// storage keeps every adjacency ascending: *out = nullptr when the rows of
// (d_off, d_adj) are already ascending, else a new device array holding each
// row sorted (a radix sort of (row, dst) keys, which keeps the rows in place).
// *out = nullptr when the rows of (d_off, d_adj) are already ascending, else a
// new device array holding each row sorted (a radix sort of (row, dst) keys,
// which keeps the rows in place); with meta-data the sort carries each
// entry's position and *out_meta receives the fields in the new order (new
// device arrays, caller frees).
static hn_status sort_rows_if_needed(int32_t nv, int64_t ne, const int64_t* d_off, const int32_t* d_adj,
                                     const MetaSet& in_meta, hipStream_t st, int32_t** out, MetaSet* out_meta) {
    *out = nullptr;
    *out_meta = meta_like(in_meta, nullptr);
    if (ne == 0 || nv == 0) return HN_OK;
    int sbits = ceil_log2_host((int64_t)nv + 1);
    if (sbits < 1) sbits = 1;
    uint64_t* ka = nullptr;
    uint64_t* kb = nullptr;
    uint32_t* ia = nullptr;
    uint32_t* ib = nullptr;
    unsigned* flag = nullptr;
    void* tmp = nullptr;
    int32_t* res = nullptr;
    MetaSet om = meta_like(in_meta, nullptr);
    hn_status s = HN_OK;
    auto cleanup = [&]() {
        dfree(ka); dfree(kb); dfree(ia); dfree(ib); dfree(flag); dfree(tmp);
    };
    auto fail = [&](hn_status e) {
        cleanup();
        dfree(res);
        for (int f = 0; f < om.n; f++) dfree(om.p[f]);
        return e;
    };
    if ((s = dalloc(&ka, (size_t)ne)) != HN_OK || (s = dalloc(&flag, 1)) != HN_OK) return fail(s);
    HN_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned), st));
    row_keys_kernel<<<(unsigned)ceil_div(ne, BLOCK), BLOCK, 0, st>>>(nv, ne, d_off, d_adj, sbits, ka, flag);
    HN_CHECK_LAUNCH();
    unsigned h = 0;
    HN_HIP(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, st));
    HN_HIP(hipStreamSynchronize(st));
    if (!h) {
        cleanup();
        return HN_OK;
    }
    if ((s = dalloc(&kb, (size_t)ne)) != HN_OK || (s = dmalloc(&tmp, sort_temp_bytes(ne))) != HN_OK ||
        (s = dalloc(&res, (size_t)ne)) != HN_OK)
        return fail(s);
    if (in_meta.any()) {
        if ((s = dalloc(&ia, (size_t)ne)) != HN_OK || (s = dalloc(&ib, (size_t)ne)) != HN_OK) return fail(s);
        for (int f = 0; f < om.n; f++)
            if ((s = dmalloc(reinterpret_cast<void**>(&om.p[f]), (size_t)ne * (size_t)om.bytes[f])) != HN_OK)
                return fail(s);
        iota_u32_kernel<<<(unsigned)ceil_div(ne, BLOCK), BLOCK, 0, st>>>(ne, ia);
        HN_CHECK_LAUNCH();
        if ((s = radix_sort_u64_kv(&ka, &kb, &ia, &ib, ne, 2 * sbits, tmp, st)) != HN_OK) return fail(s);
    } else if ((s = radix_sort_u64(&ka, &kb, ne, 2 * sbits, tmp, st)) != HN_OK) {
        return fail(s);
    }
    keys_to_adj_kernel<<<(unsigned)ceil_div(ne, BLOCK), BLOCK, 0, st>>>(ne, ka, (1ull << sbits) - 1, res, ia, om,
                                                                        in_meta);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error(std::string("row sort: ") + hipGetErrorString(e));
        return fail(HN_E_HIP);
    }
    cleanup();
    *out = res;
    *out_meta = om;
    return HN_OK;
}

hn_status graph_from_device_csr(int32_t nv, int64_t ne, const int64_t* d_off, const int32_t* d_adj,
                                double slack, hipStream_t st, hn_graph** out, const MetaSet& d_meta) {
    hn_graph* g = nullptr;
    HN_TRY(new_graph(nv, &g));
    g->meta = meta_like(d_meta, nullptr);   // layout only; setup_arena allocates
    auto fail = [&](hn_status s) { destroy_graph(g); return s; };
    g->ne = ne;
    hn_status s;
    if ((s = dalloc(&g->rec, (size_t)(nv > 0 ? nv : 1))) != HN_OK) return fail(s);
    uint64_t* uoff = nullptr;
    if ((s = dalloc(&uoff, (size_t)nv + 1)) != HN_OK) return fail(s);
    if ((s = g->s_scan.reserve(scan_temp_bytes(nv))) != HN_OK) { dfree(uoff); return fail(s); }
    if ((s = exclusive_scan<uint64_t>(BlockUnitsIn{d_off}, uoff, nv, g->s_scan.ptr, st)) != HN_OK) {
        dfree(uoff);
        return fail(s);
    }
    uint64_t used = 0;
    if ((s = read_back(g, uoff + nv, 8, &used, st)) != HN_OK) { dfree(uoff); return fail(s); }
    if ((s = setup_arena(g, used, slack, st)) != HN_OK) { dfree(uoff); return fail(s); }
    if (nv > 0) {
        set_records_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(nv, d_off, uoff, g->rec);
    }
    uint32_t* tile_lo = nullptr;
    if (ne > 0) {
        const int64_t tiles = ceil_div(ne, COPY_TILE);
        if ((s = dalloc(&tile_lo, (size_t)tiles)) != HN_OK) { dfree(uoff); return fail(s); }
        tile_starts_kernel<<<(unsigned)ceil_div(tiles, BLOCK), BLOCK, 0, st>>>(tiles, d_off, nv, tile_lo);
        const int64_t grid = std::min<int64_t>(tiles, (int64_t)num_cus() * 8);
        copy_in_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(nv, ne, d_off, tile_lo, d_adj, g->rec, g->arena, g->meta,
                                                         d_meta);
    }
    hipError_t e = hipStreamSynchronize(st);
    dfree(uoff);
    dfree(tile_lo);
    if (e != hipSuccess) {
        set_error(std::string("graph build: ") + hipGetErrorString(e));
        return fail(HN_E_HIP);
    }
    *out = g;
    return HN_OK;
}

hn_status csr_from_device_coo(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                              uint32_t flags, hipStream_t st, int64_t** d_off_out, int32_t** d_adj_out,
                              int64_t* ne_out) {
    const int sym = (flags & HN_SYMMETRIZE) ? 1 : 0;
    const int drop = (flags & HN_DROP_SELF_LOOPS) ? 1 : 0;
    const int64_t n = m * (sym ? 2 : 1);
    int sbits = ceil_log2_host((int64_t)nv + 1);
    if (sbits < 1) sbits = 1;
    const int key_bits = 2 * sbits;
    const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
    uint64_t* ka = nullptr;
    uint64_t* kb = nullptr;
    uint32_t* pos = nullptr;
    unsigned* bad = nullptr;
    void* tmp = nullptr;
    int64_t* d_off = nullptr;
    int32_t* d_adj = nullptr;
    hn_status s = HN_OK;
    auto cleanup = [&]() {
        dfree(ka); dfree(kb); dfree(pos); dfree(bad); dfree(tmp);
    };
    size_t tmpb = std::max(sort_temp_bytes(n), scan_temp_bytes(n));
    if ((s = dalloc(&ka, (size_t)n + 1)) != HN_OK || (s = dalloc(&kb, (size_t)n + 1)) != HN_OK ||
        (s = dalloc(&pos, (size_t)n + 1)) != HN_OK || (s = dalloc(&bad, 1)) != HN_OK ||
        (s = dmalloc(&tmp, tmpb)) != HN_OK) {
        cleanup();
        return s;
    }
    HN_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned), st));
    if (m > 0) {
        PackArgs a{src, dst, m, sbits, sym, drop, nv};
        pack_keys_kernel<<<(unsigned)ceil_div(m, BLOCK), BLOCK, 0, st>>>(a, ka, bad);
        HN_CHECK_LAUNCH();
    }
    if ((s = radix_sort_u64(&ka, &kb, n, key_bits, tmp, st)) != HN_OK) { cleanup(); return s; }
    if ((s = exclusive_scan<uint32_t>(UniqueFlagIn{ka, mask}, pos, n, tmp, st)) != HN_OK) {
        cleanup();
        return s;
    }
    uint32_t u32 = 0;
    unsigned hbad = 0;
    HN_HIP(hipMemcpyAsync(&u32, pos + n, 4, hipMemcpyDeviceToHost, st));
    HN_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
    HN_HIP(hipStreamSynchronize(st));
    if (hbad) {
        cleanup();
        set_error("COO contains a vertex id outside [0, nv)");
        return HN_E_INVALID_ARG;
    }
    int64_t u = u32;
    if ((s = dalloc(&d_off, (size_t)nv + 1)) != HN_OK || (s = dalloc(&d_adj, (size_t)(u > 0 ? u : 1))) != HN_OK) {
        dfree(d_off); dfree(d_adj);
        cleanup();
        return s;
    }
    // compacted keys reuse kb
    if (n > 0) {
        unique_write_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(ka, n, mask, pos, sbits, kb, d_adj);
        HN_CHECK_LAUNCH();
    }
    offsets_from_keys_kernel<<<(unsigned)ceil_div((int64_t)nv + 1, BLOCK), BLOCK, 0, st>>>(kb, u, sbits, nv, d_off);
    HN_CHECK_LAUNCH();
    hipError_t e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) {
        dfree(d_off); dfree(d_adj);
        set_error(std::string("coo->csr: ") + hipGetErrorString(e));
        return HN_E_HIP;
    }
    *d_off_out = d_off;
    *d_adj_out = d_adj;
    *ne_out = u;
    return HN_OK;
}

hn_status degree_prefix(hn_graph* g, int64_t* d_eoff, hipStream_t st) {
    HN_TRY(g->s_scan.reserve(scan_temp_bytes(g->nv)));
    return exclusive_scan<int64_t>(DegIn{g->rec}, d_eoff, g->nv, g->s_scan.ptr, st);
}

hn_status graph_is_symmetric(hn_graph* g, hipStream_t st, bool* yes) {
    // the serving loop's slot threads call this concurrently on one graph: the
    // check (which reserves the shared s_misc scratch) runs under a lock
    std::lock_guard<std::mutex> lock(g->sym_mu);
    if (g->sym == 1 || (g->sym == -1 && g->sym_version == g->version)) {
        *yes = g->sym == 1;
        return HN_OK;
    }
    uint32_t flag = 0;
    if (g->ne > 0) {
        const size_t off_bytes = ((size_t)(g->nv + 1) * 8 + 255) & ~size_t(255);
        HN_TRY(g->s_misc.reserve(off_bytes + 256));
        int64_t* eoff = g->s_misc.as<int64_t>();
        uint32_t* d_flag = reinterpret_cast<uint32_t*>(g->s_misc.as<char>() + off_bytes);
        HN_HIP(hipMemsetAsync(d_flag, 0, 4, st));
        HN_TRY(degree_prefix(g, eoff, st));
        const int64_t grid = std::min<int64_t>(ceil_div(g->ne, COPY_TILE), (int64_t)num_cus() * 16);
        symmetry_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(g->nv, g->ne, eoff, g->rec, g->arena, d_flag);
        HN_CHECK_LAUNCH();
        HN_TRY(read_back(g, d_flag, 4, &flag, st));
    }
    g->sym = flag ? -1 : 1;
    g->sym_version = g->version;
    *yes = g->sym == 1;
    return HN_OK;
}

}  // namespace hn

using namespace hn;

// ===================================================================== C ABI
extern "C" {

const char* hn_status_string(hn_status s) {
    switch (s) {
        case HN_OK: return "ok";
        case HN_E_INVALID_ARG: return "invalid argument";
        case HN_E_OOM: return "out of device memory";
        case HN_E_HIP: return "HIP runtime error";
        case HN_E_NO_DEVICE: return "no device";
        case HN_E_CAPACITY: return "output capacity too small";
        default: return "unknown status";
    }
}

const char* hn_last_error(void) { return g_last_error.c_str(); }

const char* hn_version(void) { return "hornet-hip 0.1 (gfx950)"; }

hn_status hn_device_info(int32_t* cus, int64_t* lds, int64_t* hbm, int32_t* wave) {
    int dev = 0;
    HN_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HN_HIP(hipGetDeviceProperties(&p, dev));
    if (cus) *cus = p.multiProcessorCount;
    if (lds) *lds = (int64_t)p.maxSharedMemoryPerMultiProcessor;
    if (hbm) *hbm = (int64_t)p.totalGlobalMem;
    if (wave) *wave = p.warpSize;
    return HN_OK;
}

hn_status hn_graph_create_csr(int32_t nv, int64_t ne, const void* offsets, const int32_t* edges,
                              uint32_t flags, double slack, void* stream, hn_graph** out) {
    return hn_graph_create_csr_meta(nv, ne, offsets, edges, nullptr, nullptr, flags, slack, stream, out);
}

hn_status hn_graph_create_csr_meta(int32_t nv, int64_t ne, const void* offsets, const int32_t* edges,
                                   const hn_edge_meta* layout, const void* const* meta, uint32_t flags,
                                   double slack, void* stream, hn_graph** out) {
    HN_ARG(out != nullptr, "out is null");
    *out = nullptr;
    HN_ARG(nv >= 0 && ne >= 0, "negative size");
    HN_ARG(offsets != nullptr, "offsets is null");
    HN_ARG(ne == 0 || edges != nullptr, "edges is null");
    HN_ARG(ne < (int64_t)0xFFFFFFFF, "more than 2^32-1 edges");
    MetaSet in_meta;
    if (layout && layout->nfields > 0) {
        HN_ARG(layout->nfields <= HN_MAX_EDGE_META, "more than HN_MAX_EDGE_META meta-data fields");
        HN_ARG(ne == 0 || meta != nullptr, "meta-data arrays are null");
        in_meta.n = layout->nfields;
        for (int f = 0; f < in_meta.n; f++) {
            HN_ARG(layout->bytes[f] == 4 || layout->bytes[f] == 8, "meta-data fields are 4 or 8 bytes");
            in_meta.bytes[f] = layout->bytes[f];
        }
    }
    hipStream_t st = as_stream(stream);
    const bool dev = flags & HN_DEVICE_PTRS;
    const bool o32 = flags & HN_OFFSETS_INT32;
    // validate on the host when the input is on the host
    std::vector<int64_t> hoff;
    if (!dev) {
        hoff.resize((size_t)nv + 1);
        for (int64_t v = 0; v <= nv; v++)
            hoff[v] = o32 ? (int64_t)((const int32_t*)offsets)[v] : ((const int64_t*)offsets)[v];
        HN_ARG(hoff[0] == 0 && hoff[nv] == ne, "offsets[0] must be 0 and offsets[nv] == ne");
        for (int64_t v = 0; v < nv; v++) HN_ARG(hoff[v] <= hoff[v + 1], "offsets not monotone");
        for (int64_t j = 0; j < ne; j++) HN_ARG(edges[j] >= 0 && edges[j] < nv, "edge endpoint outside [0, nv)");
    }
    int64_t* d_off = nullptr;
    int32_t* d_adj = nullptr;
    MetaSet up = meta_like(in_meta, nullptr);   // host meta-data copied to the device
    auto release = [&]() {
        dfree(d_off);
        dfree(d_adj);
        for (int f = 0; f < up.n; f++) dfree(up.p[f]);
    };
    hn_status s;
    if ((s = dalloc(&d_off, (size_t)nv + 1)) != HN_OK) return s;
    if (dev && !o32) {
        HN_HIP(hipMemcpyAsync(d_off, offsets, sizeof(int64_t) * ((size_t)nv + 1), hipMemcpyDeviceToDevice, st));
    } else if (dev && o32) {
        std::vector<int32_t> t((size_t)nv + 1);
        HN_HIP(hipMemcpyAsync(t.data(), offsets, sizeof(int32_t) * t.size(), hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        hoff.assign(t.begin(), t.end());
        HN_HIP(hipMemcpyAsync(d_off, hoff.data(), sizeof(int64_t) * hoff.size(), hipMemcpyHostToDevice, st));
    } else {
        HN_HIP(hipMemcpyAsync(d_off, hoff.data(), sizeof(int64_t) * hoff.size(), hipMemcpyHostToDevice, st));
    }
    const int32_t* adj_in = edges;
    MetaSet dmeta = in_meta;
    // a null field array reads as zeros
    for (int f = 0; f < in_meta.n; f++) dmeta.p[f] = const_cast<char*>(static_cast<const char*>(meta[f]));
    if (!dev) {
        if ((s = dalloc(&d_adj, (size_t)(ne > 0 ? ne : 1))) != HN_OK) { release(); return s; }
        if (ne > 0) HN_HIP(hipMemcpyAsync(d_adj, edges, sizeof(int32_t) * (size_t)ne, hipMemcpyHostToDevice, st));
        adj_in = d_adj;
        for (int f = 0; f < up.n; f++) {
            if (!meta[f]) continue;
            const size_t bytes = (size_t)(ne > 0 ? ne : 1) * (size_t)up.bytes[f];
            if ((s = dmalloc(reinterpret_cast<void**>(&up.p[f]), bytes)) != HN_OK) { release(); return s; }
            if (ne > 0)
                HN_HIP(hipMemcpyAsync(up.p[f], meta[f], (size_t)ne * (size_t)up.bytes[f], hipMemcpyHostToDevice, st));
            dmeta.p[f] = up.p[f];
        }
    }
    int32_t* d_sorted = nullptr;
    MetaSet sorted_meta;
    s = sort_rows_if_needed(nv, ne, d_off, adj_in, dmeta, st, &d_sorted, &sorted_meta);
    if (s == HN_OK)
        s = graph_from_device_csr(nv, ne, d_off, d_sorted ? d_sorted : adj_in, slack, st, out,
                                  d_sorted ? sorted_meta : dmeta);
    release();
    dfree(d_sorted);
    if (d_sorted)
        for (int f = 0; f < sorted_meta.n; f++) dfree(sorted_meta.p[f]);
    return s;
}

hn_status hn_graph_edge_meta(const hn_graph* g, hn_edge_meta* layout) {
    HN_ARG(g && layout, "null argument");
    layout->nfields = g->meta.n;
    for (int f = 0; f < HN_MAX_EDGE_META; f++) layout->bytes[f] = f < g->meta.n ? g->meta.bytes[f] : 0;
    return HN_OK;
}

hn_status hn_graph_create_coo(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                              uint32_t flags, double slack, void* stream, hn_graph** out) {
    HN_ARG(out != nullptr, "out is null");
    *out = nullptr;
    HN_ARG(nv >= 0 && m >= 0, "negative size");
    HN_ARG(m == 0 || (src && dst), "null COO");
    HN_ARG(flags & HN_DEVICE_PTRS, "hn_graph_create_coo takes device pointers");
    HN_ARG(m * ((flags & HN_SYMMETRIZE) ? 2 : 1) < (int64_t)0xFFFFFFFF, "COO too large");
    hipStream_t st = as_stream(stream);
    int64_t* d_off = nullptr;
    int32_t* d_adj = nullptr;
    int64_t ne = 0;
    HN_TRY(csr_from_device_coo(nv, m, src, dst, flags, st, &d_off, &d_adj, &ne));
    hn_status s = graph_from_device_csr(nv, ne, d_off, d_adj, slack, st, out);
    if (s == HN_OK && (flags & HN_SYMMETRIZE)) (*out)->sym = 1;
    dfree(d_off);
    dfree(d_adj);
    return s;
}

hn_status hn_graph_destroy(hn_graph* g) {
    if (g) {
        (void)hipDeviceSynchronize();
        destroy_graph(g);
    }
    return HN_OK;
}

hn_status hn_graph_nv(const hn_graph* g, int32_t* nv) {
    HN_ARG(g && nv, "null argument");
    *nv = g->nv;
    return HN_OK;
}

hn_status hn_graph_ne(const hn_graph* g, int64_t* ne) {
    HN_ARG(g && ne, "null argument");
    *ne = g->ne;
    return HN_OK;
}

hn_status hn_graph_device_view(const hn_graph* g, hn_dview* view) {
    HN_ARG(g && view, "null argument");
    *view = g->view();
    return HN_OK;
}

hn_status hn_graph_memory(hn_graph* g, int64_t* used_slots, int64_t* capacity_slots, void* stream) {
    HN_ARG(g, "null graph");
    AllocState a;
    HN_TRY(read_back(g, g->alloc, sizeof(AllocState), &a, as_stream(stream)));
    int64_t freed = 0;
    for (int k = MIN_CLASS; k < NCLASS; k++) freed += (int64_t)a.free_count[k] << k;
    if (used_slots) *used_slots = (int64_t)a.bump_units * UNIT_SLOTS - freed;
    if (capacity_slots) *capacity_slots = (int64_t)a.cap_units * UNIT_SLOTS;
    return HN_OK;
}

static hn_status flatten(hn_graph* g, int32_t* src, int32_t* dst, int64_t* eoff_out, hipStream_t st,
                         const MetaSet& out = MetaSet()) {
    // s_misc: the degree prefix (unless the caller's offsets take it), then
    // the tiles' first segments
    const int64_t tiles = ceil_div(g->ne, COPY_TILE);
    const size_t eoff_bytes = eoff_out ? 0 : ((sizeof(int64_t) * ((size_t)g->nv + 1) + 255) & ~size_t(255));
    HN_TRY(g->s_misc.reserve(eoff_bytes + sizeof(uint32_t) * (size_t)(tiles + 1)));
    int64_t* eoff = eoff_out ? eoff_out : g->s_misc.as<int64_t>();
    uint32_t* tile_lo = reinterpret_cast<uint32_t*>(g->s_misc.as<char>() + eoff_bytes);
    HN_TRY(degree_prefix(g, eoff, st));
    if (g->ne > 0) {
        tile_starts_kernel<<<(unsigned)ceil_div(tiles, BLOCK), BLOCK, 0, st>>>(tiles, eoff, g->nv, tile_lo);
        HN_CHECK_LAUNCH();
        int64_t grid = std::min<int64_t>(tiles, (int64_t)num_cus() * 8);
        flatten_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(g->nv, g->ne, eoff, tile_lo, g->rec, g->arena, src, dst,
                                                         out, g->meta);
        HN_CHECK_LAUNCH();
    }
    HN_HIP(hipStreamSynchronize(st));
    return HN_OK;
}

hn_status hn_graph_export_coo(hn_graph* g, int32_t* src, int32_t* dst, int64_t cap, int64_t* n_out,
                              void* stream) {
    HN_ARG(g && src && dst, "null argument");
    if (n_out) *n_out = g->ne;
    if (cap < g->ne) {
        set_error("export_coo: capacity " + std::to_string(cap) + " < ne " + std::to_string(g->ne));
        return HN_E_CAPACITY;
    }
    return flatten(g, src, dst, nullptr, as_stream(stream));
}

hn_status hn_graph_export_coo_meta(hn_graph* g, int32_t* src, int32_t* dst, void* const* meta, int64_t cap,
                                   int64_t* n_out, void* stream) {
    HN_ARG(g && src && dst, "null argument");
    if (n_out) *n_out = g->ne;
    if (cap < g->ne) {
        set_error("export_coo: capacity " + std::to_string(cap) + " < ne " + std::to_string(g->ne));
        return HN_E_CAPACITY;
    }
    MetaSet out = meta_like(g->meta, meta ? reinterpret_cast<char* const*>(meta) : nullptr);
    return flatten(g, src, dst, nullptr, as_stream(stream), meta ? out : MetaSet());
}

hn_status hn_graph_export_csr(hn_graph* g, int64_t* offsets, int32_t* adj, int64_t cap, void* stream) {
    HN_ARG(g && offsets && adj, "null argument");
    if (cap < g->ne) {
        set_error("export_csr: capacity too small");
        return HN_E_CAPACITY;
    }
    return flatten(g, nullptr, adj, offsets, as_stream(stream));
}

hn_status hn_graph_export_csr_meta(hn_graph* g, int64_t* offsets, int32_t* adj, void* const* meta, int64_t cap,
                                   void* stream) {
    HN_ARG(g && offsets && adj, "null argument");
    if (cap < g->ne) {
        set_error("export_csr: capacity too small");
        return HN_E_CAPACITY;
    }
    MetaSet out = meta_like(g->meta, meta ? reinterpret_cast<char* const*>(meta) : nullptr);
    return flatten(g, nullptr, adj, offsets, as_stream(stream), meta ? out : MetaSet());
}

hn_status hn_graph_snapshot(hn_graph* g, double slack, void* stream, hn_graph** out) {
    HN_ARG(g && out, "null argument");
    *out = nullptr;
    hipStream_t st = as_stream(stream);
    int64_t* d_off = nullptr;
    int32_t* d_adj = nullptr;
    MetaSet tm = meta_like(g->meta, nullptr);
    hn_status s = dalloc(&d_off, (size_t)g->nv + 1);
    if (s == HN_OK) s = dalloc(&d_adj, (size_t)(g->ne > 0 ? g->ne : 1));
    for (int f = 0; f < tm.n && s == HN_OK; f++)
        s = dmalloc(reinterpret_cast<void**>(&tm.p[f]), (size_t)(g->ne > 0 ? g->ne : 1) * (size_t)tm.bytes[f]);
    if (s == HN_OK) s = flatten(g, nullptr, d_adj, d_off, st, tm);
    if (s == HN_OK)
        s = graph_from_device_csr(g->nv, g->ne, d_off, d_adj, slack < 1.0 ? 1.0 : slack, st, out, tm);
    dfree(d_off);
    dfree(d_adj);
    for (int f = 0; f < tm.n; f++) dfree(tm.p[f]);
    if (s == HN_OK) {
        // the copy starts at version 0 with what is known of the source's symmetry
        const bool known = g->sym == 1 || (g->sym == -1 && g->sym_version == g->version);
        (*out)->sym = known ? g->sym : 0;
        (*out)->sym_version = 0;
    }
    return s;
}

hn_status hn_graph_degrees(hn_graph* g, int32_t* deg, void* stream) {
    HN_ARG(g && deg, "null argument");
    if (g->nv > 0) {
        degrees_kernel<<<(unsigned)ceil_div(g->nv, BLOCK), BLOCK, 0, as_stream(stream)>>>(g->nv, g->rec, deg);
        HN_CHECK_LAUNCH();
    }
    return HN_OK;
}

hn_status hn_graph_max_degree(hn_graph* g, int32_t* id, int32_t* deg, void* stream) {
    HN_ARG(g, "null graph");
    hipStream_t st = as_stream(stream);
    HN_TRY(g->s_misc.reserve(64));
    unsigned long long* best = g->s_misc.as<unsigned long long>();
    HN_HIP(hipMemsetAsync(best, 0, 8, st));
    if (g->nv > 0) {
        int64_t grid = std::min<int64_t>(ceil_div(g->nv, BLOCK), (int64_t)num_cus() * 8);
        max_degree_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(g->nv, g->rec, best);
        HN_CHECK_LAUNCH();
    }
    unsigned long long b = 0;
    HN_TRY(read_back(g, best, 8, &b, st));
    if (id) *id = g->nv > 0 ? (int32_t)(0xFFFFFFFFull - (b & 0xFFFFFFFFull)) : -1;
    if (deg) *deg = (int32_t)(b >> 32);
    return HN_OK;
}

}  // extern "C"
