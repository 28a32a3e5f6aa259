// bfs.hip -- level-synchronous top-down BFS (HornetsNest BfsTopDown).
//
// Reference: hornetsnest/src/Static/BreadthFirstSearch/TopDown.cu:49-106 runs,
// per level, forAllEdges(hornet, queue, BFSOperator, BinarySearch):
// computeWorkKernel + cub ExclusiveSum + D2H total + binarySearchKernel
// (primitives/LoadBalancing/BinarySearch.i.cuh:57-101), then a queue swap with
// a <<<1,1>>> kernel and another D2H (primitives/Queue/TwoLevelQueue.i.cuh:158-179).
// Its operator is a non-atomic check-then-set (TopDown.cu:54-60), so a vertex
// may be enqueued more than once; the distances are still exact.
//
// Here, per level:
//   1. a scan of deg(q[i]) read straight from the vertex records; its
//      down-sweep also writes, per frontier vertex, the adjacency base
//      (record offset minus prefix, so an edge's address is base + work index)
//      and, per 2048-edge tile, the frontier index holding the tile's first
//      edge -- the partition that BinarySearchLB recomputes with two global
//      binary searches per block (BinarySearchLB.i.cuh:254-264);
//   2. one expansion kernel: each tile stages its frontier slice in LDS, maps
//      every edge to (vertex, offset) with an LDS search, tests the neighbour
//      against the visited bitmap as of the start of the level (n/8 bytes,
//      read-only during the kernel, so it stays cache resident) and marks an
//      unvisited neighbour with an idempotent byte store.  No atomics: with
//      eight XCDs, each holding its own L2, device-scope atomics cannot be
//      served from cache, and claiming vertices with atomicOr measured 7x
//      slower;
//   3. one collect kernel: a lane owns 32 vertices (one bitmap word), turns
//      their marks into visited bits, distances and next-frontier entries
//      (one atomicAdd per wave), so every vertex enters a frontier once;
//   4. one 8-byte readback (next frontier size + this level's edge count).
// Distances: source 0, its neighbours 1, unreached INT32_MAX (TopDown.cu:42,81).
#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int EXP_IPT = 8;
constexpr int EXP_TILE = BLOCK * EXP_IPT;
constexpr int32_t INF = 0x7fffffff;

// The frontier is kept as (degree, adjacency offset) pairs written by the
// collect kernel, so the scan reads contiguous arrays instead of gathering
// one 16-byte vertex record per frontier vertex twice.
struct FrontierDegIn {
    const int32_t* qdeg;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const { return (uint32_t)qdeg[i]; }
};

// per frontier vertex i with exclusive edge prefix `ex` and degree d
struct FrontierOut {
    const int64_t* qoff;
    uint32_t* P;
    int64_t* base;
    uint32_t* tile_lo;
    __device__ __forceinline__ void operator()(int64_t i, uint32_t ex, uint32_t d) const {
        P[i] = ex;
        base[i] = qoff[i] - (int64_t)ex;
        if (d) {
            uint32_t t0 = (uint32_t)(((uint64_t)ex + EXP_TILE - 1) / EXP_TILE);
            uint32_t t1 = (uint32_t)(((uint64_t)ex + d + EXP_TILE - 1) / EXP_TILE);
            for (uint32_t t = t0; t < t1; t++) tile_lo[t] = (uint32_t)i;
        }
    }
};

__global__ void bfs_init_kernel(int32_t nv, int32_t* __restrict__ dist, uint32_t* __restrict__ bitmap,
                                int32_t source, int32_t* __restrict__ q) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t words = ((int64_t)nv + 31) / 32;
    if (i < nv) dist[i] = (i == source) ? 0 : INF;
    if (i < words) bitmap[i] = (i == (source >> 5)) ? (1u << (source & 31)) : 0u;
    if (i == 0) q[0] = source;
}

// Phase 1: every edge of the frontier tests the visited bitmap as of the
// start of the level (read-only in this kernel, so cacheable) and marks an
// unvisited neighbour in the byte array `next`.  The mark is an idempotent
// byte store -- the reference's check-then-set without the queue insert --
// so no atomics are needed and no update can be lost.
__global__ __launch_bounds__(BLOCK) void bfs_expand_kernel(int64_t F, const uint32_t* __restrict__ P,
                                                           const int64_t* __restrict__ segbase,
                                                           const uint32_t* __restrict__ tile_lo,
                                                           const int32_t* __restrict__ arena,
                                                           const uint32_t* __restrict__ bitmap,
                                                           uint8_t* __restrict__ next) {
    __shared__ LbShared<EXP_IPT, uint32_t> sh;
    const int64_t W = P[F];
    const int64_t ntiles = (W + EXP_TILE - 1) / EXP_TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t lo = tile_lo[t];
        const int64_t hi = (t + 1 < ntiles) ? (int64_t)tile_lo[t + 1] : F - 1;
        const int64_t w0 = t * EXP_TILE;
        const int count = lb_stage<EXP_IPT>(P, lo, hi, sh, [&](int64_t s) { return segbase[s]; });
        // phase 1: every item's adjacency slot (independent LDS searches)
        int64_t addr[EXP_IPT];
        if (count > 0) {
#pragma unroll
            for (int j = 0; j < EXP_IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                const int64_t a = sh.base[lb_locate<EXP_IPT>(sh, count, w)] + w;
                addr[j] = w < W ? a : -1;
            }
        } else {   // segment range wider than LDS (zero-degree frontier vertices)
            for (int j = 0; j < EXP_IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                addr[j] = w < W ? segbase[upper_bound_m1<int64_t>(P, lo, lo + (int64_t)(-count), w)] + w : -1;
            }
        }
        // phase 2: all neighbour ids in flight at once (streamed, not kept in L2)
        int32_t v[EXP_IPT];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++) v[j] = addr[j] >= 0 ? __builtin_nontemporal_load(arena + addr[j]) : -1;
        // phase 3: all visited words in flight at once
        uint32_t word[EXP_IPT];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++) word[j] = v[j] >= 0 ? bitmap[v[j] >> 5] : 0xFFFFFFFFu;
        // phase 4: idempotent marks (testing the mark first cut L2 write-backs
        // 10x but added a dependent read and measured slower)
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
            if (!(word[j] & (1u << (v[j] & 31)))) next[v[j]] = 1;
        __syncthreads();
    }
}

// (degree, adjacency offset) of every frontier vertex, one record gather each
__global__ void bfs_gather_kernel(int64_t F, const int32_t* __restrict__ q, const hn_vrec* __restrict__ rec,
                                  int32_t* __restrict__ qdeg, int64_t* __restrict__ qoff) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    const hn_vrec r = rec[q[i]];
    qdeg[i] = r.deg;
    qoff[i] = r.off;
}

// Phase 2: one lane owns 32 vertices (one bitmap word and 32 mark bytes):
// marked vertices get their distance, their visited bit and a slot in the
// next frontier (one atomicAdd per wave); marks are cleared for the next level.
__global__ __launch_bounds__(BLOCK) void bfs_collect_kernel(int64_t words, uint32_t* __restrict__ bitmap,
                                                            uint32_t* __restrict__ next4,
                                                            int32_t* __restrict__ dist, int32_t level,
                                                            int32_t* __restrict__ qn,
                                                            uint32_t* __restrict__ qn_count) {
    const int64_t wi = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t found = 0;
    if (wi < words) {
        const uint4* p = reinterpret_cast<const uint4*>(next4 + wi * 8);
        uint4 a = p[0], b = p[1];
        uint32_t m[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            uint32_t x = m[k];
            // byte j of word k is vertex 32*wi + 4k + j
            found |= ((x & 0xFFu) ? 1u : 0u) << (4 * k);
            found |= ((x & 0xFF00u) ? 1u : 0u) << (4 * k + 1);
            found |= ((x & 0xFF0000u) ? 1u : 0u) << (4 * k + 2);
            found |= ((x & 0xFF000000u) ? 1u : 0u) << (4 * k + 3);
        }
        if (found) {
            uint4 z = make_uint4(0, 0, 0, 0);
            uint4* pw = reinterpret_cast<uint4*>(next4 + wi * 8);
            pw[0] = z;
            pw[1] = z;
            bitmap[wi] |= found;
        }
    }
    const uint32_t cnt = (uint32_t)__popc(found);
    const uint32_t inc = wave_inclusive_scan(cnt);
    const uint32_t tot = __shfl(inc, WAVE - 1, WAVE);
    if (tot == 0) return;
    uint32_t b = 0;
    if (__lane_id() == WAVE - 1) b = atomicAdd(qn_count, tot);
    b = __shfl(b, WAVE - 1, WAVE) + inc - cnt;
    while (found) {
        const int j = __ffs(found) - 1;
        found &= found - 1;
        const int32_t v = (int32_t)(wi * 32 + j);
        dist[v] = level;
        qn[b++] = v;
    }
}

}  // namespace

// scratch layout inside g->s_bfs
// The frontier needs no ping-pong: q is gathered into (qdeg, qoff) and
// scanned into P / base / tile_lo before collect overwrites q with level L+1.
struct BfsBufs {
    uint32_t* bitmap;
    int32_t* q;
    int32_t* qdeg;
    int64_t* qoff;
    uint32_t* P;
    int64_t* base;
    uint32_t* tile_lo;
    uint8_t* next;   // one mark byte per vertex, padded to whole bitmap words
    uint32_t* cnt;   // [0] next frontier size
};

static hn_status bfs_buffers(hn_graph* g, BfsBufs* b) {
    const int64_t nv = g->nv > 0 ? g->nv : 1;
    const int64_t words = (nv + 31) / 32;
    const int64_t tiles = g->ne / EXP_TILE + 2;
    auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
    int64_t o_bitmap = 0;
    int64_t o_q = al(o_bitmap + words * 4);
    int64_t o_qdeg = al(o_q + nv * 4);
    int64_t o_qoff = al(o_qdeg + nv * 4);
    int64_t o_P = al(o_qoff + nv * 8);
    int64_t o_base = al(o_P + (nv + 1) * 4);
    int64_t o_tile = al(o_base + nv * 8);
    int64_t o_next = al(o_tile + tiles * 4);
    int64_t o_cnt = al(o_next + words * 32);
    int64_t total = al(o_cnt + 64);
    HN_TRY(g->s_bfs.reserve((size_t)total));
    char* base = g->s_bfs.as<char>();
    b->bitmap = reinterpret_cast<uint32_t*>(base + o_bitmap);
    b->q = reinterpret_cast<int32_t*>(base + o_q);
    b->qdeg = reinterpret_cast<int32_t*>(base + o_qdeg);
    b->qoff = reinterpret_cast<int64_t*>(base + o_qoff);
    b->P = reinterpret_cast<uint32_t*>(base + o_P);
    b->base = reinterpret_cast<int64_t*>(base + o_base);
    b->tile_lo = reinterpret_cast<uint32_t*>(base + o_tile);
    b->next = reinterpret_cast<uint8_t*>(base + o_next);
    b->cnt = reinterpret_cast<uint32_t*>(base + o_cnt);
    HN_TRY(g->s_scan.reserve(scan_temp_bytes(nv)));
    return HN_OK;
}

}  // namespace hn

using namespace hn;

extern "C" hn_status hn_bfs_topdown(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats,
                                    void* stream) {
    HN_ARG(g && dist, "null argument");
    HN_ARG(source >= 0 && source < g->nv, "source outside [0, nv)");
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds the 32-bit work prefix");
    hipStream_t st = as_stream(stream);
    BfsBufs b;
    HN_TRY(bfs_buffers(g, &b));
    const int64_t nv = g->nv;
    const int64_t words = (nv + 31) / 32;
    bfs_init_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(g->nv, dist, b.bitmap, source, b.q);
    HN_CHECK_LAUNCH();
    HN_HIP(hipMemsetAsync(b.next, 0, (size_t)words * 32, st));
    const int64_t grid_cap = (int64_t)num_cus() * 8;
    const bool timed = stats && stats->time_kernels;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (timed) {
        HN_HIP(hipEventCreate(&ev0));
        HN_HIP(hipEventCreate(&ev1));
    }
    int64_t F = 1;
    int32_t level = 1;
    int64_t visited = 0, scanned = 0, kns = 0;
    int32_t nlevels = 0;
    uint32_t* host = reinterpret_cast<uint32_t*>(g->h_pinned);
    hn_status rc = HN_OK;
    while (F > 0) {
        visited += F;
        nlevels++;
        bfs_gather_kernel<<<(unsigned)ceil_div(F, BLOCK), BLOCK, 0, st>>>(F, b.q, g->rec, b.qdeg, b.qoff);
        rc = exclusive_scan_fn<uint32_t>(FrontierDegIn{b.qdeg}, FrontierOut{b.qoff, b.P, b.base, b.tile_lo},
                                         b.P + F, F, g->s_scan.ptr, st);
        if (rc != HN_OK) break;
        hipError_t e = hipMemsetAsync(b.cnt, 0, 4, st);
        if (timed && e == hipSuccess) e = hipEventRecord(ev0, st);
        if (e == hipSuccess) {
            bfs_expand_kernel<<<(unsigned)grid_cap, BLOCK, 0, st>>>(F, b.P, b.base, b.tile_lo, g->arena,
                                                                      b.bitmap, b.next);
            e = hipGetLastError();
        }
        if (timed && e == hipSuccess) e = hipEventRecord(ev1, st);
        if (e == hipSuccess) {
            bfs_collect_kernel<<<(unsigned)ceil_div(words, BLOCK), BLOCK, 0, st>>>(
                words, b.bitmap, reinterpret_cast<uint32_t*>(b.next), dist, level, b.q, b.cnt);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(host, b.cnt, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(host + 1, b.P + F, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (timed && e == hipSuccess) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, ev0, ev1);
            kns += (int64_t)((double)ms * 1e6);
        }
        if (e != hipSuccess) {
            set_error(std::string("bfs level: ") + hipGetErrorString(e));
            rc = HN_E_HIP;
            break;
        }
        scanned += host[1];
        F = host[0];
        level++;
    }
    if (timed) {
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
    if (rc != HN_OK) return rc;
    if (stats) {
        stats->levels = nlevels;
        stats->visited = visited;
        stats->edges_scanned = scanned;
        stats->expand_ns = kns;
        stats->expand_launches = nlevels;
        stats->pad_ = 0;
    }
    return HN_OK;
}
