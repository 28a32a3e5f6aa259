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
//   1. a scan of deg(q[i]) read straight from the vertex records (no separate
//      work array pass);
//   2. one expansion kernel: the wave64 LB tile maps each edge of the frontier
//      to (vertex, offset), the neighbour is tested against a visited bitmap
//      (n/8 bytes, L2/MALL resident) and claimed with atomicOr, so each vertex
//      enters the next frontier exactly once; appends are aggregated per wave
//      with one 64-lane ballot and one atomicAdd;
//   3. one 8-byte readback (next frontier size + this level's edge count).
// Distances: source 0, its neighbours 1, unreached INT32_MAX (TopDown.cu:42,81).
#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int EXP_IPT = 8;
constexpr int EXP_TILE = BLOCK * EXP_IPT;
constexpr int32_t INF = 0x7fffffff;

struct FrontierDegIn {
    const int32_t* q;
    const hn_vrec* rec;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const { return (uint32_t)rec[q[i]].deg; }
};

__global__ void bfs_init_kernel(int32_t nv, int32_t* __restrict__ dist, uint32_t* __restrict__ bitmap,
                                int32_t source, int32_t* __restrict__ q) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t words = ((int64_t)nv + 31) / 32;
    if (i < nv) dist[i] = (i == source) ? 0 : INF;
    if (i < words) bitmap[i] = (i == (source >> 5)) ? (1u << (source & 31)) : 0u;
    if (i == 0) q[0] = source;
}

__global__ __launch_bounds__(BLOCK) void bfs_expand_kernel(const int32_t* __restrict__ q, int64_t F,
                                                           const uint32_t* __restrict__ P,
                                                           const hn_vrec* __restrict__ rec,
                                                           const int32_t* __restrict__ arena,
                                                           uint32_t* __restrict__ bitmap,
                                                           int32_t* __restrict__ dist, int32_t level,
                                                           int32_t* __restrict__ qn,
                                                           uint32_t* __restrict__ qn_count) {
    __shared__ LbShared<EXP_IPT, uint32_t> sh;
    const int64_t W = P[F];
    const uint64_t lt = lanemask_lt();
    for (int64_t w0 = (int64_t)blockIdx.x * EXP_TILE; w0 < W; w0 += (int64_t)gridDim.x * EXP_TILE) {
        lb_tile<EXP_IPT>(
            P, F, w0, W, sh,
            [&](int64_t s) { return rec[q[s]].off - (int64_t)P[s]; },
            [&](int64_t, int64_t w, int64_t base) {
                const int32_t v = arena[base + w];
                const uint32_t bit = 1u << (v & 31);
                uint32_t* word = bitmap + (v >> 5);
                bool won = false;
                if (!(*word & bit)) {
                    uint32_t old = atomicOr(word, bit);
                    won = !(old & bit);
                }
                if (won) dist[v] = level;
                const uint64_t m = __ballot(won);
                if (m) {
                    const int leader = __ffsll((unsigned long long)m) - 1;
                    uint32_t b = 0;
                    if (__lane_id() == leader) b = atomicAdd(qn_count, (uint32_t)__popcll(m));
                    b = __shfl(b, leader, WAVE);
                    if (won) qn[b + (uint32_t)__popcll(m & lt)] = v;
                }
            });
    }
}

}  // namespace

// scratch layout inside g->s_bfs
struct BfsBufs {
    uint32_t* bitmap;
    int32_t* qa;
    int32_t* qb;
    uint32_t* P;
    uint32_t* cnt;   // [0] next frontier size
};

static hn_status bfs_buffers(hn_graph* g, BfsBufs* b) {
    const int64_t nv = g->nv > 0 ? g->nv : 1;
    const int64_t words = (nv + 31) / 32;
    auto al = [](int64_t x) { return (x + 63) & ~int64_t(63); };
    int64_t o_bitmap = 0;
    int64_t o_qa = al(o_bitmap + words * 4);
    int64_t o_qb = al(o_qa + nv * 4);
    int64_t o_P = al(o_qb + nv * 4);
    int64_t o_cnt = al(o_P + (nv + 1) * 4);
    int64_t total = al(o_cnt + 64);
    HN_TRY(g->s_bfs.reserve((size_t)total));
    char* base = g->s_bfs.as<char>();
    b->bitmap = reinterpret_cast<uint32_t*>(base + o_bitmap);
    b->qa = reinterpret_cast<int32_t*>(base + o_qa);
    b->qb = reinterpret_cast<int32_t*>(base + o_qb);
    b->P = reinterpret_cast<uint32_t*>(base + o_P);
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
    bfs_init_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(g->nv, dist, b.bitmap, source, b.qa);
    HN_CHECK_LAUNCH();
    const int64_t grid_cap = (int64_t)num_cus() * 8;
    int64_t F = 1;
    int32_t level = 1;
    int64_t visited = 0, scanned = 0;
    int32_t nlevels = 0;
    uint32_t* host = reinterpret_cast<uint32_t*>(g->h_pinned);
    while (F > 0) {
        visited += F;
        nlevels++;
        HN_TRY((exclusive_scan<uint32_t>(FrontierDegIn{b.qa, g->rec}, b.P, F, g->s_scan.ptr, st)));
        HN_HIP(hipMemsetAsync(b.cnt, 0, 4, st));
        bfs_expand_kernel<<<(unsigned)grid_cap, BLOCK, 0, st>>>(b.qa, F, b.P, g->rec, g->arena, b.bitmap,
                                                                  dist, level, b.qb, b.cnt);
        HN_CHECK_LAUNCH();
        HN_HIP(hipMemcpyAsync(host, b.cnt, 4, hipMemcpyDeviceToHost, st));
        HN_HIP(hipMemcpyAsync(host + 1, b.P + F, 4, hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        scanned += host[1];
        F = host[0];
        std::swap(b.qa, b.qb);
        level++;
    }
    if (stats) {
        stats->levels = nlevels;
        stats->pad_ = 0;
        stats->visited = visited;
        stats->edges_scanned = scanned;
    }
    return HN_OK;
}
