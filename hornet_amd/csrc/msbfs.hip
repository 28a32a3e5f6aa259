// msbfs.hip -- batched multi-source BFS: up to 64 sources per launch.
//
// The reference runs one BfsTopDown per source (TopDown.cu:94-106), paying a
// full traversal of the graph per query.  Here up to 64 sources share one
// traversal: every vertex holds three 64-bit masks (visited, frontier, next),
// bit b standing for source b, so one adjacency read serves every source
// whose frontier touches it.  Per level, a one-thread decide kernel picks the
// direction from counters the previous level left on the device:
//   push  (sparse frontier): one wave per frontier vertex ORs its frontier
//         mask into each neighbour's `next` with a 64-bit atomicOr;
//   pull  (dense frontier): each vertex still missing some source ORs the
//         frontier masks of its neighbours, stopping once every missing bit is
//         found -- no atomics.  Vertices with degree <= T take one lane,
//         heavier ones one wave (T picked by the autotuner).
// A collect kernel turns `next` into visited/frontier bits and per-source
// distances (source-major, int32, INT32_MAX unreached) and leaves the
// counters for the next decision.  The whole level step is captured once into
// a hipGraph and replayed; the host reads a 4-byte "done" flag every few
// levels.  Sources may carry a depth limit (the serving loop's 2-hop queries
// use 2): their bits are dropped from the frontier once the limit is reached.
#include <algorithm>
#include <cstring>
#include <vector>

#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int32_t INF = 0x7fffffff;

struct MsState {
    int32_t level;                  // level being built (sources are level 0)
    int32_t mode;                   // 0 push, 1 pull
    int32_t done;
    int32_t levels_run;
    unsigned long long active;      // vertices with a non-empty frontier
    unsigned long long active_deg;  // sum of their degrees
    unsigned long long open_deg;    // sum of degrees of vertices still missing a bit
    unsigned long long edges;       // sum over sources of degrees of reached vertices
    uint64_t depth_mask[64];        // per level: bits whose depth limit allows expanding
    uint64_t all;                   // bits in use
    unsigned long long push_levels;
    unsigned long long pull_levels;
};

struct MsParams {
    int32_t T;          // pull: degree threshold lane / wave
    int32_t alpha;      // push while active_deg * alpha < open_deg
    int32_t depth_limit[64];
};

// counters for level 0 (the sources) and clearing
__global__ void ms_init_kernel(int64_t nv, uint64_t* __restrict__ visited, uint64_t* __restrict__ frontier,
                               uint64_t* __restrict__ next) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        visited[v] = 0;
        frontier[v] = 0;
        next[v] = 0;
    }
}

__global__ void ms_fill_dist_kernel(int64_t total, int32_t* __restrict__ dist) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        dist[i] = INF;
}

__global__ void ms_seed_kernel(int32_t k, const int32_t* __restrict__ src, int64_t nv, uint64_t* __restrict__ visited,
                               uint64_t* __restrict__ frontier, int32_t* __restrict__ dist, const hn_vrec* __restrict__ rec,
                               MsState* S, MsParams P) {
    // single thread: sources may repeat, so OR serially
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t all = 0;
    unsigned long long edges = 0;
    for (int b = 0; b < k; b++) {
        const int32_t s = src[b];
        visited[s] |= 1ull << b;
        frontier[s] |= 1ull << b;
        dist[(int64_t)b * nv + s] = 0;
        all |= 1ull << b;
        edges += (unsigned long long)rec[s].deg;
    }
    S->all = all;
    S->level = 0;
    S->done = 0;
    S->levels_run = 0;
    S->edges = edges;
    S->push_levels = S->pull_levels = 0;
    // distinct frontier vertices and their degrees; open degree taken as large
    unsigned long long act = 0, adeg = 0;
    for (int b = 0; b < k; b++) {
        bool first = true;
        for (int c = 0; c < b; c++)
            if (src[c] == src[b]) first = false;
        if (first) {
            act++;
            adeg += (unsigned long long)rec[src[b]].deg;
        }
    }
    S->active = act;
    S->active_deg = adeg;
    S->open_deg = ~0ull >> 1;
    for (int L = 0; L < 64; L++) {
        uint64_t m = 0;
        for (int b = 0; b < k; b++)
            if (P.depth_limit[b] <= 0 || L < P.depth_limit[b]) m |= 1ull << b;
        S->depth_mask[L] = m;
    }
}

// one thread: pick the direction of the next level or stop
__global__ void ms_decide_kernel(MsState* S, MsParams P) {
    if (S->done) return;
    if (S->active == 0) {
        S->done = 1;
        return;
    }
    S->level += 1;
    S->levels_run += 1;
    const bool push = (unsigned long long)S->active_deg * (unsigned long long)P.alpha < S->open_deg;
    S->mode = push ? 0 : 1;
    if (push) S->push_levels++; else S->pull_levels++;
    S->active = 0;
    S->active_deg = 0;
    S->open_deg = 0;
}

__device__ __forceinline__ uint64_t frontier_of(const uint64_t* frontier, const MsState* S, int64_t v) {
    const int L = S->level - 1;   // frontier holds level L vertices
    return frontier[v] & S->depth_mask[L < 63 ? L : 63];
}

// push: one wave per vertex with a non-empty (depth-limited) frontier
__global__ __launch_bounds__(BLOCK) void ms_push_kernel(int64_t nv, const hn_vrec* __restrict__ rec,
                                                        const int32_t* __restrict__ arena,
                                                        const uint64_t* __restrict__ frontier,
                                                        const uint64_t* __restrict__ visited,
                                                        unsigned long long* __restrict__ next, const MsState* S) {
    if (S->done || S->mode != 0) return;
    const int64_t wid = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const int64_t nw = (int64_t)gridDim.x * BLOCK / WAVE;
    const int l = __lane_id();
    for (int64_t v = wid; v < nv; v += nw) {
        const uint64_t f = frontier_of(frontier, S, v);
        if (!f) continue;
        const hn_vrec r = rec[v];
        for (int64_t j = l; j < r.deg; j += WAVE) {
            const int32_t u = arena[r.off + j];
            const uint64_t add = f & ~visited[u];
            if (add) atomicOr(next + u, (unsigned long long)add);
        }
    }
}

// pull, light vertices: one lane per vertex
__global__ __launch_bounds__(BLOCK) void ms_pull_light_kernel(int64_t nv, const hn_vrec* __restrict__ rec,
                                                              const int32_t* __restrict__ arena,
                                                              const uint64_t* __restrict__ frontier,
                                                              const uint64_t* __restrict__ visited,
                                                              uint64_t* __restrict__ next, const MsState* S,
                                                              MsParams P) {
    if (S->done || S->mode != 1) return;
    const uint64_t all = S->all;
    const int L = S->level - 1;
    const uint64_t dm = S->depth_mask[L < 63 ? L : 63];
    for (int64_t u = (int64_t)blockIdx.x * BLOCK + threadIdx.x; u < nv; u += (int64_t)gridDim.x * BLOCK) {
        const uint64_t miss = all & ~visited[u];
        if (!miss) continue;
        const hn_vrec r = rec[u];
        if (r.deg > P.T) continue;
        uint64_t acc = 0;
        for (int32_t j = 0; j < r.deg; j++) {
            acc |= frontier[arena[r.off + j]] & dm;
            if ((acc & miss) == miss) break;
        }
        acc &= miss;
        if (acc) next[u] = acc;
    }
}

// pull, heavy vertices: one wave per vertex, 64 neighbours per step
__global__ __launch_bounds__(BLOCK) void ms_pull_heavy_kernel(int64_t nv, const hn_vrec* __restrict__ rec,
                                                              const int32_t* __restrict__ arena,
                                                              const uint64_t* __restrict__ frontier,
                                                              const uint64_t* __restrict__ visited,
                                                              uint64_t* __restrict__ next, const MsState* S,
                                                              MsParams P, const int32_t* __restrict__ heavy,
                                                              int64_t nheavy) {
    if (S->done || S->mode != 1) return;
    const uint64_t all = S->all;
    const int L = S->level - 1;
    const uint64_t dm = S->depth_mask[L < 63 ? L : 63];
    const int64_t wid = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const int64_t nw = (int64_t)gridDim.x * BLOCK / WAVE;
    const int l = __lane_id();
    for (int64_t i = wid; i < nheavy; i += nw) {
        const int32_t u = heavy[i];
        const uint64_t miss = all & ~visited[u];
        if (!miss) continue;
        const hn_vrec r = rec[u];
        uint64_t acc = 0;
        for (int64_t j0 = 0; j0 < r.deg; j0 += WAVE) {
            const int64_t j = j0 + l;
            uint64_t f = j < r.deg ? (frontier[arena[r.off + j]] & dm) : 0ull;
#pragma unroll
            for (int d = WAVE / 2; d > 0; d >>= 1) f |= __shfl_xor(f, d, WAVE);
            acc |= f;
            if ((acc & miss) == miss) break;
        }
        acc &= miss;
        if (l == 0 && acc) next[u] = acc;
    }
}

// collect: new bits -> visited, frontier, distances; counters for the next decision
__global__ __launch_bounds__(BLOCK) void ms_collect_kernel(int64_t nv, int32_t k, const hn_vrec* __restrict__ rec,
                                                           uint64_t* __restrict__ visited,
                                                           uint64_t* __restrict__ frontier,
                                                           uint64_t* __restrict__ next, int32_t* __restrict__ dist,
                                                           MsState* S) {
    if (S->done) return;
    const uint64_t all = S->all;
    const int32_t level = S->level;
    unsigned long long act = 0, adeg = 0, odeg = 0, edges = 0;
    for (int64_t u = (int64_t)blockIdx.x * BLOCK + threadIdx.x; u < nv; u += (int64_t)gridDim.x * BLOCK) {
        const uint64_t vis = visited[u];
        const uint64_t nb = next[u] & ~vis;
        uint64_t f = 0;
        if (nb) {
            next[u] = 0;
            visited[u] = vis | nb;
            const int32_t deg = rec[u].deg;
            act += 1;
            adeg += (unsigned long long)deg;
            edges += (unsigned long long)deg * (unsigned long long)__popcll(nb);
            uint64_t m = nb;
            while (m) {
                const int b = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                dist[(int64_t)b * nv + u] = level;
            }
            f = nb;
            if ((all & ~(vis | nb)) != 0) odeg += (unsigned long long)deg;
        } else if (next[u]) {
            next[u] = 0;
            if ((all & ~vis) != 0) odeg += (unsigned long long)rec[u].deg;
        } else if ((all & ~vis) != 0) {
            odeg += (unsigned long long)rec[u].deg;
        }
        frontier[u] = f;
    }
    act = wave_reduce_sum(act);
    adeg = wave_reduce_sum(adeg);
    odeg = wave_reduce_sum(odeg);
    edges = wave_reduce_sum(edges);
    if (__lane_id() == 0) {
        if (act) atomicAdd(&S->active, act);
        if (adeg) atomicAdd(&S->active_deg, adeg);
        if (odeg) atomicAdd(&S->open_deg, odeg);
        if (edges) atomicAdd(&S->edges, edges);
    }
}

__global__ void heavy_flag_kernel(int64_t nv, const hn_vrec* __restrict__ rec, int32_t T, uint32_t* __restrict__ flag) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) flag[v] = rec[v].deg > T ? 1u : 0u;
}
__global__ void heavy_write_kernel(int64_t nv, const uint32_t* __restrict__ pos, int32_t* __restrict__ heavy) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv && pos[v + 1] != pos[v]) heavy[pos[v]] = (int32_t)v;
}

}  // namespace

// Per-graph multi-source engine state (lives in the graph handle's s_misc-free storage)
struct MsEngine {
    hn_graph* g = nullptr;
    uint64_t* visited = nullptr;
    uint64_t* frontier = nullptr;
    uint64_t* next = nullptr;
    int32_t* heavy = nullptr;
    int64_t nheavy = 0;
    int32_t heavy_T = -1;
    uint64_t heavy_version = 0;       // graph version the heavy list was built for
    MsState* S = nullptr;
    int32_t* d_src = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t exec_stream = nullptr;
    int32_t* exec_dist = nullptr;
    MsParams exec_params{};
    MsParams params{};
    hipStream_t own = nullptr;        // non-blocking stream the level graph is captured on
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int32_t* pinned = nullptr;        // engine-private readback buffer
    void release() {
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        if (own) (void)hipStreamDestroy(own);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        own = nullptr;
        ev_in = ev_out = nullptr;
        dfree(visited); dfree(frontier); dfree(next); dfree(heavy); dfree(S); dfree(d_src);
        visited = frontier = next = nullptr;
        heavy = nullptr;
        S = nullptr;
        d_src = nullptr;
    }
};

// a standalone engine: masks, state, its own stream and pinned readback
// buffer, so several engines may run on one graph from different threads
hn_status ms_engine_create(hn_graph* g, MsEngine** out) {
    MsEngine* e = new MsEngine();
    e->g = g;
    e->params.T = 32;
    e->params.alpha = 16;
    const size_t nv = (size_t)(g->nv > 0 ? g->nv : 1);
    hn_status s;
    if ((s = dalloc(&e->visited, nv)) != HN_OK || (s = dalloc(&e->frontier, nv)) != HN_OK ||
        (s = dalloc(&e->next, nv)) != HN_OK || (s = dalloc(&e->S, 1)) != HN_OK ||
        (s = dalloc(&e->d_src, 64)) != HN_OK) {
        e->release();
        delete e;
        return s;
    }
    if (hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&e->pinned), 4096, 0) != hipSuccess) {
        (void)hipGetLastError();
        e->release();
        delete e;
        set_error("multi-source BFS: stream / event / pinned buffer creation failed");
        return HN_E_HIP;
    }
    *out = e;
    return HN_OK;
}

void ms_engine_destroy(MsEngine* e) {
    if (!e) return;
    e->release();
    delete e;
}

// the graph's default engine, created on first use
static hn_status ms_engine(hn_graph* g, MsEngine** out) {
    if (!g->ms) {
        MsEngine* e = nullptr;
        HN_TRY(ms_engine_create(g, &e));
        g->ms = e;
        g->ms_release = [](void* p) { ms_engine_destroy(static_cast<MsEngine*>(p)); };
    }
    *out = static_cast<MsEngine*>(g->ms);
    return HN_OK;
}

// synchronous readback through the engine's own pinned buffer
static hn_status ms_read(MsEngine* e, const void* dptr, size_t bytes, void* host, hipStream_t st) {
    HN_HIP(hipMemcpyAsync(e->pinned, dptr, bytes, hipMemcpyDeviceToHost, st));
    HN_HIP(hipStreamSynchronize(st));
    std::memcpy(host, e->pinned, bytes);
    return HN_OK;
}

static hn_status ms_heavy_list(MsEngine* e, hipStream_t st) {
    hn_graph* g = e->g;
    // degrees change with every update: rebuild when the graph moved on
    if (e->heavy_T == e->params.T && e->heavy && e->heavy_version == g->version) return HN_OK;
    const int64_t nv = g->nv;
    uint32_t* flag = nullptr;
    uint32_t* pos = nullptr;
    void* tmp = nullptr;
    hn_status s;
    if ((s = dalloc(&flag, (size_t)nv + 1)) != HN_OK) return s;
    if ((s = dalloc(&pos, (size_t)nv + 1)) != HN_OK) { dfree(flag); return s; }
    if ((s = dmalloc(&tmp, scan_temp_bytes(nv))) != HN_OK) { dfree(flag); dfree(pos); return s; }
    if (nv > 0) heavy_flag_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(nv, g->rec, e->params.T, flag);
    s = exclusive_scan<uint32_t>(ArrayIn<uint32_t>{flag}, pos, nv, tmp, st);
    uint32_t cnt = 0;
    if (s == HN_OK) s = ms_read(e, pos + nv, 4, &cnt, st);
    if (s == HN_OK) {
        dfree(e->heavy);
        e->heavy = nullptr;
        s = dalloc(&e->heavy, (size_t)(cnt > 0 ? cnt : 1));
    }
    if (s == HN_OK && nv > 0) {
        heavy_write_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(nv, pos, e->heavy);
        if (hipStreamSynchronize(st) != hipSuccess) s = HN_E_HIP;
    }
    dfree(flag);
    dfree(pos);
    dfree(tmp);
    if (s == HN_OK) {
        e->nheavy = cnt;
        e->heavy_T = e->params.T;
        e->heavy_version = g->version;
        if (e->exec) (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    return s;
}

// record one level step on `st` (also used for capture)
static void ms_level(MsEngine* e, int32_t k, int32_t* dist, const MsParams& P, hipStream_t st) {
    hn_graph* g = e->g;
    const int64_t nv = g->nv;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(nv, BLOCK), (int64_t)num_cus() * 8);
    const unsigned wgrid = (unsigned)((int64_t)num_cus() * 8);
    ms_decide_kernel<<<1, 1, 0, st>>>(e->S, P);
    ms_push_kernel<<<wgrid, BLOCK, 0, st>>>(nv, g->rec, g->arena, e->frontier, e->visited,
                                            reinterpret_cast<unsigned long long*>(e->next), e->S);
    ms_pull_light_kernel<<<grid, BLOCK, 0, st>>>(nv, g->rec, g->arena, e->frontier, e->visited, e->next, e->S, P);
    ms_pull_heavy_kernel<<<wgrid, BLOCK, 0, st>>>(nv, g->rec, g->arena, e->frontier, e->visited, e->next, e->S, P,
                                                  e->heavy, e->nheavy);
    ms_collect_kernel<<<grid, BLOCK, 0, st>>>(nv, k, g->rec, e->visited, e->frontier, e->next, dist, e->S);
}

static hn_status ms_run_on(MsEngine* e, const int32_t* sources, int32_t k, const int32_t* depth_limit,
                           int32_t* dist, hn_msbfs_stats* stats, hipStream_t st);

hn_status ms_run(hn_graph* g, const int32_t* sources, int32_t k, const int32_t* depth_limit, int32_t* dist,
                 hn_msbfs_stats* stats, hipStream_t caller) {
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    return ms_run_engine(e, sources, k, depth_limit, dist, stats, caller);
}

hn_status ms_run_engine(MsEngine* e, const int32_t* sources, int32_t k, const int32_t* depth_limit, int32_t* dist,
                        hn_msbfs_stats* stats, hipStream_t caller) {
    // all work runs on the engine's own stream, ordered after the caller's
    const hipStream_t st = e->own;
    HN_HIP(hipEventRecord(e->ev_in, caller));
    HN_HIP(hipStreamWaitEvent(st, e->ev_in, 0));
    hn_status rs = ms_run_on(e, sources, k, depth_limit, dist, stats, st);
    HN_HIP(hipEventRecord(e->ev_out, st));
    HN_HIP(hipStreamWaitEvent(caller, e->ev_out, 0));
    return rs;
}

static hn_status ms_run_on(MsEngine* e, const int32_t* sources, int32_t k, const int32_t* depth_limit,
                           int32_t* dist, hn_msbfs_stats* stats, hipStream_t st) {
    hn_graph* g = e->g;
    HN_TRY(ms_heavy_list(e, st));
    MsParams P = e->params;
    for (int b = 0; b < 64; b++) P.depth_limit[b] = (depth_limit && b < k) ? depth_limit[b] : 0;
    const int64_t nv = g->nv;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(nv, BLOCK), (int64_t)num_cus() * 8);
    HN_HIP(hipMemcpyAsync(e->d_src, sources, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, st));
    ms_init_kernel<<<grid, BLOCK, 0, st>>>(nv, e->visited, e->frontier, e->next);
    ms_fill_dist_kernel<<<(unsigned)std::min<int64_t>(ceil_div((int64_t)k * nv, BLOCK), (int64_t)num_cus() * 16),
                          BLOCK, 0, st>>>((int64_t)k * nv, dist);
    ms_seed_kernel<<<1, 1, 0, st>>>(k, e->d_src, nv, e->visited, e->frontier, dist, g->rec, e->S, P);
    HN_CHECK_LAUNCH();
    // capture the level step once per (stream, dist buffer, parameters, k)
    const bool same = e->exec && e->exec_stream == st && e->exec_dist == dist &&
                      std::memcmp(&e->exec_params, &P, sizeof(P)) == 0;
    if (!same) {
        if (e->exec) (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
        hipGraph_t graph = nullptr;
        HN_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        ms_level(e, k, dist, P, st);
        hipError_t ce = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess) {
            set_error(std::string("level-step capture: ") + hipGetErrorString(ce));
            return HN_E_HIP;
        }
        hipError_t ie = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            e->exec = nullptr;
            set_error(std::string("level-step instantiate: ") + hipGetErrorString(ie));
            return HN_E_HIP;
        }
        e->exec_stream = st;
        e->exec_dist = dist;
        e->exec_params = P;
    }
    // replay; the done flag is read every `chunk` levels
    const int chunk = 4;
    int32_t* host = e->pinned;
    for (int64_t rep = 0; rep <= nv + 1; rep += chunk) {
        for (int c = 0; c < chunk; c++) HN_HIP(hipGraphLaunch(e->exec, st));
        HN_HIP(hipMemcpyAsync(host, &e->S->done, 4, hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        if (host[0]) break;
    }
    if (stats) {
        MsState s;
        static_assert(sizeof(MsState) <= 4096, "state readback fits the pinned buffer");
        HN_TRY(ms_read(e, e->S, sizeof(MsState), &s, st));
        stats->levels = s.levels_run;
        stats->push_levels = (int32_t)s.push_levels;
        stats->pull_levels = (int32_t)s.pull_levels;
        stats->pad_ = 0;
        stats->edges = (int64_t)s.edges;
    }
    return HN_OK;
}

}  // namespace hn

using namespace hn;

extern "C" {

hn_status hn_bfs_multi(hn_graph* g, const int32_t* sources, int32_t k, int32_t* dist, hn_msbfs_stats* stats,
                       void* stream) {
    HN_ARG(g && sources && dist, "null argument");
    HN_ARG(k >= 1 && k <= 64, "k must be in [1, 64]");
    for (int b = 0; b < k; b++) HN_ARG(sources[b] >= 0 && sources[b] < g->nv, "source outside [0, nv)");
    return ms_run(g, sources, k, nullptr, dist, stats, as_stream(stream));
}

hn_status hn_bfs_multi_depth(hn_graph* g, const int32_t* sources, const int32_t* depth_limit, int32_t k,
                             int32_t* dist, hn_msbfs_stats* stats, void* stream) {
    HN_ARG(g && sources && dist && depth_limit, "null argument");
    HN_ARG(k >= 1 && k <= 64, "k must be in [1, 64]");
    for (int b = 0; b < k; b++) {
        HN_ARG(sources[b] >= 0 && sources[b] < g->nv, "source outside [0, nv)");
        HN_ARG(depth_limit[b] >= 0 && depth_limit[b] <= 63, "depth limit must be in [0, 63] (0 = none)");
    }
    return ms_run(g, sources, k, depth_limit, dist, stats, as_stream(stream));
}

hn_status hn_bfs_multi_params(hn_graph* g, int32_t degree_threshold, int32_t push_alpha) {
    HN_ARG(g, "null graph");
    HN_ARG(degree_threshold >= 0 && push_alpha >= 0, "negative parameter");
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    if (degree_threshold > 0) e->params.T = degree_threshold;
    if (push_alpha > 0) e->params.alpha = push_alpha;
    return HN_OK;
}

hn_status hn_bfs_multi_tune(hn_graph* g, const int32_t* sources, int32_t k, int32_t* dist, int32_t* best_threshold,
                            int32_t* best_alpha, void* stream) {
    HN_ARG(g && sources && dist, "null argument");
    HN_ARG(k >= 1 && k <= 64, "k must be in [1, 64]");
    hipStream_t st = as_stream(stream);
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    const int32_t Ts[] = {8, 16, 32, 64, 128, 256};
    const int32_t As[] = {4, 16, 64};
    hipEvent_t a = nullptr, b = nullptr;
    HN_HIP(hipEventCreate(&a));
    HN_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int32_t bt = e->params.T, ba = e->params.alpha;
    hn_status s = HN_OK;
    for (int32_t T : Ts) {
        for (int32_t A : As) {
            e->params.T = T;
            e->params.alpha = A;
            // one warm run (captures the level graph), one timed run
            if ((s = ms_run(g, sources, k, nullptr, dist, nullptr, st)) != HN_OK) break;
            (void)hipEventRecord(a, st);
            if ((s = ms_run(g, sources, k, nullptr, dist, nullptr, st)) != HN_OK) break;
            (void)hipEventRecord(b, st);
            (void)hipEventSynchronize(b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            if (ms < best) {
                best = ms;
                bt = T;
                ba = A;
            }
        }
        if (s != HN_OK) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->params.T = bt;
    e->params.alpha = ba;
    if (best_threshold) *best_threshold = bt;
    if (best_alpha) *best_alpha = ba;
    return s;
}

}  // extern "C"
