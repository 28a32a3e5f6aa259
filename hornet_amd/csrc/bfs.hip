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
// Here, per level (no host round trip: counts live in a device control
// block, and the host polls its done flag every few levels):
//   1. the frontier arrives as (id, degree, adjacency offset) triples, written
//      by the previous level's collect (one 16-byte record read each);
//   2. one scan whose values pack (degree << 32 | groups), a group being four
//      adjacency slots (16 bytes).  Its down-sweep writes the group prefix P
//      and, per tile, the frontier index holding the tile's first group -- the
//      partition that BinarySearchLB recomputes with two global binary
//      searches per block (BinarySearchLB.i.cuh:254-264); its total carries
//      the level's edge count beside the group count;
//   3. expand: every edge block is a power of two of at least 32 slots and
//      128-byte aligned, so a vertex's ceil(deg/4) groups lie inside its block
//      and each is one aligned 16-byte load (slots past deg in the last group
//      are masked).  Each tile stages its frontier slice (prefix, offset,
//      degree) in LDS and maps a group to its vertex with a fixed-step LDS
//      search.  Every neighbour first probes the visited bitmap as of the
//      start of the level (n/8 bytes, read-only here, so L2 resident); the
//      survivors probe the byte-per-vertex state array (0 = never reached)
//      and an unreached one is marked with an idempotent byte store of 1:
//      probe and mark touch the same cache line, and marks made earlier in
//      the level suppress later ones.  No atomics: with eight
//      XCDs, each holding its own L2, device-scope atomics are not served from
//      cache (claiming vertices with atomicOr measured 7x slower, and even
//      XCD-private mark bitmaps with workgroup-scope atomics measured 1.5x
//      slower than byte stores);
//   4. collect: a lane owns 32 vertices (32 state bytes, one word of the
//      visited bitmap as of the start of the level); a set state byte whose
//      bitmap bit is clear is a newly reached vertex: it gets its distance, its
//      bitmap bit and one next-frontier slot (one atomicAdd per workgroup),
//      so every vertex enters a frontier once;
//   5. a one-thread step that adds the level to the statistics and makes the
//      next frontier current (done when it is empty).
// Distances: source 0, its neighbours 1, unreached INT32_MAX (TopDown.cu:42,81).
#include <algorithm>
#include <cstddef>
#include <cstring>

#include "bfs.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int EXP_G = 4;                       // adjacency slots per work unit
constexpr int EXP_IPT = 4;                     // work units per lane per tile
constexpr int EXP_TILE = BLOCK * EXP_IPT;      // work units per tile
constexpr int32_t INF = 0x7fffffff;
// The expansion keeps the first FULL_LDS_WORDS words of the full-word summary
// in LDS (12 KB: vertices below 3 * 2^20).  A neighbour whose bitmap word is known
// full was visited before the level, so its bitmap probe is skipped.
constexpr int FULL_LDS_WORDS = 3072;
static_assert(UNIT_SLOTS % EXP_G == 0, "groups must not straddle a block start");

struct FrontierIn {
    const int32_t* qdeg;
    __device__ __forceinline__ uint64_t operator()(int64_t i) const {
        const uint32_t d = (uint32_t)qdeg[i];
        return ((uint64_t)d << 32) | ((d + EXP_G - 1) / EXP_G);
    }
};

// per frontier vertex i with exclusive packed prefix `ex` and value `x`
struct FrontierOut {
    uint32_t* P;
    uint32_t* tile_lo;
    __device__ __forceinline__ void operator()(int64_t i, uint64_t ex, uint64_t x) const {
        const uint32_t g0 = (uint32_t)ex, ng = (uint32_t)x;
        P[i] = g0;
        if (ng) {
            uint32_t t0 = (uint32_t)(((uint64_t)g0 + EXP_TILE - 1) / EXP_TILE);
            uint32_t t1 = (uint32_t)(((uint64_t)g0 + ng + EXP_TILE - 1) / EXP_TILE);
            for (uint32_t t = t0; t < t1; t++) tile_lo[t] = (uint32_t)i;
        }
    }
};

__global__ void bfs_init_kernel(int32_t nv, const hn_vrec* __restrict__ rec, int32_t* __restrict__ dist,
                                uint32_t* __restrict__ bitmap, uint32_t* __restrict__ full,
                                uint8_t* __restrict__ state, int32_t source, int32_t* __restrict__ q,
                                int32_t* __restrict__ qdeg, int64_t* __restrict__ qoff, BfsCtl* __restrict__ ctl) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t words = ((int64_t)nv + 31) / 32;
    if (i < nv) dist[i] = (i == source) ? 0 : INF;
    if (i < words) bitmap[i] = (i == (source >> 5)) ? (1u << (source & 31)) : 0u;
    if (i < (words + 31) / 32) full[i] = 0u;
    if (i == 0) {
        const hn_vrec r = rec[source];
        q[0] = source;
        qdeg[0] = r.deg;
        qoff[0] = r.off;
        state[source] = 1;
        BfsCtl c{};
        c.F = 1;
        c.level = 1;
        *ctl = c;
    }
}

// End of a level: its frontier joins the statistics and the next frontier
// becomes current; an empty one ends the traversal.  Levels launched after
// the end find F == 0 and change nothing.
__global__ void bfs_advance_kernel(BfsCtl* __restrict__ ctl, const uint64_t* __restrict__ tot) {
    BfsCtl c = *ctl;
    if (c.F == 0) return;
    c.visited += c.F;
    c.scanned += tot[0] >> 32;
    c.levels += 1;
    c.F = c.nextF;
    c.nextF = 0;
    c.level += 1;
    c.done = c.F == 0 ? 1u : 0u;
    *ctl = c;
}

// the frontier slice of a tile: group prefix, adjacency offset in 128-byte
// units (every block starts on one), degree
using ExpLb = LbShared<EXP_IPT, uint32_t, uint32_t>;
struct ExpShared {
    ExpLb lb;
    int32_t deg[ExpLb::CAP];
    uint32_t full[FULL_LDS_WORDS];     // head of the full-word summary
};

// Every neighbour of the frontier is tested against the visited bitmap and,
// if not visited before this level, against state[]; unreached ones are
// marked.  state[] is only ever set to 1 here, so concurrent stores agree.
template <int G>
__global__ __launch_bounds__(BLOCK) void bfs_expand_kernel(const uint64_t* __restrict__ tot,
                                                           const uint32_t* __restrict__ P,
                                                           const uint32_t* __restrict__ tile_lo,
                                                           const int32_t* __restrict__ qdeg,
                                                           const int64_t* __restrict__ qoff,
                                                           const BfsCtl* __restrict__ ctl,
                                                           const int32_t* __restrict__ arena,
                                                           uint8_t* __restrict__ state,
                                                           const uint32_t* __restrict__ bitmap,
                                                           const uint32_t* __restrict__ full, int64_t fwords) {
    typedef int32_t Vec __attribute__((ext_vector_type(G)));
    constexpr int CAP = ExpLb::CAP;
    __shared__ ExpShared sh;
    const int64_t F = ctl->F;
    const int64_t W = F > 0 ? (int64_t)(uint32_t)tot[0] : 0;
    const int64_t ntiles = (W + EXP_TILE - 1) / EXP_TILE;
    if ((int64_t)blockIdx.x >= ntiles) return;
    for (int i = threadIdx.x; i < FULL_LDS_WORDS; i += BLOCK) sh.full[i] = i < fwords ? full[i] : 0u;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t lo = tile_lo[t];
        const int64_t hi = (t + 1 < ntiles) ? (int64_t)tile_lo[t + 1] : F - 1;
        const int64_t w0 = t * EXP_TILE;
        const int64_t span = hi - lo + 1;
        const bool staged = span <= CAP;
        if (staged) {
            for (int i = threadIdx.x; i < (int)span; i += BLOCK) {
                sh.lb.pref[i] = P[lo + i];
                sh.lb.base[i] = (uint32_t)(qoff[lo + i] / UNIT_SLOTS);
                sh.deg[i] = qdeg[lo + i];
            }
        }
        __syncthreads();
        // phase 1: every unit's slot address and valid slot count
        int64_t addr[EXP_IPT];
        int nval[EXP_IPT];
        if (staged) {
#pragma unroll
            for (int j = 0; j < EXP_IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                const int s = lb_locate<EXP_IPT>(sh.lb, (int)span, w);
                const int64_t k = (w - (int64_t)sh.lb.pref[s]) * G;
                addr[j] = (int64_t)sh.lb.base[s] * UNIT_SLOTS + k;
                nval[j] = w < W ? (int)min((int64_t)G, (int64_t)sh.deg[s] - k) : 0;
            }
        } else {   // slice wider than LDS (runs of zero-degree frontier vertices)
            for (int j = 0; j < EXP_IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                addr[j] = 0;
                nval[j] = 0;
                if (w < W) {
                    const int64_t s = upper_bound_m1<int64_t>(P, lo, hi + 1, w);
                    const int64_t k = (w - (int64_t)P[s]) * G;
                    addr[j] = qoff[s] + k;
                    nval[j] = (int)min((int64_t)G, (int64_t)qdeg[s] - k);
                }
            }
        }
        // Phases 2 and 3 load unconditionally (a unit without work reads
        // arena[0..G), a masked slot probes state[0]) so that no load sits
        // in a branch of its own, where the compiler waits for it at once.
        // phase 2: all adjacency groups in flight at once (streamed, not kept in L2)
        Vec v[EXP_IPT];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
            v[j] = __builtin_nontemporal_load(reinterpret_cast<const Vec*>(arena + (nval[j] > 0 ? addr[j] : 0)));
        // phase 3a: all visited-bitmap probes in flight at once (n/8 bytes,
        // L2 resident), folded branch-free into one bit per candidate slot
        // (a neighbour in a word the LDS summary knows full probes word 0 instead)
        uint32_t known = 0;
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++) {
                const uint32_t t = e < nval[j] ? (uint32_t)v[j][e] : 0u;
                const uint32_t fw = t >> 10;
                const uint32_t f = sh.full[fw < (uint32_t)FULL_LDS_WORDS ? fw : 0u];
                known |= (uint32_t)((fw < (uint32_t)FULL_LDS_WORDS) & ((f >> ((t >> 5) & 31)) & 1u)) << (j * G + e);
            }
        uint32_t wb[EXP_IPT][G];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++)
                wb[j][e] = bitmap[((e < nval[j]) && !((known >> (j * G + e)) & 1u)) ? (v[j][e] >> 5) : 0];
        uint32_t cand = 0;
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++)
                cand |= (uint32_t)((e < nval[j]) & !((known >> (j * G + e)) & 1u) &
                                   (((wb[j][e] >> (v[j][e] & 31)) & 1u) == 0)) << (j * G + e);
        // phase 3b: candidates (not visited before this level) probe their
        // state byte; the rest probe state[0], one shared line
        uint8_t s[EXP_IPT][G];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++) s[j][e] = state[((cand >> (j * G + e)) & 1u) ? v[j][e] : 0];
        uint32_t need = 0;
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++)
                need |= (uint32_t)(((cand >> (j * G + e)) & 1u) & (s[j][e] == 0)) << (j * G + e);
        // phase 4: idempotent marks
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++)
                if (need & (1u << (j * G + e))) state[v[j][e]] = 1;
        __syncthreads();
    }
}

// One lane owns 32 vertices (32 state bytes and one visited-bitmap word):
// set state bytes not yet in the bitmap are this level's discoveries; they
// get their distance, their bitmap bit and a slot in the next frontier.
__global__ __launch_bounds__(BLOCK) void bfs_collect_kernel(int64_t words, const uint32_t* __restrict__ state4,
                                                            uint32_t* __restrict__ bitmap,
                                                            uint32_t* __restrict__ full,
                                                            int32_t* __restrict__ dist,
                                                            const hn_vrec* __restrict__ rec,
                                                            int32_t* __restrict__ qn, int32_t* __restrict__ qdeg,
                                                            int64_t* __restrict__ qoff, BfsCtl* __restrict__ ctl) {
    if (ctl->F == 0) return;   // a level launched after the traversal ended
    const int32_t level = (int32_t)ctl->level;
    uint32_t* qn_count = &ctl->nextF;
    const int64_t wi = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t found = 0;
    bool wfull = false;
    if (wi < words) {
        const uint4* p = reinterpret_cast<const uint4*>(state4 + wi * 8);
        const uint4 a = p[0], b = p[1];
        const uint32_t m[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t set = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t x = m[k];
            // byte j of word k is vertex 32*wi + 4k + j
            set |= ((x & 0xFFu) ? 1u : 0u) << (4 * k);
            set |= ((x & 0xFF00u) ? 1u : 0u) << (4 * k + 1);
            set |= ((x & 0xFF0000u) ? 1u : 0u) << (4 * k + 2);
            set |= ((x & 0xFF000000u) ? 1u : 0u) << (4 * k + 3);
        }
        if (set) {
            const uint32_t old = bitmap[wi];
            found = set & ~old;
            if (found) bitmap[wi] = old | found;
        }
        // state bytes are never cleared during a traversal, so `set` is every
        // vertex of the word visited so far: the word is full iff all are
        wfull = set == 0xFFFFFFFFu;
    }
    {
        // the full-word summary: a wave owns 64 consecutive words = 2 summary words
        const uint64_t fb = __ballot(wfull);
        const int64_t fw = (wi - __lane_id()) >> 5;
        if (__lane_id() == 0 && fb && fw < (words + 31) / 32) full[fw] = (uint32_t)fb;
        if (__lane_id() == 1 && (fb >> 32) && fw + 1 < (words + 31) / 32) full[fw + 1] = (uint32_t)(fb >> 32);
    }
    // next-frontier slots: one atomicAdd per workgroup (the counter is a
    // single device-scope address, so fewer of them is what matters)
    __shared__ uint32_t wsum[NWAVES + 1];
    const uint32_t cnt = (uint32_t)__popc(found);
    const uint32_t inc = wave_inclusive_scan(cnt);
    if (__lane_id() == WAVE - 1) wsum[threadIdx.x / WAVE] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int k = 0; k < NWAVES; k++) {
            const uint32_t x = wsum[k];
            wsum[k] = acc;
            acc += x;
        }
        wsum[NWAVES] = acc ? atomicAdd(qn_count, acc) : 0u;
    }
    __syncthreads();
    if (__shfl(inc, WAVE - 1, WAVE) == 0) return;   // nothing found by this wave
    const uint32_t b = wsum[NWAVES] + wsum[threadIdx.x / WAVE] + inc - cnt;
    // Writes, transposed so that consecutive lanes write consecutive vertices:
    // in round r lane l handles bit (l & 31) of the wave's word 2r + (l >> 5),
    // fetching that word's found bits and first queue slot from its owner.
    const int lane = __lane_id();
    const int64_t wbase = wi - lane;
    const int bit = lane & 31;
    const uint32_t below = (1u << bit) - 1u;
#pragma unroll 4
    for (int r = 0; r < WAVE / 2; r++) {
        const int owner = 2 * r + (lane >> 5);
        const uint32_t fw = __shfl(found, owner, WAVE);
        const uint32_t bo = __shfl(b, owner, WAVE);
        if ((fw >> bit) & 1u) {
            const int32_t v = (int32_t)((wbase + owner) * 32 + bit);
            const uint32_t slot = bo + (uint32_t)__popc(fw & below);
            const hn_vrec r = rec[v];
            dist[v] = level;
            qn[slot] = v;
            qdeg[slot] = r.deg;
            qoff[slot] = r.off;
        }
    }
}

}  // namespace

// Lays the buffers of a traversal of a graph with nv vertices and ne edges
// out from `base` (nullptr: sizes only); returns the bytes they take.
int64_t bfs_layout(int64_t nv_in, int64_t ne, char* base, BfsBufs* b) {
    const int64_t nv = nv_in > 0 ? nv_in : 1;
    const int64_t words = (nv + 31) / 32;
    const int64_t tiles = (ne / EXP_G + nv) / EXP_TILE + 2;
    auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
    const int64_t o_bitmap = 0;
    const int64_t o_full = al(o_bitmap + words * 4);
    const int64_t o_q = al(o_full + (words + 31) / 32 * 4);
    const int64_t o_qdeg = al(o_q + nv * 4);
    const int64_t o_qoff = al(o_qdeg + nv * 4);
    const int64_t o_P = al(o_qoff + nv * 8);
    const int64_t o_tile = al(o_P + (nv + 1) * 4);
    const int64_t o_state = al(o_tile + tiles * 4);
    const int64_t o_tot = al(o_state + words * 32);
    const int64_t o_ctl = al(o_tot + 8);
    const int64_t o_scan = al(o_ctl + (int64_t)sizeof(BfsCtl));
    const int64_t total = al(o_scan + (int64_t)scan_temp_bytes(nv));
    if (base) {
        b->bitmap = reinterpret_cast<uint32_t*>(base + o_bitmap);
        b->full = reinterpret_cast<uint32_t*>(base + o_full);
        b->q = reinterpret_cast<int32_t*>(base + o_q);
        b->qdeg = reinterpret_cast<int32_t*>(base + o_qdeg);
        b->qoff = reinterpret_cast<int64_t*>(base + o_qoff);
        b->P = reinterpret_cast<uint32_t*>(base + o_P);
        b->tile_lo = reinterpret_cast<uint32_t*>(base + o_tile);
        b->state = reinterpret_cast<uint8_t*>(base + o_state);
        b->tot = reinterpret_cast<uint64_t*>(base + o_tot);
        b->ctl = reinterpret_cast<BfsCtl*>(base + o_ctl);
        b->scan_tmp = base + o_scan;
    }
    return total;
}

__global__ void zero_u32_kernel(int64_t n, uint32_t* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The level loop over laid-out buffers.  Levels are queued without a host
// round trip; the host reads the done flag after as many levels as the last
// traversal of this graph took, then after 1, 2, 4, 4, ... more (levels
// queued past the end change nothing).  Nothing here allocates (kernel timing,
// which creates its events, only with `may_time`).
hn_status bfs_run(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats, hipStream_t st,
                  const BfsBufs& b, bool may_time, uint64_t* host) {
    const int64_t nv = g->nv;
    const int64_t words = (nv + 31) / 32;
    // state starts at zero (a fill kernel, not a memset node, so a captured
    // copy of this sequence replays correctly)
    zero_u32_kernel<<<(unsigned)ceil_div(words * 8, BLOCK), BLOCK, 0, st>>>(words * 8,
                                                                          reinterpret_cast<uint32_t*>(b.state));
    bfs_init_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(g->nv, g->rec, dist, b.bitmap, b.full, b.state,
                                                                      source, b.q, b.qdeg, b.qoff, b.ctl);
    HN_CHECK_LAUNCH();
    const int64_t grid_cap = (int64_t)num_cus() * 64;
    const bool timed = may_time && stats && stats->time_kernels;
    // kernel timing: one event pair around each queued expansion of a chunk
    constexpr int MAX_CHUNK = hn_graph::BFS_EVENTS / 2;
    hipEvent_t* ev = g->bfs_ev;
    hn_status rc = HN_OK;
    if (timed)
        for (int i = 0; i < 2 * MAX_CHUNK && rc == HN_OK; i++)
            if (!ev[i] && hipEventCreate(&ev[i]) != hipSuccess) {
                (void)hipGetLastError();
                ev[i] = nullptr;
                set_error("bfs: event creation failed");
                rc = HN_E_HIP;
            }
    const int last = g->bfs_last_levels.load(std::memory_order_relaxed);
    int64_t next = std::min<int64_t>(last > 0 ? last : 4, MAX_CHUNK);
    int64_t ramp = 1, queued = 0, kns = 0;
    uint32_t levels_seen = 0;
    static_assert(offsetof(BfsCtl, done) == offsetof(BfsCtl, levels) + 4, "levels, done adjacent");
    while (rc == HN_OK) {
        hipError_t e = hipSuccess;
        for (int64_t c = 0; c < next && e == hipSuccess; c++) {
            e = exclusive_scan_dev_count<uint64_t>(FrontierIn{b.qdeg}, FrontierOut{b.P, b.tile_lo}, b.tot,
                                                    &b.ctl->F, nv, b.scan_tmp, st);
            if (timed && e == hipSuccess) e = hipEventRecord(ev[2 * c], st);
            if (e == hipSuccess) {
                bfs_expand_kernel<EXP_G><<<(unsigned)grid_cap, BLOCK, 0, st>>>(
                    b.tot, b.P, b.tile_lo, b.qdeg, b.qoff, b.ctl, g->arena, b.state, b.bitmap, b.full,
                    (words + 31) / 32);
                e = hipGetLastError();
            }
            if (timed && e == hipSuccess) e = hipEventRecord(ev[2 * c + 1], st);
            if (e == hipSuccess) {
                bfs_collect_kernel<<<(unsigned)ceil_div(words, BLOCK), BLOCK, 0, st>>>(
                    words, reinterpret_cast<const uint32_t*>(b.state), b.bitmap, b.full, dist, g->rec, b.q, b.qdeg,
                    b.qoff, b.ctl);
                e = hipGetLastError();
            }
            if (e == hipSuccess) {
                bfs_advance_kernel<<<1, 1, 0, st>>>(b.ctl, b.tot);
                e = hipGetLastError();
            }
        }
        queued += next;
        // [0] = levels | done << 32
        if (e == hipSuccess) e = hipMemcpyAsync(host, &b.ctl->levels, 8, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        const uint32_t levels_now = (uint32_t)host[0];
        const bool done = (host[0] >> 32) != 0;
        if (timed && e == hipSuccess)
            // only the expansions of non-empty levels count
            for (uint32_t c = 0; c < levels_now - levels_seen && e == hipSuccess; c++) {
                float ms = 0.f;
                e = hipEventElapsedTime(&ms, ev[2 * c], ev[2 * c + 1]);
                kns += (int64_t)((double)ms * 1e6);
            }
        if (e != hipSuccess) {
            set_error(std::string("bfs level: ") + hipGetErrorString(e));
            rc = HN_E_HIP;
            break;
        }
        levels_seen = levels_now;
        if (done) break;
        if (queued > nv + 1) {   // cannot happen: every level reaches a new vertex
            set_error("bfs: level loop did not end");
            rc = HN_E_HIP;
            break;
        }
        next = ramp;
        ramp = std::min<int64_t>(2 * ramp, 4);
    }
    if (rc != HN_OK) return rc;
    g->bfs_last_levels.store((int)levels_seen, std::memory_order_relaxed);
    if (stats) {
        BfsCtl c;
        HN_HIP(hipMemcpyAsync(host, b.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        std::memcpy(&c, host, sizeof(BfsCtl));
        stats->levels = (int32_t)c.levels;
        stats->visited = (int64_t)c.visited;
        stats->edges_scanned = (int64_t)c.scanned;
        stats->expand_ns = kns;
        stats->expand_launches = (int32_t)c.levels;
        stats->pad_ = 0;
    }
    return HN_OK;
}

}  // namespace hn

using namespace hn;

extern "C" hn_status hn_bfs_topdown(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats,
                                    void* stream) {
    HN_ARG(g && dist, "null argument");
    HN_ARG(source >= 0 && source < g->nv, "source outside [0, nv)");
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds the 32-bit work prefix");
    BfsBufs b;
    HN_TRY(g->s_bfs.reserve((size_t)bfs_layout(g->nv, g->ne, nullptr, &b)));
    bfs_layout(g->nv, g->ne, g->s_bfs.as<char>(), &b);
    return bfs_run(g, source, dist, stats, as_stream(stream), b, true, g->h_pinned);
}


// ---------------------------------------------------------------------------
// BFS TopDown without allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops).  Its traversal scratch is one
// static device array per GPU, held until the process ends; its device address
// is looked up once per device with hipGetSymbolAddress, and only that address
// is handed to kernels (the array's name is a host-side stand-in).  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace hn {
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up on its first use.
hn_status noalloc_scratch(char** out) {
    int dev = -1;
    HN_HIP(hipGetDevice(&dev));
    HN_ARG(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device ordinal beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("no-allocation scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return HN_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return HN_OK;
}

}  // namespace
}  // namespace hn

extern "C" hn_status hn_bfs_topdown_noalloc_limit(int64_t* scratch_bytes) {
    HN_ARG(scratch_bytes, "null argument");
    *scratch_bytes = (int64_t)NOALLOC_SCRATCH_BYTES;
    return HN_OK;
}

extern "C" hn_status hn_bfs_topdown_noalloc(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats,
                                            void* stream) {
    HN_ARG(g && dist, "null argument");
    HN_ARG(source >= 0 && source < g->nv, "source outside [0, nv)");
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds the 32-bit work prefix");
    BfsBufs b;
    const int64_t need = bfs_layout(g->nv, g->ne, nullptr, &b);
    if (need > (int64_t)NOALLOC_SCRATCH_BYTES) {
        set_error("hn_bfs_topdown_noalloc: the graph needs " + std::to_string(need) +
                  " bytes of traversal scratch, more than the static " +
                  std::to_string(NOALLOC_SCRATCH_BYTES) + " (use hn_bfs_topdown)");
        return HN_E_CAPACITY;
    }
    char* base = nullptr;
    HN_TRY(noalloc_scratch(&base));
    bfs_layout(g->nv, g->ne, base, &b);
    return bfs_run(g, source, dist, stats, as_stream(stream), b, false, g->h_pinned);
}
