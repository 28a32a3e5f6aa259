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
// block, which the host reads every few levels), two launches:
//   1. expand: the frontier arrives as (degree, adjacency offset, group
//      prefix) triples plus, per tile of the expansion, the frontier index
//      holding the tile's first group -- the partition that BinarySearchLB
//      recomputes with two global binary searches per block
//      (BinarySearchLB.i.cuh:254-264); a group is four adjacency slots (16
//      bytes).  Every edge block is a power of two of at least 32 slots and
//      128-byte aligned, so a vertex's ceil(deg/4) groups lie inside its block
//      and each is one aligned 16-byte load (slots past deg in the last group
//      are masked).  Each tile stages its frontier slice in LDS and maps a
//      group to its vertex with a fixed-step LDS search.  A neighbour in a
//      bitmap word that the LDS summary knows full was reached before the
//      level; every other one first probes the visited bitmap as of the start
//      of the level (n/8 bytes, read-only here, so L2 resident), and the
//      survivors probe the byte-per-vertex state array (0 = never reached) --
//      on levels where most of the adjacency still leads to unreached vertices
//      the bitmap probe is skipped and the state probe alone decides.  An
//      unreached neighbour is marked with an idempotent byte store of 1: probe
//      and mark touch the same cache line, and marks made earlier in the level
//      suppress later ones.  No atomics:
//      with eight XCDs, each holding its own L2, device-scope atomics are not
//      served from cache (claiming vertices with atomicOr measured 7x slower,
//      and even XCD-private mark bitmaps with workgroup-scope atomics measured
//      1.5x slower than byte stores);
//   2. collect: a lane owns 32 vertices (32 state bytes, one word of the
//      visited bitmap as of the start of the level); a set state byte whose
//      bitmap bit is clear is a newly reached vertex: it gets its distance and
//      its bitmap bit, so every vertex enters a frontier once.  A workgroup
//      sums its discoveries' degrees and groups, claims their frontier slots
//      and group range with one 64-bit atomicAdd of (count << 32 | groups), and
//      writes each discovery's triple and the tile starts inside its groups:
//      the exclusive scan of the frontier's groups that the reference runs as
//      a separate pass (cub ExclusiveSum) falls out of the slot claims.
// The frontier counters alternate by level parity, so no launch is needed to
// make the next frontier current: the host passes each launch its level.
// Distances: source 0, its neighbours 1, unreached INT32_MAX (TopDown.cu:42,81).
#include <algorithm>
#include <cstddef>
#include <cstring>

#include <hip/hip_ext.h>

#include "bfs.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int EXP_G = 4;                       // adjacency slots per work unit
#ifndef BFS_EXP_IPT
#define BFS_EXP_IPT 4
#endif
#ifndef BFS_FULL_LDS_WORDS
#define BFS_FULL_LDS_WORDS 7168
#endif
#ifndef BFS_GRID_PER_CU
#define BFS_GRID_PER_CU 64   // expansion grid, in 256-lane workgroups per CU
#endif
#ifndef BFS_EXP_BLOCK
#define BFS_EXP_BLOCK 512
#endif
constexpr int EXP_IPT = BFS_EXP_IPT;           // work units per lane per tile
constexpr int EXP_BLOCK = BFS_EXP_BLOCK;       // expansion workgroup size
constexpr int EXP_WAVES = EXP_BLOCK / WAVE;
constexpr int EXP_TILE = EXP_BLOCK * EXP_IPT;  // work units per tile
constexpr int32_t INF = 0x7fffffff;
// The expansion keeps the first FULL_LDS_WORDS words of the full-word summary
// in LDS (28 KB: vertices below 7 * 2^20).  A neighbour whose bitmap word is known
// full was visited before the level, so its bitmap probe is skipped.  512-lane
// workgroups share one copy per 8 waves: with the 24 KB load-balancing stage,
// three workgroups fill a CU's 160 KB.
constexpr int FULL_LDS_WORDS = BFS_FULL_LDS_WORDS;
#ifndef BFS_DIRECT_OPEN_PCT
#define BFS_DIRECT_OPEN_PCT 25   // skip bitmap probes while more than this % of the adjacency volume is unreached
#endif
static_assert(UNIT_SLOTS % EXP_G == 0, "groups must not straddle a block start");

__global__ void bfs_init_kernel(int32_t nv, const hn_vrec* __restrict__ rec, int32_t* __restrict__ dist,
                                uint32_t* __restrict__ bitmap, uint32_t* __restrict__ full,
                                uint32_t* __restrict__ state4, int32_t source, int32_t* __restrict__ qdeg,
                                uint32_t* __restrict__ qbase, uint32_t* __restrict__ P,
                                uint32_t* __restrict__ tile_lo, uint64_t* __restrict__ part,
                                BfsCtl* __restrict__ ctl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    const int64_t words = ((int64_t)nv + 31) / 32;
    for (int64_t k = i; k < (words + BLOCK - 1) / BLOCK; k += nthr) part[k] = 0u;
    for (int64_t k = i; k < (words + 31) / 32; k += nthr) full[k] = 0u;
    for (int64_t k = i; k < words; k += nthr) bitmap[k] = (k == (source >> 5)) ? (1u << (source & 31)) : 0u;
    // distances: 16-byte stores where the caller's array allows, the source's
    // entry set within its own vector
    const int64_t n4 = (reinterpret_cast<uintptr_t>(dist) & 15) == 0 ? (int64_t)nv / 4 : 0;
    for (int64_t k = i; k < n4; k += nthr) {
        int4 q = make_int4(INF, INF, INF, INF);
        if (k == (source >> 2)) {
            const int o = source & 3;
            q.x = o == 0 ? 0 : q.x;
            q.y = o == 1 ? 0 : q.y;
            q.z = o == 2 ? 0 : q.z;
            q.w = o == 3 ? 0 : q.w;
        }
        reinterpret_cast<int4*>(dist)[k] = q;
    }
    for (int64_t k = 4 * n4 + i; k < nv; k += nthr) dist[k] = (k == source) ? 0 : INF;
    // state bytes (4 per word, words * 8 words, 256-byte aligned): the
    // source's alone set
    for (int64_t k = i; k < words * 2; k += nthr) {
        uint4 q = make_uint4(0u, 0u, 0u, 0u);
        if (k == (source >> 4)) {
            const uint32_t b = 1u << (8 * (source & 3));
            const int o = (source >> 2) & 3;
            q.x = o == 0 ? b : 0u;
            q.y = o == 1 ? b : 0u;
            q.z = o == 2 ? b : 0u;
            q.w = o == 3 ? b : 0u;
        }
        reinterpret_cast<uint4*>(state4)[k] = q;
    }
    // the first frontier: the source alone, owning every tile of its groups
    const hn_vrec r = rec[source];
    const int64_t ng = ((int64_t)r.deg + EXP_G - 1) / EXP_G;
    for (int64_t t = i; t < (ng + EXP_TILE - 1) / EXP_TILE; t += (int64_t)gridDim.x * blockDim.x) tile_lo[t] = 0u;
    if (i == 0) {
        qdeg[0] = r.deg;
        qbase[0] = (uint32_t)(r.off / UNIT_SLOTS);
        P[0] = 0u;
        BfsCtl c{};
        c.fc[1] = ((uint64_t)1 << 32) | (uint64_t)ng;
        c.scanned[1] = (uint64_t)r.deg;   // [0]: nothing visited before level 1
        c.src_deg = (uint64_t)r.deg;
        *ctl = c;
    }
}

// Probes go through buffer loads: a lane without a probe to make gets an
// offset past the buffer's end, which the hardware's range check answers with
// 0 and no cache access (a plain load of a shared dummy address would still
// take a cache-tag lookup in every group of four lanes that holds one).
constexpr uint32_t PROBE_NONE = 0xFFFFFFF0u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t probe_rsrc(const void* p, uint32_t bytes) {
    // raw buffer (stride 0), gfx9-family descriptor word 3
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t probe_b32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ uint8_t probe_u8(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint8_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0);
}

// the frontier slice of a tile: group prefix, adjacency offset in 128-byte
// units (every block starts on one), degree
// (LbShared sizes its stage for BLOCK threads: EXP_TILE + 1 entries)
using ExpLb = LbShared<EXP_TILE / BLOCK, uint32_t, uint32_t>;
static_assert(ExpLb::CAP == EXP_TILE + 1, "stage holds a tile's segments");
struct ExpShared {
    ExpLb lb;
    int32_t deg[ExpLb::CAP];
    uint32_t full[FULL_LDS_WORDS];     // head of the full-word summary
};

// Every neighbour of the frontier is tested against the visited bitmap and,
// if not visited before this level, against state[]; unreached ones are
// marked.  state[] is only ever set to 1 here, so concurrent stores agree.
template <int G>
__global__ __launch_bounds__(EXP_BLOCK) void bfs_expand_kernel(uint32_t L, const uint32_t* __restrict__ P,
                                                           const uint32_t* __restrict__ tile_lo,
                                                           const int32_t* __restrict__ qdeg,
                                                           const uint32_t* __restrict__ qbase,
                                                           BfsCtl* __restrict__ ctl,
                                                           const int32_t* __restrict__ arena,
                                                           uint8_t* __restrict__ state,
                                                           const uint32_t* __restrict__ bitmap,
                                                           const uint32_t* __restrict__ full, int64_t words,
                                                           int64_t fwords,
                                                           const uint64_t* __restrict__ part, int64_t nparts,
                                                           int64_t ne) {
    typedef int32_t Vec __attribute__((ext_vector_type(G)));
    constexpr int CAP = ExpLb::CAP;
    __shared__ ExpShared sh;
    const __amdgpu_buffer_rsrc_t rbm = probe_rsrc(bitmap, (uint32_t)(words * 4));
    const __amdgpu_buffer_rsrc_t rst = probe_rsrc(state, (uint32_t)(words * 32));
    const uint64_t fcur = ctl->fc[L & 1];
    const int64_t F = (int64_t)(fcur >> 32);
    const int64_t W = F > 0 ? (int64_t)(uint32_t)fcur : 0;
    // Skip the bitmap probes while most adjacency entries still lead to
    // unreached vertices: the state bytes are cumulative (every vertex reached
    // so far has its byte set), so the state probe alone decides, and the
    // bitmap would only filter the few reached targets.  The share of entries
    // leading to reached vertices is estimated by the adjacency volume reached
    // before this level (written by the previous expansion) plus this
    // frontier's (4 slots per group).
    const uint64_t vol = ctl->scanned[(L + 1) & 1] + (uint64_t)W * G;
    const bool direct = (uint64_t)(100 - BFS_DIRECT_OPEN_PCT) * (uint64_t)ne > 100 * vol;
    // level bookkeeping by workgroup 0: this level's collect counts into the
    // other counter (nothing else touches it during this launch), and the
    // degrees of every discovery so far, one partial per collect workgroup,
    // join the source's in `scanned`
    if (blockIdx.x == 0 && F > 0) {
        __shared__ uint64_t red[EXP_WAVES];
        uint64_t x = 0;
        for (int64_t i = threadIdx.x; i < nparts; i += EXP_BLOCK) x += part[i];
        x = wave_reduce_sum(x);
        if (__lane_id() == 0) red[threadIdx.x / WAVE] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = ctl->src_deg;
            for (int k = 0; k < EXP_WAVES; k++) t += red[k];
            ctl->scanned[L & 1] = t;
            ctl->levels = L;
            ctl->visited += (uint64_t)F;
            ctl->dsum += (uint64_t)F * (uint64_t)(L - 1);   // level L's frontier lies at distance L - 1
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->fc[(L + 1) & 1] = 0;
    const int64_t ntiles = (W + EXP_TILE - 1) / EXP_TILE;
    if ((int64_t)blockIdx.x >= ntiles) return;
    for (int i = threadIdx.x; i < FULL_LDS_WORDS; i += EXP_BLOCK) sh.full[i] = i < fwords ? full[i] : 0u;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t lo = tile_lo[t];
        const int64_t hi = (t + 1 < ntiles) ? (int64_t)tile_lo[t + 1] : F - 1;
        const int64_t w0 = t * EXP_TILE;
        const int64_t span = hi - lo + 1;
        const bool staged = span <= CAP;
        if (staged) {
            for (int i = threadIdx.x; i < (int)span; i += EXP_BLOCK) {
                sh.lb.pref[i] = P[lo + i];
                sh.lb.base[i] = qbase[lo + i];
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
                const int64_t w = w0 + (int64_t)j * EXP_BLOCK + threadIdx.x;
                const int s = lb_locate(sh.lb, (int)span, w);
                const int64_t k = (w - (int64_t)sh.lb.pref[s]) * G;
                addr[j] = (int64_t)sh.lb.base[s] * UNIT_SLOTS + k;
                nval[j] = w < W ? (int)min((int64_t)G, (int64_t)sh.deg[s] - k) : 0;
            }
        } else {   // slice wider than LDS (runs of zero-degree frontier vertices)
            for (int j = 0; j < EXP_IPT; j++) {
                const int64_t w = w0 + (int64_t)j * EXP_BLOCK + threadIdx.x;
                addr[j] = 0;
                nval[j] = 0;
                if (w < W) {
                    const int64_t s = upper_bound_m1<int64_t>(P, lo, hi + 1, w);
                    const int64_t k = (w - (int64_t)P[s]) * G;
                    addr[j] = (int64_t)qbase[s] * UNIT_SLOTS + k;
                    nval[j] = (int)min((int64_t)G, (int64_t)qdeg[s] - k);
                }
            }
        }
        // Phases 2 and 3 load unconditionally (a unit without work reads
        // arena[0..G), a slot without a probe gets an out-of-range buffer
        // offset) so that no load sits in a branch of its own, where the
        // compiler waits for it at once.
        // phase 2: all adjacency groups in flight at once (streamed, not kept in L2)
        Vec v[EXP_IPT];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
            v[j] = __builtin_nontemporal_load(reinterpret_cast<const Vec*>(arena + (nval[j] > 0 ? addr[j] : 0)));
        // phase 3a: all visited-bitmap probes in flight at once (n/8 bytes,
        // L2 resident), folded branch-free into one bit per candidate slot
        // (a neighbour in a word the LDS summary knows full makes no probe); on
        // direct levels the state probe alone decides
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
        uint32_t cand = 0;
        if (direct) {
#pragma unroll
            for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
                for (int e = 0; e < G; e++)
                    cand |= (uint32_t)((e < nval[j]) & !((known >> (j * G + e)) & 1u)) << (j * G + e);
        } else {
            uint32_t wb[EXP_IPT][G];
#pragma unroll
            for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
                for (int e = 0; e < G; e++)
                    wb[j][e] = probe_b32(rbm, ((e < nval[j]) && !((known >> (j * G + e)) & 1u))
                                                   ? ((uint32_t)v[j][e] >> 5) * 4u : PROBE_NONE);
#pragma unroll
            for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
                for (int e = 0; e < G; e++)
                    cand |= (uint32_t)((e < nval[j]) & !((known >> (j * G + e)) & 1u) &
                                       (((wb[j][e] >> (v[j][e] & 31)) & 1u) == 0)) << (j * G + e);
        }
        // phase 3b: candidates (not visited before this level) probe their
        // state byte; the rest make no probe
        uint8_t s[EXP_IPT][G];
#pragma unroll
        for (int j = 0; j < EXP_IPT; j++)
#pragma unroll
            for (int e = 0; e < G; e++)
                s[j][e] = probe_u8(rst, ((cand >> (j * G + e)) & 1u) ? (uint32_t)v[j][e] : PROBE_NONE);
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
// get their distance, their bitmap bit, a slot in the next frontier and a
// range of its groups.
constexpr int COL_CHUNK = 8;   // rounds whose record loads are in flight together
__global__ __launch_bounds__(BLOCK) void bfs_collect_kernel(uint32_t L, int64_t words,
                                                            const uint32_t* __restrict__ state4,
                                                            uint32_t* __restrict__ bitmap,
                                                            uint32_t* __restrict__ full,
                                                            int32_t* __restrict__ dist,
                                                            const hn_vrec* __restrict__ rec,
                                                            int32_t* __restrict__ qdeg, uint32_t* __restrict__ qbase,
                                                            uint32_t* __restrict__ P, uint32_t* __restrict__ tile_lo,
                                                            uint64_t* __restrict__ part, BfsCtl* __restrict__ ctl) {
    if ((ctl->fc[L & 1] >> 32) == 0) return;   // a level launched after the traversal ended
    const int32_t level = (int32_t)L;
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
    // The wave's discoveries are visited transposed, so that consecutive lanes
    // handle consecutive vertices: in round r lane l takes bit (l & 31) of the
    // wave's word 2r + (l >> 5), fetching that word's found bits from its owner.
    // A lane without a discovery in a round loads rec[0] (one shared line), so
    // that a chunk's record loads are all in flight before any is used.
    const int lane = __lane_id();
    const int64_t vbase = (wi - lane + (lane >> 5)) * 32 + (lane & 31);   // vertex of round 0
    const int bit = lane & 31;
    const bool any = __ballot(found != 0) != 0;
    // pass 1: the groups and adjacency entries of the wave's discoveries
    uint32_t cg = 0;
    uint64_t ce = 0;
    if (any) {
        for (int r0 = 0; r0 < WAVE / 2; r0 += COL_CHUNK) {
            bool f[COL_CHUNK];
            int32_t d[COL_CHUNK];
#pragma unroll
            for (int k = 0; k < COL_CHUNK; k++)
                f[k] = (__shfl(found, 2 * (r0 + k) + (lane >> 5), WAVE) >> bit) & 1u;
#pragma unroll
            for (int k = 0; k < COL_CHUNK; k++) d[k] = rec[f[k] ? vbase + 64 * (r0 + k) : 0].deg;
#pragma unroll
            for (int k = 0; k < COL_CHUNK; k++) {
                const uint32_t dk = f[k] ? (uint32_t)d[k] : 0u;
                cg += (dk + EXP_G - 1) / EXP_G;
                ce += dk;
            }
        }
    }
    const uint32_t wc = wave_reduce_sum((uint32_t)__popc(found));
    const uint32_t wg = wave_reduce_sum(cg);
    const uint64_t we = wave_reduce_sum(ce);
    // One atomicAdd per workgroup claims its frontier slots and groups
    // (count << 32 | groups: a level's groups stay below 2^32, so the low half
    // never carries); their entries add to the workgroup's own partial.
    __shared__ uint64_t wsum[NWAVES + 1];
    __shared__ uint64_t went[NWAVES];
    const int w = threadIdx.x / WAVE;
    if (lane == 0) {
        wsum[w] = ((uint64_t)wc << 32) | wg;
        went[w] = we;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0, acce = 0;
        for (int k = 0; k < NWAVES; k++) {
            const uint64_t x = wsum[k];
            wsum[k] = acc;
            acc += x;
            acce += went[k];
        }
        wsum[NWAVES] = acc ? atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->fc[(L + 1) & 1]),
                                       (unsigned long long)acc)
                           : 0ull;
        if (acce) part[blockIdx.x] += acce;
    }
    __syncthreads();
    if (!any) return;
    // pass 2: each discovery's distance, frontier triple and tile starts
    const uint64_t b = wsum[NWAVES] + wsum[w];
    uint32_t sb = (uint32_t)(b >> 32), gb = (uint32_t)b;
    const uint64_t lt = lanemask_lt();
    for (int r0 = 0; r0 < WAVE / 2; r0 += COL_CHUNK) {
        bool f[COL_CHUNK];
        uint64_t fm[COL_CHUNK];
        uint64_t cm = 0;
#pragma unroll
        for (int k = 0; k < COL_CHUNK; k++) {
            f[k] = (__shfl(found, 2 * (r0 + k) + (lane >> 5), WAVE) >> bit) & 1u;
            fm[k] = __ballot(f[k]);
            cm |= fm[k];
        }
        if (!cm) continue;
        hn_vrec rv[COL_CHUNK];
#pragma unroll
        for (int k = 0; k < COL_CHUNK; k++) rv[k] = rec[f[k] ? vbase + 64 * (r0 + k) : 0];
#pragma unroll
        for (int k = 0; k < COL_CHUNK; k++) {
            const uint32_t ng = f[k] ? ((uint32_t)rv[k].deg + EXP_G - 1) / EXP_G : 0u;
            const uint32_t inc = wave_inclusive_scan(ng);
            if (f[k]) {
                const int64_t v = vbase + 64 * (r0 + k);
                const uint32_t slot = sb + (uint32_t)__popcll(fm[k] & lt);
                const uint64_t g0 = (uint64_t)gb + inc - ng;
                dist[v] = level;
                qdeg[slot] = rv[k].deg;
                qbase[slot] = (uint32_t)(rv[k].off / UNIT_SLOTS);
                P[slot] = (uint32_t)g0;
                for (uint64_t t = (g0 + EXP_TILE - 1) / EXP_TILE; t * EXP_TILE < g0 + ng; t++)
                    tile_lo[t] = slot;
            }
            sb += (uint32_t)__popcll(fm[k]);
            gb += (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
        }
    }
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_bfs_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&bfs_init_kernel)) != hipSuccess) (void)hipGetLastError();
}

// Lays the buffers of a traversal of a graph with nv vertices and ne edges
// out from `base` (nullptr: sizes only); returns the bytes they take.
int64_t bfs_layout(int64_t nv_in, int64_t ne, char* base, BfsBufs* b) {
    const int64_t nv = nv_in > 0 ? nv_in : 1;
    const int64_t words = (nv + 31) / 32;
    const int64_t tiles = (ne / EXP_G + nv) / EXP_TILE + 2;
    auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
    const int64_t o_bitmap = 0;
    const int64_t o_full = al(o_bitmap + words * 4);
    const int64_t o_qdeg = al(o_full + (words + 31) / 32 * 4);
    const int64_t o_qbase = al(o_qdeg + nv * 4);
    const int64_t o_P = al(o_qbase + nv * 4);
    const int64_t o_tile = al(o_P + nv * 4);
    const int64_t o_state = al(o_tile + tiles * 4);
    const int64_t o_part = al(o_state + words * 32);
    const int64_t o_ctl = al(o_part + ceil_div(words, BLOCK) * 8);
    const int64_t total = al(o_ctl + (int64_t)sizeof(BfsCtl));
    if (base) {
        b->bitmap = reinterpret_cast<uint32_t*>(base + o_bitmap);
        b->full = reinterpret_cast<uint32_t*>(base + o_full);
        b->qdeg = reinterpret_cast<int32_t*>(base + o_qdeg);
        b->qbase = reinterpret_cast<uint32_t*>(base + o_qbase);
        b->P = reinterpret_cast<uint32_t*>(base + o_P);
        b->tile_lo = reinterpret_cast<uint32_t*>(base + o_tile);
        b->state = reinterpret_cast<uint8_t*>(base + o_state);
        b->part = reinterpret_cast<uint64_t*>(base + o_part);
        b->ctl = reinterpret_cast<BfsCtl*>(base + o_ctl);
    }
    return total;
}

// The level loop over laid-out buffers: two launches per level, queued
// without a host round trip.  The host reads the control block after as many
// levels as the last traversal of this graph took, then after 1, 2, 4, 4, ...
// more (levels queued past the end change nothing), and ends when the
// frontier of the next level to launch is empty.  Nothing here allocates (kernel timing,
// which creates its events, only with `may_time`).
hn_status bfs_run(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats, hipStream_t st,
                  const BfsBufs& b, bool may_time, uint64_t* host, int32_t max_levels) {
    const int64_t nv = g->nv;
    const int64_t words = (nv + 31) / 32;
    // frontier adjacency offsets are kept in 128-byte units as 32-bit words
    // (2^32 units = 512 GiB of arena, more than one device holds)
    HN_ARG(g->cap_units < ((int64_t)1 << 32), "edge arena beyond 2^32 128-byte units");
    // kernels only (no memset node), so a captured copy of this sequence replays correctly
    bfs_init_kernel<<<(unsigned)std::min<int64_t>(ceil_div(nv, BLOCK * 8), 4096), BLOCK, 0, st>>>(
        g->nv, g->rec, dist, b.bitmap, b.full, reinterpret_cast<uint32_t*>(b.state), source, b.qdeg, b.qbase, b.P,
        b.tile_lo, b.part, b.ctl);
    HN_CHECK_LAUNCH();
    const int64_t grid_cap = (int64_t)num_cus() * BFS_GRID_PER_CU * BLOCK / EXP_BLOCK;
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
    if (max_levels > 0) next = std::min<int64_t>(next, max_levels);
    int64_t ramp = 1, kns = 0;
    uint32_t L = 0;   // levels launched
    uint32_t levels_seen = 0;
    BfsCtl c{};
    while (rc == HN_OK) {
        hipError_t e = hipSuccess;
        for (int64_t k = 0; k < next && e == hipSuccess; k++) {
            L++;
            // timed: the events take the dispatch's own start and end
            // (hipExtLaunchKernel), with no marker between the launches
            hipExtLaunchKernelGGL(bfs_expand_kernel<EXP_G>, dim3((unsigned)grid_cap), dim3(EXP_BLOCK), 0, st,
                                  timed ? ev[2 * k] : nullptr, timed ? ev[2 * k + 1] : nullptr, 0, L,
                                  (const uint32_t*)b.P, (const uint32_t*)b.tile_lo, (const int32_t*)b.qdeg,
                                  (const uint32_t*)b.qbase, b.ctl, (const int32_t*)g->arena, b.state,
                                  (const uint32_t*)b.bitmap, (const uint32_t*)b.full, (int64_t)words,
                                  (int64_t)((words + 31) / 32),
                                  (const uint64_t*)b.part, (int64_t)ceil_div(words, BLOCK), (int64_t)g->ne);
            e = hipGetLastError();
            if (e == hipSuccess) {
                bfs_collect_kernel<<<(unsigned)ceil_div(words, BLOCK), BLOCK, 0, st>>>(
                    L, words, reinterpret_cast<const uint32_t*>(b.state), b.bitmap, b.full, dist, g->rec, b.qdeg,
                    b.qbase, b.P, b.tile_lo, b.part, b.ctl);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) e = hipMemcpyAsync(host, b.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess) std::memcpy(&c, host, sizeof(BfsCtl));
        // done: the frontier of the next level to launch is empty, or the
        // level limit is reached
        const bool done = (c.fc[(L + 1) & 1] >> 32) == 0 || (max_levels > 0 && (int64_t)L >= max_levels);
        if (timed && e == hipSuccess)
            // only the expansions of non-empty levels count
            for (uint32_t k = 0; k < c.levels - levels_seen && e == hipSuccess; k++) {
                float ms = 0.f;
                e = hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
                kns += (int64_t)((double)ms * 1e6);
            }
        if (e != hipSuccess) {
            set_error(std::string("bfs level: ") + hipGetErrorString(e));
            rc = HN_E_HIP;
            break;
        }
        levels_seen = c.levels;
        if (done) break;
        if (L > nv + 1) {   // cannot happen: every level reaches a new vertex
            set_error("bfs: level loop did not end");
            rc = HN_E_HIP;
            break;
        }
        next = ramp;
        ramp = std::min<int64_t>(2 * ramp, 4);
        if (max_levels > 0) next = std::min<int64_t>(next, max_levels - (int64_t)L);
    }
    if (rc != HN_OK) return rc;
    // the first poll of the next full traversal follows this one's depth
    if (max_levels <= 0) g->bfs_last_levels.store((int)levels_seen, std::memory_order_relaxed);
    if (stats) {
        stats->levels = (int32_t)c.levels;
        stats->visited = (int64_t)c.visited;
        stats->edges_scanned = (int64_t)c.scanned[c.levels & 1];
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
