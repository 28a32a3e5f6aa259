// msbfs.hip -- batched multi-source BFS: up to 64 sources per launch.
//
// The reference runs one BfsTopDown per source (TopDown.cu:94-106), paying a
// full traversal of the graph per query.  Here up to 64 sources share one
// traversal: every vertex holds three 64-bit masks (visited, frontier, next),
// bit b standing for source b, so one adjacency read serves every source
// whose frontier touches it.
//
// One level step, captured once into a hipGraph and replayed until the
// frontiers empty (the host reads a 4-byte "done" flag every few levels):
//   decide   one workgroup reduces the previous collect's per-block counters
//            and picks the direction: push while frontier degree * alpha <
//            degree of the vertices still missing a source, else pull;
//   list     a scan of per-word popcounts of the frontier (push) or
//            light-open (pull) bitset and an emit kernel turn the bitset
//            into a compact work list carrying (vertex, degree, adjacency
//            offset, mask) -- writes transposed so lanes write consecutive
//            entries;
//   scan     a scan of the list's 16-byte adjacency groups whose element
//            count is read on the device (so it stays inside the graph),
//            writing the group prefix and the tile partition;
//   expand   edge-granular, the single-source BFS's load balancing: a lane
//            loads four adjacency slots at once (every block is a power of
//            two of >= 32 slots, 128-byte aligned) and probes the four
//            neighbours' masks together.
//              push: add = frontier(u) & ~visited(w), ORed into next(w) with
//                    a 64-bit atomicOr when non-zero;
//              pull: an open vertex u ORs its neighbours' frontier masks
//                    (restricted to the bits it misses); a tile reduces each
//                    vertex's bits with LDS atomics and stores them, with a
//                    global atomicOr only for the two vertices it shares with
//                    neighbouring tiles;
//   heavy    pull only: open vertices of degree > T take one wave each and
//            stop as soon as every missing bit is found (T is tuned);
//   collect  one lane per vertex: new bits -> visited, frontier (masked by
//            the per-source depth limits) and per-source distances (source
//            major, int32, INT32_MAX unreached; written bit by bit so lanes
//            store consecutive words), the three bitsets of the next level
//            (frontier, light-open, heavy-open) from ballots, and per-block
//            counters.  No atomics.
// Sources may carry a depth limit (the serving loop's 2-hop queries use 2).
#include <algorithm>
#include <cstring>
#include <vector>

#include <type_traits>

#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int32_t INF = 0x7fffffff;
constexpr int MS_G = 4;                          // adjacency slots per work unit (16 bytes)
// Expansion tiles hold BLOCK * IPT work units; IPT (units per lane) is a
// load-balancing parameter the tuner picks per graph from MS_IPTS.
constexpr int MS_IPT_DEFAULT = 4;
constexpr int MS_IPT_MIN = 2;                    // buffers are sized for the most tiles
constexpr int MS_IPTS[] = {2, 4, 6};
constexpr int MS_COLLECT_PER_CU = 8;             // collect workgroups per CU
constexpr int32_t MS_NO_HEAVY = 1 << 30;         // T: every open vertex goes through the edge-granular pull
static_assert(UNIT_SLOTS % MS_G == 0, "groups must not straddle a block start");

struct MsState {
    int32_t level;                  // level being built (sources are level 0)
    int32_t mode;                   // 0 push, 1 pull
    int32_t done;
    int32_t levels_run;
    uint32_t wn;                    // entries in this level's work list
    uint32_t pad;
    uint64_t tot;                   // group scan total: edges << 32 | groups
    unsigned long long active;      // vertices with a non-empty frontier
    unsigned long long active_deg;  // sum of their degrees
    unsigned long long open_deg;    // sum of degrees of vertices still missing a bit
    unsigned long long edges;       // sum over sources of degrees of reached vertices
    uint64_t depth_mask[64];        // per level: bits whose depth limit allows expanding
    uint64_t all;                   // bits in use
    uint64_t live;                  // bits with a non-empty frontier (OR over the frontier)
    unsigned long long push_levels;
    unsigned long long pull_levels;
};

struct MsParams {
    int32_t T;          // pull: open vertices of degree > T take one wave each
    int32_t alpha;      // push while active_deg * alpha < open_deg
    int32_t ipt;        // expansion work units per lane per tile (one of MS_IPTS)
    int32_t push_only;  // adjacency not symmetric: pulling over out-edges would be wrong
    int32_t digest;     // no distance matrix: per-source digests only
    int32_t depth_limit[64];
};

__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v |= __shfl_xor(v, d, WAVE);
    return v;
}

// The 64 x 64 bit matrix whose row l is lane l's word, transposed across the
// wave: lane b returns column b (bit l = bit b of lane l's word).  Each step
// swaps the off-diagonal j x j blocks of every 2j x 2j block: the upper lanes
// of a pair keep their left columns and take the partner's left columns into
// their right ones, the lower lanes the reverse.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t x) {
    const int lane = __lane_id();
    constexpr uint64_t M[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                               0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int j = 32 >> k;
        const uint64_t m = M[k];
        const uint64_t p = __shfl_xor(x, j, WAVE);
        x = (lane & j) ? ((x & ~m) | ((p & ~m) >> j)) : ((x & m) | ((p & m) << j));
    }
    return x;
}

// Bits that can still be found: sources whose frontier is non-empty (the
// frontier is already cut by the depth limits).  A source that stops (depth
// limit, or its component is exhausted) no longer keeps every vertex it did
// not reach on the pull list.
constexpr int MS_NPART = 5;   // per collect block: active, active degree, open degree, edges, OR of frontier

__global__ void ms_init_kernel(int64_t nv, uint64_t* __restrict__ visited, uint64_t* __restrict__ frontier,
                               uint64_t* __restrict__ next, int64_t nbits, uint32_t* __restrict__ bits) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
        visited[v] = 0;
        frontier[v] = 0;
        next[v] = 0;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbits; i += stride) bits[i] = 0;
}

__global__ void ms_fill_dist_kernel(int64_t total, int32_t* __restrict__ dist) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n4 = ((uintptr_t)dist & 15) == 0 ? total / 4 : 0;
    int4* d4 = reinterpret_cast<int4*>(dist);
    const int4 inf4 = make_int4(INF, INF, INF, INF);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) d4[i] = inf4;
    for (int64_t i = n4 * 4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) dist[i] = INF;
}

// one wave: lane b seeds source b; lane L builds the depth mask of level L
__global__ __launch_bounds__(WAVE) void ms_seed_kernel(int32_t k, const int32_t* __restrict__ src, int64_t nv,
                                                       uint64_t* __restrict__ visited, uint64_t* __restrict__ frontier,
                                                       uint32_t* __restrict__ fbits, int32_t* __restrict__ dist,
                                                       const hn_vrec* __restrict__ rec, MsState* S, MsParams P) {
    const int b = threadIdx.x;
    unsigned long long deg = 0, first = 0;
    if (b < k) {
        const int32_t s = src[b];
        atomicOr(reinterpret_cast<unsigned long long*>(visited + s), 1ull << b);
        atomicOr(reinterpret_cast<unsigned long long*>(frontier + s), 1ull << b);
        atomicOr(fbits + (s >> 5), 1u << (s & 31));
        if (dist) dist[(int64_t)b * nv + s] = 0;
        deg = (unsigned long long)rec[s].deg;
        first = 1;
        for (int c = 0; c < b; c++)
            if (src[c] == s) first = 0;
    }
    uint64_t m = 0;
    for (int c = 0; c < k; c++)
        if (P.depth_limit[c] <= 0 || b < P.depth_limit[c]) m |= 1ull << c;
    S->depth_mask[b] = m;
    const unsigned long long edges = wave_reduce_sum(deg);
    const unsigned long long act = wave_reduce_sum(first);
    const unsigned long long adeg = wave_reduce_sum(first * deg);
    const uint64_t all = (uint64_t)__ballot(b < k);
    if (b == 0) {
        S->all = all;
        S->live = all;
        S->level = 0;
        S->mode = 0;
        S->done = 0;
        S->levels_run = 0;
        S->wn = 0;
        S->tot = 0;
        S->edges = edges;
        S->push_levels = S->pull_levels = 0;
        S->active = act;
        S->active_deg = adeg;
        S->open_deg = ~0ull >> 1;   // the first expansion is always a push
    }
}

// one workgroup: counters of the previous collect -> direction of this level
__global__ __launch_bounds__(BLOCK) void ms_decide_kernel(MsState* S, MsParams P,
                                                          const unsigned long long* __restrict__ part,
                                                          int32_t nparts) {
    // every thread reads the state before thread 0 changes it
    const int32_t done = S->done, level = S->level;
    __syncthreads();
    if (done) return;
    __shared__ unsigned long long red[MS_NPART][NWAVES];
    if (level > 0) {
        unsigned long long c[MS_NPART] = {0, 0, 0, 0, 0};
        for (int i = threadIdx.x; i < nparts; i += BLOCK) {
#pragma unroll
            for (int j = 0; j < 4; j++) c[j] += part[(int64_t)i * MS_NPART + j];
            c[4] |= part[(int64_t)i * MS_NPART + 4];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) c[j] = wave_reduce_sum(c[j]);
        c[4] = wave_or64(c[4]);
        if (__lane_id() == 0)
#pragma unroll
            for (int j = 0; j < MS_NPART; j++) red[j][threadIdx.x / WAVE] = c[j];
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t[MS_NPART] = {0, 0, 0, 0, 0};
            for (int w = 0; w < NWAVES; w++) {
                for (int j = 0; j < 4; j++) t[j] += red[j][w];
                t[4] |= red[4][w];
            }
            S->active = t[0];
            S->active_deg = t[1];
            S->open_deg = t[2];
            S->edges += t[3];
            S->live = t[4];
        }
    }
    if (threadIdx.x != 0) return;
    if (S->active == 0) {
        S->done = 1;
        return;
    }
    S->level += 1;
    S->levels_run += 1;
    const bool push =
        P.push_only || (unsigned long long)S->active_deg * (unsigned long long)P.alpha < S->open_deg;
    S->mode = push ? 0 : 1;
    if (push) S->push_levels++; else S->pull_levels++;
}

// popcount of the word of this level's list bitset
struct ListWordIn {
    const MsState* S;
    const uint32_t* fbits;
    const uint32_t* obits;
    __device__ __forceinline__ uint32_t operator()(int64_t w) const {
        if (S->done) return 0u;
        return (uint32_t)__popc(S->mode == 0 ? fbits[w] : obits[w]);
    }
};

// bitset -> work list (vertex, degree, offset, mask); lane per bitset word,
// writes transposed so that consecutive lanes write consecutive entries
__global__ __launch_bounds__(BLOCK) void ms_emit_kernel(int64_t words, const MsState* S,
                                                        const uint32_t* __restrict__ fbits,
                                                        const uint32_t* __restrict__ obits,
                                                        const uint32_t* __restrict__ wpre,
                                                        const hn_vrec* __restrict__ rec,
                                                        const uint64_t* __restrict__ visited,
                                                        const uint64_t* __restrict__ frontier,
                                                        const uint8_t* __restrict__ dirty,
                                                        int32_t* __restrict__ list, int32_t* __restrict__ qdeg,
                                                        int64_t* __restrict__ qoff, uint64_t* __restrict__ qmask) {
    if (S->done) return;
    const bool push = S->mode == 0;
    const uint64_t live = S->live;
    const int64_t wi = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t found = wi < words ? (push ? fbits[wi] : obits[wi]) : 0u;
    if (__ballot(found != 0) == 0) return;
    const uint32_t b = wi < words ? wpre[wi] : 0u;
    const int lane = __lane_id();
    const int64_t wbase = wi - lane;
    const int bit = lane & 31;
    const uint32_t below = (1u << bit) - 1u;
    for (int r = 0; r < WAVE / 2; r++) {
        const int owner = 2 * r + (lane >> 5);
        const uint32_t fw = __shfl(found, owner, WAVE);
        if (__ballot(fw != 0) == 0) continue;
        const uint32_t bo = __shfl(b, owner, WAVE);
        if ((fw >> bit) & 1u) {
            const int32_t v = (int32_t)((wbase + owner) * 32 + bit);
            const uint32_t pos = bo + (uint32_t)__popc(fw & below);
            const hn_vrec rv = rec[v];
            const uint64_t m = push ? frontier[v] : (live & ~visited[v]);
            list[pos] = v;
            qdeg[pos] = m ? rv.deg : 0;
            qoff[pos] = rv.off;
            qmask[pos] = m;
        }
    }
}

struct GroupIn {
    const int32_t* qdeg;
    __device__ __forceinline__ uint64_t operator()(int64_t i) const {
        const uint32_t d = (uint32_t)qdeg[i];
        return ((uint64_t)d << 32) | ((d + MS_G - 1) / MS_G);
    }
};

struct GroupOut {
    uint32_t* P;
    uint32_t* tile_lo;
    uint32_t tile;   // work units per expansion tile
    __device__ __forceinline__ void operator()(int64_t i, uint64_t ex, uint64_t x) const {
        const uint32_t g0 = (uint32_t)ex, ng = (uint32_t)x;
        P[i] = g0;
        if (ng) {
            const uint32_t t0 = (uint32_t)(((uint64_t)g0 + tile - 1) / tile);
            const uint32_t t1 = (uint32_t)(((uint64_t)g0 + ng + tile - 1) / tile);
            for (uint32_t t = t0; t < t1; t++) tile_lo[t] = (uint32_t)i;
        }
    }
};

template <int IPT>
struct MsShared {
    static constexpr int CAP = LbShared<IPT, uint32_t>::CAP;
    LbShared<IPT, uint32_t> lb;      // pref = group prefix, base = adjacency offset
    int32_t deg[CAP];
    uint64_t mask[CAP];              // push: frontier mask; pull: missing bits
    unsigned long long acc[CAP];     // pull: bits found this tile
    uint64_t crange[CAP];            // window: range of the vertex's changed pairs (lo << 32 | hi), 0 if none
};

// Multi-version window (served stream): the changed pairs of the window,
// sorted by (u << sbits | w), with cS = the sources whose graph version holds
// the pair.  A vertex flagged dirty has some; its range in the sorted list:
__device__ __forceinline__ uint64_t changed_range(const uint64_t* __restrict__ ckey, uint32_t n, int32_t u,
                                                  int sbits) {
    const uint64_t k0 = (uint64_t)(uint32_t)u << sbits, k1 = (uint64_t)((uint32_t)u + 1) << sbits;
    uint32_t a = 0, b = n;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (ckey[m] < k0) a = m + 1; else b = m;
    }
    uint32_t c = a, d = n;
    while (c < d) {
        const uint32_t m = (c + d) >> 1;
        if (ckey[m] < k1) c = m + 1; else d = m;
    }
    return c > a ? ((uint64_t)a << 32) | c : 0ull;
}

// Edge-granular expansion of this level's work list (see the file comment).
template <bool PULL, int IPT>
__global__ __launch_bounds__(BLOCK) void ms_expand_kernel(const MsState* S, const uint32_t* __restrict__ P,
                                                          const uint32_t* __restrict__ tile_lo,
                                                          const int32_t* __restrict__ qdeg,
                                                          const int64_t* __restrict__ qoff,
                                                          const uint64_t* __restrict__ qmask,
                                                          const int32_t* __restrict__ list,
                                                          const int32_t* __restrict__ arena,
                                                          const uint64_t* __restrict__ probe,
                                                          unsigned long long* __restrict__ next,
                                                          const uint8_t* __restrict__ dirty,
                                                          const uint64_t* __restrict__ ckey,
                                                          const uint64_t* __restrict__ cS,
                                                          const uint32_t* __restrict__ nchg, int sbits) {
    if (S->done || S->mode != (PULL ? 1 : 0)) return;
    typedef int32_t Vec __attribute__((ext_vector_type(MS_G)));
    __shared__ MsShared<IPT> sh;
    constexpr int TILE = BLOCK * IPT;
    const uint32_t nc = *nchg;
    const int64_t F = S->wn;
    const int64_t W = (uint32_t)S->tot;
    const int64_t ntiles = (W + TILE - 1) / TILE;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t lo = tile_lo[t];
        const int64_t hi = (t + 1 < ntiles) ? (int64_t)tile_lo[t + 1] : F - 1;
        const int64_t w0 = t * TILE;
        const int64_t span = hi - lo + 1;
        const bool staged = span <= MsShared<IPT>::CAP;
        if (staged) {
            for (int i = threadIdx.x; i < (int)span; i += BLOCK) {
                sh.lb.pref[i] = P[lo + i];
                sh.lb.base[i] = qoff[lo + i];
                sh.deg[i] = qdeg[lo + i];
                sh.mask[i] = qmask[lo + i];
                if (PULL) sh.acc[i] = 0;
                const int32_t u = list[lo + i];
                sh.crange[i] = (nc && dirty[u]) ? changed_range(ckey, nc, u, sbits) : 0ull;
            }
        }
        __syncthreads();
        // phase 1: every unit's segment, slot address, valid count and mask
        int64_t addr[IPT];
        int nval[IPT];
        int64_t seg[IPT];
        uint64_t msk[IPT];
        uint64_t crg[IPT];
        if (staged) {
#pragma unroll
            for (int j = 0; j < IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                const int s = lb_locate<IPT>(sh.lb, (int)span, w);
                const int64_t k = (w - (int64_t)sh.lb.pref[s]) * MS_G;
                addr[j] = sh.lb.base[s] + k;
                nval[j] = w < W ? (int)min((int64_t)MS_G, (int64_t)sh.deg[s] - k) : 0;
                seg[j] = s;
                msk[j] = sh.mask[s];
                crg[j] = sh.crange[s];
            }
        } else {   // slice wider than LDS (runs of work-free list entries)
            for (int j = 0; j < IPT; j++) {
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                addr[j] = 0;
                nval[j] = 0;
                seg[j] = lo;
                msk[j] = 0;
                crg[j] = 0;
                if (w < W) {
                    const int64_t s = upper_bound_m1<int64_t>(P, lo, hi + 1, w);
                    const int64_t k = (w - (int64_t)P[s]) * MS_G;
                    addr[j] = qoff[s] + k;
                    nval[j] = (int)min((int64_t)MS_G, (int64_t)qdeg[s] - k);
                    seg[j] = s;
                    msk[j] = qmask[s];
                    const int32_t u = list[s];
                    crg[j] = (nc && dirty[u]) ? changed_range(ckey, nc, u, sbits) : 0ull;
                }
            }
        }
        // phase 2: adjacency groups, all in flight (a unit without work reads arena[0..4))
        Vec v[IPT];
#pragma unroll
        for (int j = 0; j < IPT; j++)
            v[j] = __builtin_nontemporal_load(reinterpret_cast<const Vec*>(arena + (nval[j] > 0 ? addr[j] : 0)));
        // phase 3: neighbour masks, all in flight (masked slots read probe[0])
        uint64_t pr[IPT][MS_G];
#pragma unroll
        for (int j = 0; j < IPT; j++)
#pragma unroll
            for (int e = 0; e < MS_G; e++) pr[j][e] = probe[e < nval[j] ? v[j][e] : 0];
        // window: a changed pair carries only the sources whose version holds
        // it (push: the others count as visited; pull: their bits are dropped)
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            if (!crg[j]) continue;
            const uint32_t clo = (uint32_t)(crg[j] >> 32), chi = (uint32_t)crg[j];
            const uint64_t ub = ckey[clo] & ~((1ull << sbits) - 1);   // u << sbits
#pragma unroll
            for (int e = 0; e < MS_G; e++) {
                if (e >= nval[j]) continue;
                const uint64_t key = ub | (uint32_t)v[j][e];
                uint32_t a = clo, b = chi;
                while (a < b) {
                    const uint32_t m = (a + b) >> 1;
                    if (ckey[m] < key) a = m + 1; else b = m;
                }
                if (a < chi && ckey[a] == key) pr[j][e] = PULL ? (pr[j][e] & cS[a]) : (pr[j][e] | ~cS[a]);
            }
        }
        if (!PULL) {
            // push: bits of the frontier vertex its neighbour has not seen
#pragma unroll
            for (int j = 0; j < IPT; j++)
#pragma unroll
                for (int e = 0; e < MS_G; e++) {
                    const uint64_t add = msk[j] & ~pr[j][e];
                    if (e < nval[j] && add) atomicOr(next + v[j][e], (unsigned long long)add);
                }
        } else {
            // pull: bits the open vertex misses, found in its neighbours' frontiers
#pragma unroll
            for (int j = 0; j < IPT; j++) {
                uint64_t a = 0;
#pragma unroll
                for (int e = 0; e < MS_G; e++) a |= e < nval[j] ? pr[j][e] : 0ull;
                a &= msk[j];
                if (a) {
                    if (staged) atomicOr(&sh.acc[seg[j]], (unsigned long long)a);
                    else atomicOr(next + list[seg[j]], (unsigned long long)a);
                }
            }
            __syncthreads();
            if (staged) {
                for (int i = threadIdx.x; i < (int)span; i += BLOCK) {
                    const unsigned long long a = sh.acc[i];
                    if (!a) continue;
                    const int32_t u = list[lo + i];
                    // the first and last vertex may share their groups with
                    // neighbouring tiles; the others belong to this tile only
                    if (i == 0 || i == (int)span - 1) atomicOr(next + u, a);
                    else next[u] = a;
                }
            }
        }
        __syncthreads();
    }
}

// pull, heavy open vertices (degree > T): one wave each, 256 slots per step,
// stopping once every missing bit is found
__global__ __launch_bounds__(BLOCK) void ms_pull_heavy_kernel(int64_t words, const MsState* S,
                                                              const uint32_t* __restrict__ hbits,
                                                              const hn_vrec* __restrict__ rec,
                                                              const int32_t* __restrict__ arena,
                                                              const uint64_t* __restrict__ frontier,
                                                              const uint64_t* __restrict__ visited,
                                                              const uint8_t* __restrict__ dirty,
                                                              uint64_t* __restrict__ next) {
    if (S->done || S->mode != 1) return;
    typedef int32_t Vec __attribute__((ext_vector_type(MS_G)));
    const uint64_t live = S->live;
    const int64_t wid = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const int64_t nw = (int64_t)gridDim.x * BLOCK / WAVE;
    const int l = __lane_id();
    for (int64_t wd = wid; wd < words; wd += nw) {
        uint32_t hw = hbits[wd];
        while (hw) {
            const int j = __ffs(hw) - 1;
            hw &= hw - 1;
            const int32_t u = (int32_t)(wd * 32 + j);
            if (dirty[u]) continue;   // dirty vertices take the masked grouped path
            const uint64_t miss = live & ~visited[u];
            if (!miss) continue;
            const hn_vrec r = rec[u];
            uint64_t acc = 0;
            for (int64_t j0 = 0; j0 < r.deg; j0 += (int64_t)WAVE * MS_G) {
                const int64_t k = j0 + (int64_t)l * MS_G;
                const int n = (int)min((int64_t)MS_G, (int64_t)r.deg - k);
                const Vec nb = __builtin_nontemporal_load(reinterpret_cast<const Vec*>(arena + (n > 0 ? r.off + k : 0)));
                uint64_t f = 0;
#pragma unroll
                for (int e = 0; e < MS_G; e++) {
                    const uint64_t x = frontier[e < n ? nb[e] : 0];
                    f |= e < n ? x : 0ull;
                }
                acc |= wave_or64(f & miss);
                if ((acc & miss) == miss) break;
            }
            if (l == 0 && acc) next[u] = acc;
        }
    }
}

// collect: new bits -> visited, frontier, distances (or, in digest mode,
// this block's per-source reached counts); next level's bitsets and this
// block's counters (active, active degree, open degree, edges)
__global__ __launch_bounds__(BLOCK) void ms_collect_kernel(int64_t nv, const int32_t* __restrict__ degs,
                                                           uint64_t* __restrict__ visited,
                                                           uint64_t* __restrict__ frontier,
                                                           uint64_t* __restrict__ next, int32_t* __restrict__ dist,
                                                           const MsState* S, int32_t T, const uint8_t* __restrict__ dirty,
                                                           uint32_t* __restrict__ fbits,
                                                           uint32_t* __restrict__ obits, uint32_t* __restrict__ hbits,
                                                           unsigned long long* __restrict__ part, int32_t digest,
                                                           unsigned long long* __restrict__ dgp) {
    if (S->done) return;
    __shared__ unsigned int dcnt[64];
    if (threadIdx.x < 64) dcnt[threadIdx.x] = 0;
    __syncthreads();
    const int32_t level = S->level;
    const uint64_t dmn = S->depth_mask[level < 63 ? level : 63];
    // bits still expandable after this level: a superset of the next live set
    const uint64_t live = S->live & dmn;
    const int lane = __lane_id();
    unsigned long long act = 0, adeg = 0, odeg = 0, edges = 0, fany = 0;
    for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < nv; base += (int64_t)gridDim.x * BLOCK) {
        const int64_t u = base + threadIdx.x;
        const bool valid = u < nv;
        uint64_t vis = 0, nx = 0, fo = 0;
        if (valid) {
            vis = visited[u];
            nx = next[u];
            fo = frontier[u];
        }
        const uint64_t nb = nx & ~vis;
        if (nx) next[u] = 0;
        const uint64_t visn = vis | nb;
        if (nb) visited[u] = visn;
        const uint64_t f = nb & dmn;
        if (f != fo) frontier[u] = f;
        const uint64_t miss = valid ? (live & ~visn) : 0ull;
        const int32_t deg = (f | miss | nb) ? degs[u] : 0;
        const bool inF = f != 0;
        // a vertex with changed pairs in a window takes the edge-granular pull
        const bool heavy = miss != 0 && deg > T && !dirty[u];
        const bool inO = miss != 0 && deg > 0 && !heavy;
        const bool inH = heavy;
        act += inF ? 1ull : 0ull;
        adeg += inF ? (unsigned long long)deg : 0ull;
        odeg += miss ? (unsigned long long)deg : 0ull;
        edges += (unsigned long long)deg * (unsigned long long)__popcll(nb);
        fany |= f;
        // distances, one source bit at a time so that lanes store consecutive
        // words; in digest mode one count per source bit and wave
        if (digest) {
            // lane b counts the wave's vertices newly reached from source b
            const unsigned int c = (unsigned int)__popcll((unsigned long long)wave_transpose64(nb));
            if (c) atomicAdd(&dcnt[lane], c);
        } else {
            uint64_t M = wave_or64(nb);
            while (M) {
                const int b = __ffsll((unsigned long long)M) - 1;
                M &= M - 1;
                if ((nb >> b) & 1ull) dist[(int64_t)b * nv + u] = level;
            }
        }
        // next level's bitsets: one 64-bit ballot = two bitset words
        const uint64_t bf = (uint64_t)__ballot(inF), bo = (uint64_t)__ballot(inO), bh = (uint64_t)__ballot(inH);
        const int64_t wv = base + (threadIdx.x & ~(WAVE - 1));   // first vertex of this wave
        if (lane == 0 && wv < nv) {
            const int64_t wq = wv / 64;                            // 64-vertex word pair
            reinterpret_cast<uint64_t*>(fbits)[wq] = bf;
            reinterpret_cast<uint64_t*>(obits)[wq] = bo;
            reinterpret_cast<uint64_t*>(hbits)[wq] = bh;
        }
    }
    __shared__ unsigned long long red[MS_NPART][NWAVES];
    unsigned long long c[MS_NPART] = {act, adeg, odeg, edges, fany};
#pragma unroll
    for (int j = 0; j < 4; j++) c[j] = wave_reduce_sum(c[j]);
    c[4] = wave_or64(c[4]);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < MS_NPART; j++) red[j][threadIdx.x / WAVE] = c[j];
    __syncthreads();
    if (threadIdx.x < MS_NPART) {
        unsigned long long t = 0;
        for (int w = 0; w < NWAVES; w++) t = threadIdx.x == 4 ? (t | red[4][w]) : t + red[threadIdx.x][w];
        part[(int64_t)blockIdx.x * MS_NPART + threadIdx.x] = t;
    }
    // digest mode: this block's slot accumulates over the levels (it always
    // covers the same vertices), so no atomics
    if (digest && threadIdx.x < 64) {
        const unsigned int cnt = dcnt[threadIdx.x];
        if (cnt) {
            unsigned long long* p = dgp + ((int64_t)blockIdx.x * 64 + threadIdx.x) * 3;
            p[0] += cnt;
            p[1] += (unsigned long long)cnt * (unsigned long long)level;
            p[2] = (unsigned long long)level;   // levels only grow
        }
    }
}

// digest[b] = {reached, distance sum, eccentricity} of source b: the source
// itself plus every collect block's slot
__global__ __launch_bounds__(BLOCK) void ms_digest_kernel(const unsigned long long* __restrict__ dgp, int64_t nparts,
                                                          unsigned long long* __restrict__ digest) {
    const int b = blockIdx.x;
    unsigned long long r = 0, s = 0, m = 0;
    for (int64_t i = threadIdx.x; i < nparts; i += BLOCK) {
        const unsigned long long* p = dgp + (i * 64 + b) * 3;
        r += p[0];
        s += p[1];
        m = p[2] > m ? p[2] : m;
    }
    r = wave_reduce_sum(r);
    s = wave_reduce_sum(s);
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        const unsigned long long o = __shfl_xor(m, d, WAVE);
        m = o > m ? o : m;
    }
    __shared__ unsigned long long red[3][NWAVES];
    if (__lane_id() == 0) {
        red[0][threadIdx.x / WAVE] = r;
        red[1][threadIdx.x / WAVE] = s;
        red[2][threadIdx.x / WAVE] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tr = 1, ts = 0, tm = 0;   // the source, at distance 0
        for (int w = 0; w < NWAVES; w++) {
            tr += red[0][w];
            ts += red[1][w];
            tm = red[2][w] > tm ? red[2][w] : tm;
        }
        digest[b * 3] = tr;
        digest[b * 3 + 1] = ts;
        digest[b * 3 + 2] = tm;
    }
}

// ---- multi-version windows (the served stream) -----------------------------
// One launch may serve queries of several graph versions: storage holds the
// newest version of the window.  A pair whose existence differs between the
// window's versions is a changed pair: the main kernels mask it per edge
// with the sources whose version holds it, and the changed pairs absent from
// storage are side records relaxed by one more kernel per level.

struct SideSV {
    uint64_t v[64];   // sources whose graph version is window version t
};

// side records: changed pairs absent from the newest version (so not in
// storage), each relaxed u -> w for its sources (the reverse pair is its own
// record, the window's batches being symmetric)
__global__ __launch_bounds__(BLOCK) void ms_side_kernel(const MsState* S, const uint32_t* __restrict__ nside,
                                                        const int2* __restrict__ uw,
                                                        const uint64_t* __restrict__ sm,
                                                        const uint64_t* __restrict__ frontier,
                                                        const uint64_t* __restrict__ visited,
                                                        unsigned long long* __restrict__ next) {
    if (S->done) return;
    const int64_t n = *nside;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const int2 e = uw[i];
        const uint64_t fu = frontier[e.x] & sm[i];
        if (!fu) continue;
        const uint64_t add = fu & ~visited[e.y];
        if (add) atomicOr(next + e.y, (unsigned long long)add);
    }
}

// window entry i: pair key << 7 | update index (1..63) << 1 | 1 for an insert
__global__ void win_pack_kernel(int64_t n, const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                                const uint8_t* __restrict__ tt, int sbits, uint64_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((((uint64_t)(uint32_t)src[i] << sbits) | (uint32_t)dst[i]) << 7) | tt[i];
}

// every entry (u, w) of update t needs (w, u) of the same kind in update t:
// side records relax both directions, which needs every version symmetric
__global__ void win_sym_kernel(int64_t n, const uint64_t* __restrict__ keys, int sbits, uint32_t* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i], pk = k >> 7, msk = (1ull << sbits) - 1;
    const uint64_t rk = ((((pk & msk) << sbits) | (pk >> sbits)) << 7) | (k & 127);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < rk) lo = mid + 1; else hi = mid;
    }
    if (lo >= n || keys[lo] != rk) *bad = 1u;
}

struct PairHeadIn {
    const uint64_t* keys;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const {
        return (i == 0 || (keys[i] >> 7) != (keys[i - 1] >> 7)) ? 1u : 0u;
    }
};

__global__ void win_seg_kernel(int64_t n, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ hpos,
                               uint32_t* __restrict__ seg_start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || (keys[i] >> 7) != (keys[i - 1] >> 7)) seg_start[hpos[i]] = (uint32_t)i;
    if (i == n - 1) seg_start[hpos[n]] = (uint32_t)n;
}

__device__ __forceinline__ bool sorted_has(const int32_t* __restrict__ a, int32_t n, int32_t x) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == x;
}

// membership of every touched pair before the window's first update
__global__ void win_e0_kernel(const uint32_t* __restrict__ nseg, const uint32_t* __restrict__ seg_start,
                              const uint64_t* __restrict__ keys, int sbits, const hn_vrec* __restrict__ rec,
                              const int32_t* __restrict__ arena, uint8_t* __restrict__ e0) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)*nseg) return;
    const uint64_t pk = keys[seg_start[j]] >> 7;
    const int32_t u = (int32_t)(pk >> sbits), w = (int32_t)(pk & ((1ull << sbits) - 1));
    const hn_vrec r = rec[u];
    e0[j] = sorted_has(arena + r.off, r.deg, w) ? 1 : 0;
}

// existence of each touched pair in window versions 0..m (update t makes
// version t; failed updates change nothing); a pair whose existence varies
// is changed and makes its source dirty
__global__ void win_exist_kernel(const uint32_t* __restrict__ nseg, const uint32_t* __restrict__ seg_start,
                                 const uint64_t* __restrict__ keys, const uint8_t* __restrict__ e0, int32_t m,
                                 uint64_t failed, int sbits, uint64_t* __restrict__ E, uint32_t* __restrict__ chg,
                                 uint8_t* __restrict__ dirty) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)*nseg) return;
    uint32_t p = seg_start[j];
    const uint32_t end = seg_start[j + 1];
    uint32_t state = e0[j];
    uint64_t e = 0;
    for (int t = 0; t <= m; t++) {
        while (p < end && (int)((keys[p] >> 1) & 63) == t) {
            if (!((failed >> t) & 1ull)) state = (uint32_t)(keys[p] & 1);
            p++;
        }
        e |= (uint64_t)state << t;
    }
    const uint64_t all = m >= 63 ? ~0ull : ((2ull << m) - 1);
    const bool changed = e != 0 && e != all;
    E[j] = e;
    chg[j] = changed ? 1u : 0u;
    if (changed) dirty[(keys[seg_start[j]] >> 7) >> sbits] = 1;
}

struct FlagIn {
    const uint32_t* f;
    const uint32_t* n;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const { return i < (int64_t)*n ? f[i] : 0u; }
};

__global__ void win_chg_write_kernel(const uint32_t* __restrict__ nseg, const uint32_t* __restrict__ chg,
                                     const uint32_t* __restrict__ cpos, const uint32_t* __restrict__ seg_start,
                                     const uint64_t* __restrict__ keys, const uint64_t* __restrict__ E,
                                     uint64_t* __restrict__ ckey, uint64_t* __restrict__ cE) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)*nseg || !chg[j]) return;
    ckey[cpos[j]] = keys[seg_start[j]] >> 7;
    cE[cpos[j]] = E[j];
}

// per changed pair: its sources, and whether it is absent from storage (a side record)
__global__ void win_cs_kernel(const uint32_t* __restrict__ nchg, const uint64_t* __restrict__ cE, int32_t m,
                              SideSV sv, uint64_t* __restrict__ cS, uint32_t* __restrict__ sflag) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)*nchg) return;
    const uint64_t e = cE[c];
    uint64_t s = 0;
    for (int t = 0; t <= m; t++)
        if ((e >> t) & 1ull) s |= sv.v[t];
    cS[c] = s;
    sflag[c] = ((e >> m) & 1ull) ? 0u : 1u;
}

__global__ void win_side_write_kernel(const uint32_t* __restrict__ nchg, const uint32_t* __restrict__ sflag,
                                      const uint32_t* __restrict__ spos, const uint64_t* __restrict__ ckey,
                                      const uint64_t* __restrict__ cS, int sbits, int2* __restrict__ uw,
                                      uint64_t* __restrict__ sm, uint32_t* __restrict__ nside) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = *nchg;
    if (c == 0) *nside = spos[n];
    if (c >= n || !sflag[c]) return;
    const uint64_t pk = ckey[c];
    uw[spos[c]] = make_int2((int32_t)(pk >> sbits), (int32_t)(pk & ((1ull << sbits) - 1)));
    sm[spos[c]] = cS[c];
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

// the engine's compact degree copy (4 bytes per vertex instead of a 16-byte record)
__global__ void ms_degrees_kernel(int64_t nv, const hn_vrec* __restrict__ rec, int32_t* __restrict__ degs) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) degs[v] = rec[v].deg;
}

// engine buffer sizes
int64_t ms_words(int64_t nv) { return (nv + 31) / 32; }
int64_t ms_bit_words(int64_t nv) { return ((nv + 63) / 64) * 2; }   // whole 64-vertex pairs
int64_t ms_tiles(int64_t nv, int64_t ne) { return (ne / MS_G + nv) / (BLOCK * MS_IPT_MIN) + 2; }
int64_t ms_collect_grid(int64_t nv) {
    return std::max<int64_t>(1, std::min<int64_t>(ceil_div(nv, BLOCK), (int64_t)num_cus() * MS_COLLECT_PER_CU));
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_msbfs_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&ms_init_kernel)) != hipSuccess) (void)hipGetLastError();
}

// Per-engine device state: masks, bitsets, work list and load-balancing
// buffers, own stream and pinned readback buffer, captured level graph.
struct MsEngine {
    hn_graph* g = nullptr;
    char* mem = nullptr;              // one allocation for everything below
    size_t mem_bytes = 0;
    uint64_t* visited = nullptr;
    uint64_t* frontier = nullptr;
    uint64_t* next = nullptr;
    uint32_t* fbits = nullptr;
    uint32_t* obits = nullptr;
    uint32_t* hbits = nullptr;
    uint32_t* wpre = nullptr;
    int32_t* list = nullptr;
    int32_t* qdeg = nullptr;
    int64_t* qoff = nullptr;
    uint64_t* qmask = nullptr;
    uint32_t* P = nullptr;
    uint32_t* tile_lo = nullptr;
    unsigned long long* part = nullptr;
    int32_t* degs = nullptr;          // degree copy, valid for graph version degs_version
    uint64_t degs_version = ~0ull;
    void* scan_tmp = nullptr;
    MsState* S = nullptr;
    int32_t* d_src = nullptr;
    int64_t nv = 0, tile_cap = 0, nparts = 0;
    hipGraphExec_t exec = nullptr;
    hipStream_t exec_stream = nullptr;
    int32_t* exec_dist = nullptr;
    const int32_t* exec_arena = nullptr;   // edge arena the captured level step reads (it moves when it grows)
    MsParams exec_params{};
    MsParams params{};
    hipStream_t own = nullptr;        // non-blocking stream the level graph is captured on
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int32_t* pinned = nullptr;        // engine-private readback buffer
    int32_t last_steps = 0;           // level steps the last unlimited launch ran until done
    // multi-version window state (served stream)
    uint8_t* dirty = nullptr;         // nv flags (vertex has changed pairs), in `mem`
    uint32_t* d_nside = nullptr;      // side record count, in `mem`
    uint32_t* d_nchg = nullptr;       // changed pair count, in `mem`
    unsigned long long* dg_part = nullptr;   // digest mode: per collect block x source {reached, sum, max}
    unsigned long long* digest = nullptr;    // digest mode: per source {reached, sum, max}
    Scratch side_uw, side_sm;         // side records (u, w) and source masks
    Scratch w_keys, w_alt, w_tmp, w_seg, w_e0, w_E, w_chg, w_ckey, w_cE, w_cS, w_sflag;
    void* exec_cS = nullptr;          // window buffers the captured level step was built with
    void* exec_ckey = nullptr;
    int64_t w_n = 0;                  // entries of the open window
    uint64_t* w_sorted = nullptr;     // its sorted keys (in w_keys or w_alt)
    bool has_window = false;
    void release_window() {
        for (Scratch* x : {&side_uw, &side_sm, &w_keys, &w_alt, &w_tmp, &w_seg, &w_e0, &w_E, &w_chg, &w_ckey,
                           &w_cE, &w_cS, &w_sflag})
            x->release();
    }
    void release_mem() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        dfree(mem);
        mem = nullptr;
        mem_bytes = 0;
    }
    void release() {
        release_mem();
        release_window();
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        if (own) (void)hipStreamDestroy(own);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        own = nullptr;
        ev_in = ev_out = nullptr;
    }
};

namespace {

// Lay out (and with `base` set, assign) the engine buffers; returns bytes.
size_t ms_layout(int64_t nv, int64_t tiles, int64_t nparts, MsEngine* e, char* base) {
    const int64_t n = nv > 0 ? nv : 1;
    const int64_t words = ms_words(n), bw = ms_bit_words(n);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) & ~size_t(255);
        return base ? base + at : nullptr;
    };
    char* p_vis = take((size_t)n * 8);
    char* p_fr = take((size_t)n * 8);
    char* p_nx = take((size_t)n * 8);
    char* p_fb = take((size_t)bw * 4 * 3);   // frontier, light-open, heavy-open: contiguous
    char* p_wp = take((size_t)words * 4);
    char* p_li = take((size_t)n * 4);
    char* p_qd = take((size_t)n * 4);
    char* p_qo = take((size_t)n * 8);
    char* p_qm = take((size_t)n * 8);
    char* p_P = take((size_t)(n + 1) * 4);
    char* p_tl = take((size_t)tiles * 4);
    char* p_pa = take((size_t)nparts * MS_NPART * 8);
    char* p_dg = take((size_t)n * 4);
    char* p_dy = take((size_t)n);
    char* p_ns = take(4);
    char* p_nc = take(4);
    char* p_dp = take((size_t)nparts * 64 * 3 * 8);
    char* p_dd = take(64 * 3 * 8);
    char* p_st = take(std::max(scan_temp_bytes(n), scan_temp_bytes(words)));
    char* p_S = take(sizeof(MsState));
    char* p_src = take(64 * 4);
    if (base) {
        e->visited = reinterpret_cast<uint64_t*>(p_vis);
        e->frontier = reinterpret_cast<uint64_t*>(p_fr);
        e->next = reinterpret_cast<uint64_t*>(p_nx);
        e->fbits = reinterpret_cast<uint32_t*>(p_fb);
        e->obits = e->fbits + bw;
        e->hbits = e->fbits + 2 * bw;
        e->wpre = reinterpret_cast<uint32_t*>(p_wp);
        e->list = reinterpret_cast<int32_t*>(p_li);
        e->qdeg = reinterpret_cast<int32_t*>(p_qd);
        e->qoff = reinterpret_cast<int64_t*>(p_qo);
        e->qmask = reinterpret_cast<uint64_t*>(p_qm);
        e->P = reinterpret_cast<uint32_t*>(p_P);
        e->tile_lo = reinterpret_cast<uint32_t*>(p_tl);
        e->part = reinterpret_cast<unsigned long long*>(p_pa);
        e->degs = reinterpret_cast<int32_t*>(p_dg);
        e->degs_version = ~0ull;
        e->dirty = reinterpret_cast<uint8_t*>(p_dy);
        e->d_nside = reinterpret_cast<uint32_t*>(p_ns);
        e->d_nchg = reinterpret_cast<uint32_t*>(p_nc);
        e->dg_part = reinterpret_cast<unsigned long long*>(p_dp);
        e->digest = reinterpret_cast<unsigned long long*>(p_dd);
        e->scan_tmp = p_st;
        e->S = reinterpret_cast<MsState*>(p_S);
        e->d_src = reinterpret_cast<int32_t*>(p_src);
    }
    return off;
}

// tile capacity with room for edge growth, so updates rarely force a resize
int64_t ms_tile_cap(int64_t nv, int64_t ne) { return ms_tiles(nv, ne + ne / 4 + 4096); }

// (re)allocate the buffers when the graph outgrew them
hn_status ms_reserve(MsEngine* e) {
    hn_graph* g = e->g;
    const int64_t need = ms_tiles(g->nv, g->ne);
    if (e->mem && e->nv == g->nv && e->tile_cap >= need) return HN_OK;
    e->release_mem();
    const int64_t cap = ms_tile_cap(g->nv, g->ne);
    const int64_t nparts = ms_collect_grid(g->nv);
    const size_t bytes = ms_layout(g->nv, cap, nparts, e, nullptr);
    char* mem = nullptr;
    HN_TRY(dmalloc(reinterpret_cast<void**>(&mem), bytes));
    e->mem = mem;
    e->mem_bytes = bytes;
    e->nv = g->nv;
    e->tile_cap = cap;
    e->nparts = nparts;
    ms_layout(g->nv, cap, nparts, e, mem);
    // no window: nothing dirty, no side records
    HN_HIP(hipMemset(e->dirty, 0, (size_t)(g->nv > 0 ? g->nv : 1)));
    HN_HIP(hipMemset(e->d_nside, 0, 4));
    HN_HIP(hipMemset(e->d_nchg, 0, 4));
    e->has_window = false;
    return HN_OK;
}

}  // namespace

int64_t ms_engine_bytes(int64_t nv, int64_t ne) {
    return (int64_t)ms_layout(nv, ms_tile_cap(nv, ne), ms_collect_grid(nv), nullptr, nullptr) + 4096;
}

// a standalone engine: buffers, state, its own stream and pinned readback
// buffer, so several engines may run on one graph from different threads
hn_status ms_engine_create(hn_graph* g, MsEngine** out) {
    MsEngine* e = new MsEngine();
    e->g = g;
    e->params.T = MS_NO_HEAVY;
    e->params.alpha = 16;
    e->params.ipt = MS_IPT_DEFAULT;
    hn_status s = ms_reserve(e);
    if (s != HN_OK) {
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

// record one level step on `st` (also used for capture); returns the first launch error
static hipError_t ms_level(MsEngine* e, int32_t* dist, const MsParams& P, hipStream_t st) {
    hn_graph* g = e->g;
    const int64_t nv = g->nv;
    const int64_t words = ms_words(nv);
    const unsigned wgrid = (unsigned)((int64_t)num_cus() * 8);
    const unsigned xgrid = (unsigned)((int64_t)num_cus() * 16);
    ms_decide_kernel<<<1, BLOCK, 0, st>>>(e->S, P, e->part, (int32_t)e->nparts);
    hipError_t err = exclusive_scan_dev(ListWordIn{e->S, e->fbits, e->obits}, ArrayOut<uint32_t>{e->wpre}, &e->S->wn,
                                        words, e->scan_tmp, st);
    if (err != hipSuccess) return err;
    ms_emit_kernel<<<(unsigned)ceil_div(words, BLOCK), BLOCK, 0, st>>>(words, e->S, e->fbits, e->obits, e->wpre,
                                                                       g->rec, e->visited, e->frontier, e->dirty,
                                                                       e->list, e->qdeg, e->qoff, e->qmask);
    err = exclusive_scan_dev_count<uint64_t>(GroupIn{e->qdeg}, GroupOut{e->P, e->tile_lo, (uint32_t)(BLOCK * P.ipt)},
                                             &e->S->tot, &e->S->wn, nv,
                                             e->scan_tmp, st);
    if (err != hipSuccess) return err;
    const int sbits = std::max(1, ceil_log2_host((int64_t)nv + 1));
    auto expand = [&](auto ipt_c) {
        constexpr int IPT = decltype(ipt_c)::value;
        ms_expand_kernel<false, IPT><<<xgrid, BLOCK, 0, st>>>(
            e->S, e->P, e->tile_lo, e->qdeg, e->qoff, e->qmask, e->list, g->arena, e->visited,
            reinterpret_cast<unsigned long long*>(e->next), e->dirty, e->w_ckey.as<uint64_t>(), e->w_cS.as<uint64_t>(),
            e->d_nchg, sbits);
        ms_expand_kernel<true, IPT><<<xgrid, BLOCK, 0, st>>>(
            e->S, e->P, e->tile_lo, e->qdeg, e->qoff, e->qmask, e->list, g->arena, e->frontier,
            reinterpret_cast<unsigned long long*>(e->next), e->dirty, e->w_ckey.as<uint64_t>(), e->w_cS.as<uint64_t>(),
            e->d_nchg, sbits);
    };
    if (P.ipt == 2) expand(std::integral_constant<int, 2>());
    else if (P.ipt == 6) expand(std::integral_constant<int, 6>());
    else expand(std::integral_constant<int, 4>());
    e->exec_cS = e->w_cS.ptr;
    e->exec_ckey = e->w_ckey.ptr;
    ms_pull_heavy_kernel<<<wgrid, BLOCK, 0, st>>>(words, e->S, e->hbits, g->rec, g->arena, e->frontier, e->visited,
                                                  e->dirty, e->next);
    ms_side_kernel<<<wgrid, BLOCK, 0, st>>>(e->S, e->d_nside, e->side_uw.as<int2>(), e->side_sm.as<uint64_t>(),
                                            e->frontier, e->visited, reinterpret_cast<unsigned long long*>(e->next));
    ms_collect_kernel<<<(unsigned)e->nparts, BLOCK, 0, st>>>(nv, e->degs, e->visited, e->frontier, e->next, dist, e->S,
                                                             P.T, e->dirty, e->fbits, e->obits, e->hbits,
                                                             e->part, P.digest, e->dg_part);
    return hipGetLastError();
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
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds the 32-bit work prefix");
    HN_TRY(ms_reserve(e));
    bool sym = false;
    HN_TRY(graph_is_symmetric(g, st, &sym));
    MsParams P = e->params;
    P.push_only = sym ? 0 : 1;
    P.digest = dist ? 0 : 1;
    for (int b = 0; b < 64; b++) P.depth_limit[b] = (depth_limit && b < k) ? depth_limit[b] : 0;
    const int64_t nv = g->nv;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(nv, BLOCK), (int64_t)num_cus() * 8);
    HN_HIP(hipMemcpyAsync(e->d_src, sources, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, st));
    if (e->degs_version != g->version) {   // degrees change with every applied update
        if (nv > 0) ms_degrees_kernel<<<(unsigned)ceil_div(nv, BLOCK), BLOCK, 0, st>>>(nv, g->rec, e->degs);
        e->degs_version = g->version;
    }
    ms_init_kernel<<<grid, BLOCK, 0, st>>>(nv, e->visited, e->frontier, e->next, 3 * ms_bit_words(nv), e->fbits);
    if (dist)
        ms_fill_dist_kernel<<<(unsigned)std::min<int64_t>(ceil_div((int64_t)k * nv, 4 * BLOCK),
                                                          (int64_t)num_cus() * 16),
                              BLOCK, 0, st>>>((int64_t)k * nv, dist);
    else
        HN_HIP(hipMemsetAsync(e->dg_part, 0, (size_t)e->nparts * 64 * 3 * 8, st));
    ms_seed_kernel<<<1, WAVE, 0, st>>>(k, e->d_src, nv, e->visited, e->frontier, e->fbits, dist, g->rec, e->S, P);
    HN_CHECK_LAUNCH();
    // capture the level step once per (stream, dist buffer, parameters, buffers)
    const bool same = e->exec && e->exec_stream == st && e->exec_dist == dist && e->exec_arena == g->arena &&
                      std::memcmp(&e->exec_params, &P, sizeof(P)) == 0;
    if (!same) {
        if (e->exec) (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
        hipGraph_t graph = nullptr;
        HN_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t le = ms_level(e, dist, P, st);
        hipError_t ce = hipStreamEndCapture(st, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            set_error(std::string("level-step capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
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
        e->exec_arena = g->arena;
        e->exec_params = P;
    }
    // Replay.  The first poll of the done flag comes after as many steps as
    // the last launch on this graph needed (a graph's depth barely changes
    // between launches; depth-limited launches need at most limit + 2), the
    // next ones after 1, 2, 4, 4, ... more steps: with a fixed poll every 4
    // steps a launch ran up to 3 steps past its end, each one 11 kernels that
    // return at once but still cost their launch gaps.
    int32_t max_limit = 0;
    bool limited = k > 0;
    for (int b = 0; b < k; b++) {
        limited = limited && P.depth_limit[b] > 0;
        max_limit = std::max(max_limit, P.depth_limit[b]);
    }
    int64_t next = e->last_steps > 0 ? e->last_steps : 4;
    if (limited) next = std::min<int64_t>(next, max_limit + 2);
    int64_t ramp = 1;
    int32_t* host = e->pinned;   // [0] done, [1] level steps that did work
    static_assert(offsetof(MsState, levels_run) == offsetof(MsState, done) + 4, "done, levels_run adjacent");
    for (int64_t rep = 0; rep <= nv + 1;) {
        for (int64_t c = 0; c < next; c++) HN_HIP(hipGraphLaunch(e->exec, st));
        rep += next;
        HN_HIP(hipMemcpyAsync(host, &e->S->done, 8, hipMemcpyDeviceToHost, st));
        HN_HIP(hipStreamSynchronize(st));
        if (host[0]) {
            if (!limited) e->last_steps = host[1] + 1;
            break;
        }
        next = ramp;
        ramp = std::min<int64_t>(2 * ramp, 4);
    }
    if (!dist) {
        ms_digest_kernel<<<(unsigned)k, BLOCK, 0, st>>>(e->dg_part, e->nparts, e->digest);
        HN_CHECK_LAUNCH();
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

// ---- multi-version windows: called by the serving loop ----------------------

hn_status ms_window_begin(MsEngine* e, const int32_t* d_src, const int32_t* d_dst, const uint8_t* d_tt, int64_t n,
                          hipStream_t st, bool* symmetric) {
    hn_graph* g = e->g;
    const int sbits = std::max(1, ceil_log2_host((int64_t)g->nv + 1));
    HN_ARG(2 * sbits + 7 <= 64, "window keys need 2 * ceil(log2(nv + 1)) + 7 <= 64 bits");
    HN_ARG(n >= 0 && n < (int64_t)0x7FFFFFFF, "window too large");
    *symmetric = true;
    e->w_n = n;
    e->has_window = true;
    if (n == 0) return HN_OK;
    HN_TRY(e->w_keys.reserve((size_t)n * 8));
    HN_TRY(e->w_alt.reserve((size_t)n * 8 + 512));
    HN_TRY(e->w_tmp.reserve(std::max(sort_temp_bytes(n), scan_temp_bytes(std::max<int64_t>(n, g->nv)))));
    HN_TRY(e->w_seg.reserve((size_t)(n + 1) * 8 + 256));
    HN_TRY(e->w_e0.reserve((size_t)n));
    uint64_t* keys = e->w_keys.as<uint64_t>();
    uint64_t* alt = e->w_alt.as<uint64_t>();
    win_pack_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(n, d_src, d_dst, d_tt, sbits, keys);
    HN_CHECK_LAUNCH();
    HN_TRY(radix_sort_u64(&keys, &alt, n, 2 * sbits + 7, e->w_tmp.ptr, st));
    e->w_sorted = keys;
    {
        uint32_t* d_bad = reinterpret_cast<uint32_t*>(e->w_seg.as<char>() + (size_t)(n + 1) * 8);
        HN_HIP(hipMemsetAsync(d_bad, 0, 4, st));
        win_sym_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(n, keys, sbits, d_bad);
        HN_CHECK_LAUNCH();
        uint32_t bad = 0;
        HN_TRY(ms_read(e, d_bad, 4, &bad, st));
        if (bad) {   // not a symmetric window: the caller applies these updates one by one
            *symmetric = false;
            e->has_window = false;
            e->w_n = 0;
            return HN_OK;
        }
    }
    uint32_t* hpos = e->w_seg.as<uint32_t>();
    uint32_t* seg_start = hpos + (n + 1);
    HN_TRY(exclusive_scan<uint32_t>(PairHeadIn{keys}, hpos, n, e->w_tmp.ptr, st));
    win_seg_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(n, keys, hpos, seg_start);
    HN_CHECK_LAUNCH();
    // membership before the first update of the window
    win_e0_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(hpos + n, seg_start, keys, sbits, g->rec,
                                                                  g->arena, e->w_e0.as<uint8_t>());
    HN_CHECK_LAUNCH();
    HN_HIP(hipStreamSynchronize(st));
    return HN_OK;
}

hn_status ms_window_end(MsEngine* e, int32_t m, uint64_t failed, const uint64_t* sv, hipStream_t st) {
    hn_graph* g = e->g;
    HN_ARG(e->has_window, "no open window");
    HN_ARG(m >= 0 && m <= 63, "a window holds at most 63 updates");
    HN_TRY(ms_reserve(e));   // may re-lay out the buffers after the updates: before any flag is set
    e->has_window = true;
    const int64_t nv = g->nv, n = e->w_n;
    const int sbits = std::max(1, ceil_log2_host((int64_t)nv + 1));
    HN_TRY(e->w_tmp.reserve(scan_temp_bytes(std::max<int64_t>(n, nv))));
    HN_HIP(hipMemsetAsync(e->dirty, 0, (size_t)(nv > 0 ? nv : 1), st));
    uint32_t nchg = 0;
    if (n > 0) {
        HN_TRY(e->w_E.reserve((size_t)n * 8));
        HN_TRY(e->w_chg.reserve((size_t)(2 * n + 1) * 4 + 256));
        HN_TRY(e->w_ckey.reserve((size_t)n * 8));
        HN_TRY(e->w_cE.reserve((size_t)n * 8));
        uint32_t* hpos = e->w_seg.as<uint32_t>();
        uint32_t* seg_start = hpos + (n + 1);
        const uint32_t* nseg = hpos + n;
        uint32_t* chg = e->w_chg.as<uint32_t>();
        uint32_t* cpos = chg + n;
        win_exist_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(nseg, seg_start, e->w_sorted,
                                                                         e->w_e0.as<uint8_t>(), m, failed, sbits,
                                                                         e->w_E.as<uint64_t>(), chg, e->dirty);
        HN_CHECK_LAUNCH();
        HN_TRY(exclusive_scan<uint32_t>(FlagIn{chg, nseg}, cpos, n, e->w_tmp.ptr, st));
        win_chg_write_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(nseg, chg, cpos, seg_start, e->w_sorted,
                                                                             e->w_E.as<uint64_t>(),
                                                                             e->w_ckey.as<uint64_t>(),
                                                                             e->w_cE.as<uint64_t>());
        HN_CHECK_LAUNCH();
        HN_TRY(ms_read(e, cpos + n, 4, &nchg, st));
    }
    // sources per changed pair; the pairs absent from storage become side records
    HN_TRY(e->w_cS.reserve((size_t)(nchg > 0 ? nchg : 1) * 8));
    HN_TRY(e->w_sflag.reserve((size_t)(2 * (int64_t)nchg + 2) * 4 + 256));
    void* old_uw = e->side_uw.ptr;
    void* old_sm = e->side_sm.ptr;
    HN_TRY(e->side_uw.reserve((size_t)(nchg > 0 ? nchg : 1) * 8));
    HN_TRY(e->side_sm.reserve((size_t)(nchg > 0 ? nchg : 1) * 8));
    if (e->side_uw.ptr != old_uw || e->side_sm.ptr != old_sm || e->w_cS.ptr != e->exec_cS ||
        e->w_ckey.ptr != e->exec_ckey) {   // the captured level step holds these pointers
        if (e->exec) (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    set_u32_kernel<<<1, 1, 0, st>>>(e->d_nchg, nchg);
    HN_CHECK_LAUNCH();
    if (nchg > 0) {
        SideSV v;
        for (int t = 0; t < 64; t++) v.v[t] = t <= m ? sv[t] : 0ull;
        uint32_t* sflag = e->w_sflag.as<uint32_t>();
        uint32_t* spos = sflag + (nchg + 1);
        win_cs_kernel<<<(unsigned)ceil_div(nchg, BLOCK), BLOCK, 0, st>>>(e->d_nchg, e->w_cE.as<uint64_t>(), m, v,
                                                                         e->w_cS.as<uint64_t>(), sflag);
        HN_CHECK_LAUNCH();
        HN_TRY(exclusive_scan<uint32_t>(ArrayIn<uint32_t>{sflag}, spos, nchg, e->w_tmp.ptr, st));
        win_side_write_kernel<<<(unsigned)ceil_div(nchg, BLOCK), BLOCK, 0, st>>>(
            e->d_nchg, sflag, spos, e->w_ckey.as<uint64_t>(), e->w_cS.as<uint64_t>(), sbits,
            e->side_uw.as<int2>(), e->side_sm.as<uint64_t>(), e->d_nside);
        HN_CHECK_LAUNCH();
    } else {
        set_u32_kernel<<<1, 1, 0, st>>>(e->d_nside, 0u);
        HN_CHECK_LAUNCH();
    }
    HN_HIP(hipStreamSynchronize(st));
    return HN_OK;
}

const unsigned long long* ms_engine_digest(const MsEngine* e) { return e->digest; }

hn_status ms_window_clear(MsEngine* e, hipStream_t st) {
    if (!e->has_window) return HN_OK;
    HN_HIP(hipMemsetAsync(e->dirty, 0, (size_t)(e->nv > 0 ? e->nv : 1), st));
    HN_HIP(hipMemsetAsync(e->d_nside, 0, 4, st));
    HN_HIP(hipMemsetAsync(e->d_nchg, 0, 4, st));
    e->has_window = false;
    e->w_n = 0;
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

hn_status hn_bfs_multi_lb_tile(hn_graph* g, int32_t units_per_lane) {
    HN_ARG(g, "null graph");
    bool ok = false;
    for (int v : MS_IPTS) ok |= v == units_per_lane;
    HN_ARG(ok, "units per lane must be 2, 4 or 6");
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    e->params.ipt = units_per_lane;
    return HN_OK;
}

hn_status hn_bfs_multi_get_params(hn_graph* g, int32_t* degree_threshold, int32_t* push_alpha,
                                  int32_t* units_per_lane) {
    HN_ARG(g, "null graph");
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    if (degree_threshold) *degree_threshold = e->params.T;
    if (push_alpha) *push_alpha = e->params.alpha;
    if (units_per_lane) *units_per_lane = e->params.ipt;
    return HN_OK;
}

hn_status hn_bfs_multi_tune(hn_graph* g, const int32_t* sources, int32_t k, int32_t* dist, int32_t* best_threshold,
                            int32_t* best_alpha, void* stream) {
    HN_ARG(g && sources && dist, "null argument");
    HN_ARG(k >= 1 && k <= 64, "k must be in [1, 64]");
    hipStream_t st = as_stream(stream);
    MsEngine* e = nullptr;
    HN_TRY(ms_engine(g, &e));
    const int32_t Ts[] = {64, 256, 1024, MS_NO_HEAVY};
    const int32_t As[] = {4, 16, 64};
    hipEvent_t a = nullptr, b = nullptr;
    HN_HIP(hipEventCreate(&a));
    HN_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int32_t bt = e->params.T, ba = e->params.alpha, bi = MS_IPT_DEFAULT;
    hn_status s = HN_OK;
    // one warm run (captures the level graph), one timed run
    auto timed = [&](float* ms) -> hn_status {
        hn_status r = ms_run(g, sources, k, nullptr, dist, nullptr, st);
        if (r != HN_OK) return r;
        (void)hipEventRecord(a, st);
        if ((r = ms_run(g, sources, k, nullptr, dist, nullptr, st)) != HN_OK) return r;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        *ms = 0.f;
        (void)hipEventElapsedTime(ms, a, b);
        return HN_OK;
    };
    // the degree threshold and switch point over a grid at the default tile,
    // then the tile size at the best pair
    e->params.ipt = MS_IPT_DEFAULT;
    for (int32_t T : Ts) {
        for (int32_t A : As) {
            e->params.T = T;
            e->params.alpha = A;
            float ms = 0.f;
            if ((s = timed(&ms)) != HN_OK) break;
            if (ms < best) {
                best = ms;
                bt = T;
                ba = A;
            }
        }
        if (s != HN_OK) break;
    }
    e->params.T = bt;
    e->params.alpha = ba;
    for (int ipt : MS_IPTS) {
        if (s != HN_OK || ipt == MS_IPT_DEFAULT) continue;
        e->params.ipt = ipt;
        float ms = 0.f;
        if ((s = timed(&ms)) != HN_OK) break;
        if (ms < best) {
            best = ms;
            bi = ipt;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->params.ipt = bi;
    if (best_threshold) *best_threshold = bt;
    if (best_alpha) *best_alpha = ba;
    return s;
}

}  // extern "C"
