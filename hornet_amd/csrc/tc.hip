// tc.hip -- triangle counting (HornetsNest TriangleCounting2).
//
// Reference: TriangleCounting2::run -> forAllAdjUnions
// (primitives/Operator++.i.cuh:369-475) bins every stored edge by estimated
// work into 2,048 bins with 64-bit atomics, copies the bin sizes to the host,
// and launches up to 39 merge-path / binary-search intersection kernels
// (:97-246), each followed by a device sync; the operator
// (hornetsnest/src/Static/TriangleCounting/triangle2.cu:75-124) skips
// src == dst and deg(src) < 2, intersects N(u) with N(v) where u is the
// endpoint with the smaller degree (ties -> the destination), and credits u.
// The driver stores the DIRECTED_BY_DEGREE orientation
// (hornetsnest/test/TriangleTest2.cu:86-87, xlib/src/Graph/GraphStd.cpp:247-304),
// so the sum is the triangle count.
//
// Here:
//   * orientation of the symmetric storage is a flag + scan compaction (no
//     atomics): keep (u,v) iff (deg u, u) < (deg v, v); the oriented offsets
//     are the scan positions at the symmetric offsets;
//   * intersection work min(d(u), d(v)) splits every edge into one of three
//     wave64 classes: one lane per edge (<= 16 entries in the shorter list),
//     sixteen lanes per edge (<= 128), where each lane binary-searches
//     entries of the shorter list in the longer one, or a whole wave per edge
//     intersecting 64-entry windows of both lists (tc_window_kernel; both
//     equal the merge count on simple graphs);
//   * the two deferred classes are queued by the one-lane pass; their waves
//     claim 16 queued edges at a time from a ticket, which keeps the waves
//     busy to the end of the skewed queues;
//   * counts are reduced per wave; the total takes one 64-bit atomicAdd per
//     wave, per-vertex credits (optional) one per edge with a non-zero count.
#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr int TC_IPT = 2;
constexpr int TC_TILE = BLOCK * TC_IPT;
constexpr int LIGHT_MAX = 16;
// class limits on the shorter list: one lane per edge up to LIGHT_MAX
// entries, MEDIUM_LANES lanes up to MEDIUM_MAX, the window intersection
// above; measured on R-MAT s24 with binary searches in every class (16 / 256:
// 783 ms per recount; 8 / 512: 819 ms; 8 / 256: 805 ms; 32 / 512: 784 ms;
// 4 / 512: 1057 ms), then with the window class (window from 257 entries:
// 451 ms, from 129: 431 ms, from 193: 448 ms; from 97 or 65 when the longer
// list is at most 8 or 16 times longer: 430-435 ms)
constexpr int MEDIUM_MAX = 128;
constexpr int MEDIUM_LANES = 16;

__device__ __forceinline__ bool bsearch_has(const int32_t* __restrict__ a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        int32_t y = a[mid];
        if (y < x) lo = mid + 1;
        else if (y > x) hi = mid;
        else return true;
    }
    return false;
}

struct TcCsr {
    const int64_t* off;
    const int32_t* adj;
};

struct TcOut {
    unsigned long long* total;
    unsigned long long* per_vertex;   // may be null
    uint32_t* med_q;                  // edge ids for the medium class
    uint32_t* heavy_q;                // edge ids for the heavy class
    uint32_t* cnt;                    // [0] medium, [1] heavy, [2] heavy ticket, [3] medium ticket
    int32_t* src_of;                  // edge -> source (written by the light pass)
};

// The light edges of one tile (at most LIGHT_MAX entries in the shorter
// list): their lists, and one work item per entry of the shorter list.
struct LightShared {
    uint32_t s[TC_TILE];        // first entry of the shorter list in adj
    uint32_t l[TC_TILE];        // first entry of the longer list
    int32_t ln[TC_TILE];        // length of the longer list
    int32_t credit[TC_TILE];    // the vertex credited
    uint32_t pre[TC_TILE + 1];  // shorter-list lengths, then their exclusive prefix
    uint32_t cnt[TC_TILE];      // matches per light edge
    uint32_t sc[NWAVES + 1];
    uint32_t n;
};

// Light pass over all stored edges: defers medium / heavy edges to queues and
// counts the light ones.  A light edge's up to LIGHT_MAX binary searches are
// spread over the workgroup (one search per lane, the tile's light edges
// flattened into one item list), not run one after another by the lane that
// found the edge: the slowest lane of a wave had chained up to
// TC_IPT x LIGHT_MAX searches.
__global__ __launch_bounds__(BLOCK) void tc_light_kernel(int32_t nv, int64_t ne, TcCsr g, TcOut o) {
    __shared__ LbShared<TC_IPT, int64_t> sh;
    __shared__ LightShared ls;
    unsigned long long acc = 0;
    const uint64_t lt = lanemask_lt();
    for (int64_t w0 = (int64_t)blockIdx.x * TC_TILE; w0 < ne; w0 += (int64_t)gridDim.x * TC_TILE) {
        if (threadIdx.x == 0) ls.n = 0;   // lb_tile synchronises before it calls the items
        lb_tile<TC_IPT>(
            g.off, (int64_t)nv, w0, ne, sh, [&](int64_t s) { return (int64_t)0; },
            [&](int64_t s, int64_t e, int64_t) {
                const int32_t u0 = (int32_t)s;
                const int32_t v0 = g.adj[e];
                const int64_t su = g.off[u0], du = g.off[u0 + 1] - su;
                const int64_t sv = g.off[v0], dv = g.off[v0 + 1] - sv;
                int cls = 0;   // 0 done here, 1 medium, 2 heavy
                o.src_of[e] = u0;
                if (u0 != v0 && du >= 2) {
                    const bool src_smaller = du < dv;
                    const int64_t sn = src_smaller ? du : dv;
                    if (sn <= LIGHT_MAX) {
                        if (sn > 0) {
                            const uint32_t k = atomicAdd(&ls.n, 1u);
                            ls.s[k] = (uint32_t)(src_smaller ? su : sv);
                            ls.l[k] = (uint32_t)(src_smaller ? sv : su);
                            ls.ln[k] = (int32_t)(src_smaller ? dv : du);
                            ls.credit[k] = src_smaller ? u0 : v0;
                            ls.pre[k] = (uint32_t)sn;
                            ls.cnt[k] = 0;
                        }
                    } else {
                        cls = sn <= MEDIUM_MAX ? 1 : 2;
                    }
                }
                // wave-aggregated queue appends for deferred edges
                for (int q = 1; q <= 2; q++) {
                    const uint64_t m = __ballot(cls == q);
                    if (m) {
                        const int leader = __ffsll((unsigned long long)m) - 1;
                        uint32_t b = 0;
                        if (__lane_id() == leader) b = atomicAdd(o.cnt + (q - 1), (uint32_t)__popcll(m));
                        b = __shfl(b, leader, WAVE);
                        if (cls == q) (q == 1 ? o.med_q : o.heavy_q)[b + (uint32_t)__popcll(m & lt)] = (uint32_t)e;
                    }
                }
            });
        __syncthreads();
        // exclusive prefix of the shorter-list lengths (thread t owns entries
        // t * TC_IPT .. t * TC_IPT + TC_IPT - 1)
        const uint32_t n = ls.n;
        uint32_t len[TC_IPT], sum = 0;
#pragma unroll
        for (int j = 0; j < TC_IPT; j++) {
            const uint32_t k = (uint32_t)threadIdx.x * TC_IPT + j;
            len[j] = k < n ? ls.pre[k] : 0u;
            sum += len[j];
        }
        uint32_t items;
        uint32_t run = block_exclusive_scan(sum, ls.sc, &items);
#pragma unroll
        for (int j = 0; j < TC_IPT; j++) {
            const uint32_t k = (uint32_t)threadIdx.x * TC_IPT + j;
            if (k < n) ls.pre[k] = run;
            run += len[j];
        }
        __syncthreads();
        // one binary search per item: item t is entry t - pre[k] of light edge k
        for (uint32_t t = threadIdx.x; t < items; t += BLOCK) {
            uint32_t lo = 0, hi = n;   // largest k with pre[k] <= t (pre strictly increasing)
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ls.pre[mid] <= t) lo = mid; else hi = mid;
            }
            const int32_t x = g.adj[(int64_t)ls.s[lo] + (t - ls.pre[lo])];
            if (bsearch_has(g.adj + ls.l[lo], ls.ln[lo], x)) atomicAdd(&ls.cnt[lo], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += BLOCK) {
            const uint32_t c = ls.cnt[k];
            if (c) {
                if (o.per_vertex) atomicAdd(o.per_vertex + ls.credit[k], (unsigned long long)c);
                acc += c;
            }
        }
        __syncthreads();   // the next tile reuses ls
    }
    acc = wave_reduce_sum(acc);
    if (__lane_id() == 0 && acc) atomicAdd(o.total, acc);
}

// `lanes` lanes cooperate on one deferred edge
constexpr int64_t TC_CHUNK = 16;
template <int LANES>
__global__ __launch_bounds__(BLOCK) void tc_group_kernel(TcCsr g, const uint32_t* __restrict__ q,
                                                         const uint32_t* __restrict__ qcount,
                                                         const int32_t* __restrict__ src_of, TcOut o) {
    const uint32_t n = *qcount;
    const int sub = threadIdx.x % LANES;
    unsigned long long acc = 0;
    // one deferred edge; the LANES lanes of a group share k, so a group is
    // either wholly active or wholly done and the group shuffles never read an
    // exited lane
    auto one = [&](int64_t k) {
        const int64_t e = q[k];
        const int32_t u0 = src_of[e];
        const int32_t v0 = g.adj[e];
        const int64_t su = g.off[u0], du = g.off[u0 + 1] - su;
        const int64_t sv = g.off[v0], dv = g.off[v0 + 1] - sv;
        const bool src_smaller = du < dv;
        const int32_t* sp = g.adj + (src_smaller ? su : sv);
        const int64_t sn = src_smaller ? du : dv;
        const int32_t* lp = g.adj + (src_smaller ? sv : su);
        const int64_t ln = src_smaller ? dv : du;
        uint32_t c = 0;
        for (int64_t i = sub; i < sn; i += LANES) c += bsearch_has(lp, ln, sp[i]) ? 1u : 0u;
        const int32_t credit = src_smaller ? u0 : v0;
        // reduce over the group
#pragma unroll
        for (int d = LANES / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, WAVE);
        if (sub == 0 && c && o.per_vertex) atomicAdd(o.per_vertex + credit, (unsigned long long)c);
        if (sub == 0) acc += c;
    };
    // edges differ widely in work: each wave claims TC_CHUNK edges per group
    // at a time from its class's ticket until the queue is drained; the
    // groups of a wave take the claimed edges in turn
    constexpr int G = WAVE / LANES;
    const int grp = (int)(__lane_id() / LANES);
    uint32_t* ticket = o.cnt + (LANES == WAVE ? 2 : 3);
    for (;;) {
        uint32_t t = 0;
        if (__lane_id() == 0) t = atomicAdd(ticket, 1u);
        t = __shfl(t, 0, WAVE);
        const int64_t k0 = (int64_t)t * TC_CHUNK * G;
        if (k0 >= (int64_t)n) break;
        const int64_t k1 = k0 + TC_CHUNK * G < (int64_t)n ? k0 + TC_CHUNK * G : (int64_t)n;
        for (int64_t k = k0 + grp; k < k1; k += G) one(k);
    }
    acc = wave_reduce_sum(acc);
    if (__lane_id() == 0 && acc) atomicAdd(o.total, acc);
}

// A whole wave per deferred edge, intersecting by sorted windows: lane l
// holds element l of a 64-entry window of the shorter list S and of the
// longer list L; each lane finds its S element in the L window with six
// cross-lane steps, then the S window moves on if its last element is not
// above the L window's, else the L window does.  Every load is one coalesced
// 256-byte window, where the binary searches scatter their lanes over a
// list's cache lines.  Once the L window has moved past an S window, the
// S entries up to the old L window's last element were already checked, so
// they count only above it: each S entry counts once if L holds it, as with
// the binary searches, repeated entries of a multigraph included (a host
// model of this loop was checked against per-entry binary searches on random
// lists with and without repeats).
//
// A wave runs TC_SLOTS edges at once, each in its own slot, so that the
// slots' dependent cross-lane steps and window loads overlap; a slot whose
// edge is done takes the next one.  Windows past a list's end repeat its last
// entry (they stay sorted and their last lane holds their largest entry),
// and every iteration loads the next window of both lists of every slot (the
// same one again when a list does not move on), so the loads need no
// selects and are in flight while the slots are matched.
// (R-MAT s24 recount: 1 slot 472 ms, 2 slots 450 ms, 4 slots 490 ms, whose
// registers leave half the waves per SIMD)
constexpr int SLOTS = 2;

struct WinSlot {
    const int32_t* sp;   // shorter list
    const int32_t* lp;   // longer list
    int64_t sn, ln;      // their lengths (>= 1 while live)
    int64_t i, j;        // window starts
    int32_t a, b;        // this lane's entry of the current windows
    int32_t an, bn;      // ... of the next windows
    int32_t lim;         // S entries up to lim were checked against an earlier L window
    int32_t credit;      // vertex credited with the count
    uint32_t c;          // this lane's matches
    bool live;
};

__device__ __forceinline__ int32_t window_entry(const int32_t* __restrict__ x, int64_t n, int64_t k) {
    return x[k < n ? k : n - 1];
}

__global__ __launch_bounds__(BLOCK) void tc_window_kernel(TcCsr g, const uint32_t* __restrict__ q,
                                                          const uint32_t* __restrict__ qcount,
                                                          const int32_t* __restrict__ src_of, TcOut o) {
    const uint32_t n = *qcount;
    const int lane = __lane_id();
    unsigned long long acc = 0;
    uint32_t* ticket = o.cnt + 2;
    int64_t k = 0, k1 = 0;   // the wave's claimed edges [k, k1)
    bool drained = false;
    // the next queued edge into slot w (or the slot goes idle)
    auto fetch = [&](WinSlot& w) {
        if (k >= k1 && !drained) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(ticket, 1u);
            t = __shfl(t, 0, WAVE);
            k = (int64_t)t * TC_CHUNK;
            k1 = k + TC_CHUNK < (int64_t)n ? k + TC_CHUNK : (int64_t)n;
            drained = k >= (int64_t)n;
        }
        if (drained || k >= k1) {
            w.live = false;
            w.sp = w.lp = g.adj;   // an idle slot keeps loading one valid entry
            w.sn = w.ln = 1;
            w.i = w.j = 0;
            w.a = w.b = w.an = w.bn = 0;
            w.lim = -1;
            w.c = 0;
            w.credit = 0;
            return;
        }
        const int64_t e = q[k++];
        const int32_t u0 = src_of[e];
        const int32_t v0 = g.adj[e];
        const int64_t su = g.off[u0], du = g.off[u0 + 1] - su;
        const int64_t sv = g.off[v0], dv = g.off[v0 + 1] - sv;
        const bool src_smaller = du < dv;
        w.sp = g.adj + (src_smaller ? su : sv);
        w.sn = src_smaller ? du : dv;
        w.lp = g.adj + (src_smaller ? sv : su);
        w.ln = src_smaller ? dv : du;
        w.credit = src_smaller ? u0 : v0;
        w.live = true;   // a deferred edge has both lists above LIGHT_MAX entries
        w.i = w.j = 0;
        w.a = window_entry(w.sp, w.sn, lane);
        w.b = window_entry(w.lp, w.ln, lane);
        w.an = window_entry(w.sp, w.sn, WAVE + lane);
        w.bn = window_entry(w.lp, w.ln, WAVE + lane);
        w.lim = -1;
        w.c = 0;
    };
    WinSlot ws[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) fetch(ws[s]);
    for (;;) {
        bool any = false;
#pragma unroll
        for (int s = 0; s < SLOTS; s++) any = any || ws[s].live;
        if (!any) break;
        // first index p of each L window with b[p] >= a, the slots interleaved
        int p[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; s++) p[s] = 0;
#pragma unroll
        for (int st = 32; st > 0; st >>= 1)
#pragma unroll
            for (int s = 0; s < SLOTS; s++) {
                const int32_t x = __shfl(ws[s].b, p[s] + st - 1, WAVE);
                p[s] += x < ws[s].a ? st : 0;
            }
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
            WinSlot& w = ws[s];
            const int32_t y = __shfl(w.b, p[s] & (WAVE - 1), WAVE);
            w.c += (w.live && p[s] < WAVE && y == w.a && w.i + lane < w.sn && w.a > w.lim) ? 1u : 0u;
            const int32_t amax = __builtin_amdgcn_readlane(w.a, WAVE - 1);
            const int32_t bmax = __builtin_amdgcn_readlane(w.b, WAVE - 1);
            const bool adv_a = amax <= bmax;
            w.i += adv_a ? WAVE : 0;
            w.j += adv_a ? 0 : WAVE;
            w.a = adv_a ? w.an : w.a;
            w.b = adv_a ? w.b : w.bn;
            w.lim = adv_a ? -1 : bmax;
            w.an = window_entry(w.sp, w.sn, w.i + WAVE + lane);
            w.bn = window_entry(w.lp, w.ln, w.j + WAVE + lane);
        }
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
            WinSlot& w = ws[s];
            if (w.live && (w.i >= w.sn || w.j >= w.ln)) {
                const uint32_t tot = wave_reduce_sum(w.c);
                if (lane == 0) {
                    if (tot && o.per_vertex) atomicAdd(o.per_vertex + w.credit, (unsigned long long)tot);
                    acc += tot;
                }
                fetch(w);
            }
        }
    }
    if (lane == 0 && acc) atomicAdd(o.total, acc);
}

// ---- orientation of the symmetric storage
constexpr int OR_IPT = 8;
constexpr int OR_TILE = BLOCK * OR_IPT;

__global__ __launch_bounds__(BLOCK) void orient_flags_kernel(int32_t nv, int64_t ne, const int64_t* __restrict__ eoff,
                                                             const hn_vrec* __restrict__ rec,
                                                             const int32_t* __restrict__ arena,
                                                             int32_t* __restrict__ adj_tmp,
                                                             uint8_t* __restrict__ keep) {
    __shared__ LbShared<OR_IPT, int64_t> sh;
    for (int64_t w0 = (int64_t)blockIdx.x * OR_TILE; w0 < ne; w0 += (int64_t)gridDim.x * OR_TILE) {
        lb_tile<OR_IPT>(
            eoff, (int64_t)nv, w0, ne, sh, [&](int64_t s) { return rec[s].off - eoff[s]; },
            [&](int64_t s, int64_t w, int64_t base) {
                const int32_t v = arena[base + w];
                const int32_t u = (int32_t)s;
                const int32_t du = (int32_t)(eoff[u + 1] - eoff[u]);
                const int32_t dv = rec[v].deg;
                adj_tmp[w] = v;
                keep[w] = (du < dv || (du == dv && u < v)) ? 1 : 0;
            });
    }
}

__global__ void orient_compact_kernel(int64_t ne, const uint8_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                                      const int32_t* __restrict__ adj_tmp, int32_t* __restrict__ adjp) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne && keep[e]) adjp[pos[e]] = adj_tmp[e];
}

__global__ void orient_offsets_kernel(int32_t nv, const int64_t* __restrict__ eoff, const uint32_t* __restrict__ pos,
                                      int64_t* __restrict__ offp) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v <= nv) offp[v] = pos[eoff[v]];
}

struct U8In {
    const uint8_t* a;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const { return a[i]; }
};

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_tc_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&tc_light_kernel)) != hipSuccess) (void)hipGetLastError();
}

static int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

// bytes of tc_core's buffers for ne stored edges
static int64_t tc_core_bytes(int64_t ne) {
    const int64_t ne1 = ne > 0 ? ne : 1;
    return al256(512 + ne1 * 4) + al256(ne1 * 4) + al256(ne1 * 4);
}

// Triangle counting core over a device CSR (sorted adjacency); its buffers
// start at `base` (tc_core_bytes(ne) bytes of the handle's TC scratch).
static hn_status tc_core(hn_graph* g, int32_t nv, int64_t ne, const int64_t* off, const int32_t* adj,
                         uint64_t* total_out, uint64_t* per_vertex, hipStream_t st, char* base) {
    const int64_t ne1 = ne > 0 ? ne : 1;
    int64_t o_tot = 0;
    int64_t o_cnt = 256;
    int64_t o_src = 512;
    int64_t o_med = al256(o_src + ne1 * 4);
    int64_t o_heavy = al256(o_med + ne1 * 4);
    TcOut o;
    o.total = reinterpret_cast<unsigned long long*>(base + o_tot);
    o.cnt = reinterpret_cast<uint32_t*>(base + o_cnt);
    o.src_of = reinterpret_cast<int32_t*>(base + o_src);
    o.med_q = reinterpret_cast<uint32_t*>(base + o_med);
    o.heavy_q = reinterpret_cast<uint32_t*>(base + o_heavy);
    o.per_vertex = reinterpret_cast<unsigned long long*>(per_vertex);
    HN_HIP(hipMemsetAsync(base, 0, 512, st));
    if (per_vertex && nv > 0) HN_HIP(hipMemsetAsync(per_vertex, 0, sizeof(uint64_t) * (size_t)nv, st));
    TcCsr c{off, adj};
    const int64_t cap = (int64_t)num_cus() * 8;
    if (ne > 0) {
        int64_t grid = std::min<int64_t>(ceil_div(ne, TC_TILE), cap);
        tc_light_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(nv, ne, c, o);
        HN_CHECK_LAUNCH();
        tc_group_kernel<MEDIUM_LANES><<<(unsigned)cap, BLOCK, 0, st>>>(c, o.med_q, o.cnt, o.src_of, o);
        HN_CHECK_LAUNCH();
        tc_window_kernel<<<(unsigned)cap, BLOCK, 0, st>>>(c, o.heavy_q, o.cnt + 1, o.src_of, o);
        HN_CHECK_LAUNCH();
    }
    unsigned long long t = 0;
    HN_TRY(read_back(g, o.total, 8, &t, st));
    *total_out = t;
    return HN_OK;
}

}  // namespace hn

using namespace hn;

extern "C" {

// The TC entry points keep every buffer in the handle's grow-only scratch
// (s_tc): the CSR they count on, then tc_core's buffers.

hn_status hn_tc2_stored(hn_graph* g, uint64_t* total, uint64_t* per_vertex, void* stream) {
    HN_ARG(g && total, "null argument");
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds 32-bit queue ids");
    hipStream_t st = as_stream(stream);
    const int64_t nv = g->nv, ne = g->ne, ne1 = ne > 0 ? ne : 1;
    const int64_t o_off = 0;
    const int64_t o_adj = al256(o_off + (nv + 1) * 8);
    const int64_t o_core = al256(o_adj + ne1 * 4);
    HN_TRY(g->s_tc.reserve((size_t)(o_core + tc_core_bytes(ne))));
    char* base = g->s_tc.as<char>();
    int64_t* off = reinterpret_cast<int64_t*>(base + o_off);
    int32_t* adj = reinterpret_cast<int32_t*>(base + o_adj);
    HN_TRY(hn_graph_export_csr(g, off, adj, ne, stream));
    return tc_core(g, g->nv, ne, off, adj, total, per_vertex, st, base + o_core);
}

hn_status hn_tc_count(hn_graph* g, uint64_t* total, uint64_t* per_vertex, void* stream) {
    HN_ARG(g && total, "null argument");
    HN_ARG(g->ne < (int64_t)0xFFFFFFFF, "edge count exceeds the 32-bit scan");
    hipStream_t st = as_stream(stream);
    const int32_t nv = g->nv;
    const int64_t ne = g->ne, ne1 = ne > 0 ? ne : 1;
    // the oriented CSR holds at most half the entries of the symmetric one
    // (a multigraph's repeated entries may all point up: size for ne)
    const int64_t o_eoff = 0;
    const int64_t o_offp = al256(o_eoff + ((int64_t)nv + 1) * 8);
    const int64_t o_tmp = al256(o_offp + ((int64_t)nv + 1) * 8);
    const int64_t o_adjp = al256(o_tmp + ne1 * 4);
    const int64_t o_keep = al256(o_adjp + ne1 * 4);
    const int64_t o_pos = al256(o_keep + ne1);
    const int64_t o_scan = al256(o_pos + (ne1 + 1) * 4);
    const int64_t o_core = al256(o_scan + (int64_t)scan_temp_bytes(ne1));
    HN_TRY(g->s_tc.reserve((size_t)(o_core + tc_core_bytes(ne))));
    char* base = g->s_tc.as<char>();
    int64_t* eoff = reinterpret_cast<int64_t*>(base + o_eoff);
    int64_t* offp = reinterpret_cast<int64_t*>(base + o_offp);
    int32_t* adj_tmp = reinterpret_cast<int32_t*>(base + o_tmp);
    int32_t* adjp = reinterpret_cast<int32_t*>(base + o_adjp);
    uint8_t* keep = reinterpret_cast<uint8_t*>(base + o_keep);
    uint32_t* pos = reinterpret_cast<uint32_t*>(base + o_pos);
    void* tmp = base + o_scan;
    HN_TRY(degree_prefix(g, eoff, st));
    if (ne > 0) {
        int64_t grid = std::min<int64_t>(ceil_div(ne, OR_TILE), (int64_t)num_cus() * 16);
        orient_flags_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(nv, ne, eoff, g->rec, g->arena, adj_tmp, keep);
        HN_CHECK_LAUNCH();
    }
    HN_TRY(exclusive_scan<uint32_t>(U8In{keep}, pos, ne, tmp, st));
    if (ne > 0) {
        orient_compact_kernel<<<(unsigned)ceil_div(ne, BLOCK), BLOCK, 0, st>>>(ne, keep, pos, adj_tmp, adjp);
        HN_CHECK_LAUNCH();
    }
    orient_offsets_kernel<<<(unsigned)ceil_div((int64_t)nv + 1, BLOCK), BLOCK, 0, st>>>(nv, eoff, pos, offp);
    HN_CHECK_LAUNCH();
    uint32_t nep = 0;
    HN_TRY(read_back(g, pos + ne, 4, &nep, st));
    return tc_core(g, nv, (int64_t)nep, offp, adjp, total, per_vertex, st, base + o_core);
}

}  // extern "C"
