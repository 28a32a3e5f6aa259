// update.hip -- BatchUpdate insert / erase on the device.
//
// Reference pipeline (hornet/include/Core/BatchUpdate/BatchUpdate.i.cuh and
// BatchUpdateKernels.cuh):
//   insert  sort (2x thrust sort_by_key) -> unique_copy -> graph dedup
//           (RLE, scans, mark_duplicate_edges_kernel) -> realloc queue ->
//           HOST allocator loop -> move_adjacency_lists -> append at old_deg+k
//           (adjacency order not kept), ~8 host syncs per batch;
//   erase   sort -> unique -> multiplicities (markUniqueOffsets) ->
//           locate_erased_edges (atomicSub) -> copy_if -> tail fill of the
//           holes (order destroyed) -> shrink realloc.
// Semantics kept: insert without dedup is a multiset union; removeBatchDuplicates
// drops in-batch repeats; removeGraphDuplicates drops the FIRST batch copy of
// every (u,v) already stored (lower_bound_left, BatchUpdateKernels.cuh:27-58);
// erase removes min(multiplicity in batch, stored copies) and ignores absent
// edges (BatchUpdateKernels.cuh:263-300).  Self loops are not filtered.
//
// Redesign for the device: one 64-bit LSD radix sort of packed (src,dst) keys,
// flag + scan compactions, and every touched vertex rewritten in place (when
// its block still fits) or into a block from the device allocator:
//   insert  every batch entry is located once in its stored list A; the old
//           entries from the first one that moves go to i + #(B < A[i]), in
//           place in one pass, then batch entry q to q + #(A <= B[q]) (stored
//           copies first on ties), so adjacency stays sorted;
//   erase   survivors are compacted in order, moved as runs between deleted
//           ranges (one run per distinct deleted key plus a tail run per
//           vertex), in place in one pass.
// In both passes the few entries whose slot a neighbouring copy tile may not
// have read yet go through a small side buffer.  A device error word (bad
// vertex id, arena exhausted) turns every later kernel into a no-op and rolls
// the allocator back, so a failed batch leaves the graph unchanged.  Two
// host synchronisations per batch (error word with the side-buffer size,
// final status).
#include <cstdlib>

#include "graph.hpp"
#include "prims.hpp"

namespace hn {
namespace {

constexpr uint32_t ERR_BAD_ID = 1u;
constexpr uint32_t ERR_OOM = 2u;     // AllocState::error bit 0 is mirrored here
// Segmented copies: 4 items per lane, at most BLOCK + 1 runs staged in LDS
// per tile (a tile with more runs searches them in global memory), 8
// workgroups per CU striding over the tiles.  A one-entry-per-item LDS stage
// (49 KB at 8 items per lane) left three workgroups per CU and halved the
// copy rate; measured on a 2 M-entry erase at R-MAT s24, 5.0 -> 3.6 ms.
constexpr int COPY_IPT = 4;
constexpr int COPY_TILE = BLOCK * COPY_IPT;
constexpr int COPY_CAP_IPT = 1;
// The erase compaction's tiles are twice as long (8 items per lane): measured
// on a 2 M-entry erase at R-MAT s24, 2.16 -> 2.04 ms; the insert move is
// fastest at 4 (0.97 ms, 1.04 at 8).
constexpr int ERASE_IPT = 8;
constexpr int ERASE_TILE = BLOCK * ERASE_IPT;
constexpr int COPY_GRID_PER_CU = 8;

// bits per endpoint of a packed (src, dst) key: ids are below nv, and no key
// needs a sentinel, so 2^24 vertices pack into 48 bits (6 radix passes)
inline int key_bits_for(int32_t nv) { return std::max(1, ceil_log2_host((int64_t)nv)); }

struct KeyFmt {
    int sbits;
    __device__ __forceinline__ int32_t src(uint64_t k) const { return (int32_t)(k >> sbits); }
    __device__ __forceinline__ int32_t dst(uint64_t k) const {
        return (int32_t)(k & ((1ull << sbits) - 1));
    }
};

struct UpdScal {
    uint32_t err;
    uint32_t asym;                // a batch pair left (u,w) and (w,u) with different membership
    unsigned long long changed;   // entries inserted / erased
    long long side;               // entries deferred to the side buffer
};

// idx (non-null with edge meta-data) records each entry's batch position
__global__ void pack_batch_kernel(int64_t n, const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                                  int32_t nv, int sbits, uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                  UpdScal* sc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (idx) idx[i] = (uint32_t)i;
    int32_t s = src[i], d = dst[i];
    if (s < 0 || s >= nv || d < 0 || d >= nv) {
        atomicOr(&sc->err, ERR_BAD_ID);
        s = 0;
        d = 0;
    }
    keys[i] = ((uint64_t)s << sbits) | (uint64_t)d;
}

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Per sorted batch entry (u, d) of an insert: ub = #(A_u <= d) on u's stored
// list A_u, found once and used three times -- graph dedup (stored iff
// A_u[ub - 1] == d), the first slot of A_u that moves (ub of u's smallest
// kept entry) and the entry's slot in the merged list (q + ub for the q-th
// kept entry of u, stored copies first on ties).  Keep flag: in-batch dedup
// keeps the first copy of a key; graph dedup drops the first batch copy of a
// stored pair (lower_bound_left, BatchUpdateKernels.cuh:27-58).
__global__ void insert_locate_kernel(int64_t n, const uint64_t* __restrict__ keys, KeyFmt f, int dedup_batch,
                                     int dedup_graph, const hn_vrec* __restrict__ rec,
                                     const int32_t* __restrict__ arena, uint32_t* __restrict__ keep,
                                     uint32_t* __restrict__ ub, const UpdScal* sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sc->err) {   // ids were out of range: nothing is inserted
        keep[i] = 0;
        ub[i] = 0;
        return;
    }
    const uint64_t x = keys[i];
    const bool first = (i == 0) || keys[i - 1] != x;
    const hn_vrec r = rec[f.src(x)];
    const int32_t d = f.dst(x);
    const int32_t* row = arena + r.off;
    const int64_t lb = lower_bound_i32(row, r.deg, d);
    const bool stored = lb < r.deg && row[lb] == d;
    // d < nv <= 2^31 - 1, so d + 1 does not overflow
    const int64_t u = stored ? lb + lower_bound_i32(row + lb, r.deg - lb, d + 1) : lb;
    keep[i] = ((dedup_batch && !first) || (dedup_graph && first && stored)) ? 0u : 1u;
    ub[i] = (uint32_t)u;
}

// keep flag for the sorted batch (erase): optional in-batch dedup
struct FirstKeepIn {
    const uint64_t* k;
    int dedup;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const {
        return (!dedup || i == 0 || k[i - 1] != k[i]) ? 1u : 0u;
    }
};

// kept keys (and, when given, the values of one or two parallel arrays)
__global__ void compact_keys_kernel(int64_t n, const uint64_t* __restrict__ in, const uint32_t* __restrict__ pos,
                                    uint64_t* __restrict__ out, const uint32_t* __restrict__ iin,
                                    uint32_t* __restrict__ iout, const uint32_t* __restrict__ jin,
                                    uint32_t* __restrict__ jout) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pos[i + 1] != pos[i]) {
        out[pos[i]] = in[i];
        if (iin) iout[pos[i]] = iin[i];
        if (jin) jout[pos[i]] = jin[i];
    }
}

// segment starts over the first `*cnt` keys: key i starts a segment when its
// group (given by `shift`: the source, or the whole key) differs from key i-1
struct SegStartIn {
    const uint64_t* k;
    const uint32_t* cnt;
    int shift;   // 64 -> whole key compared
    __device__ __forceinline__ uint32_t operator()(int64_t i) const {
        if (i >= (int64_t)*cnt) return 0u;
        if (i == 0) return 1u;
        return shift >= 64 ? (k[i] != k[i - 1] ? 1u : 0u) : ((k[i] >> shift) != (k[i - 1] >> shift) ? 1u : 0u);
    }
};

__global__ void seg_write_kernel(int64_t n, const uint64_t* __restrict__ k, const uint32_t* __restrict__ cnt,
                                 const uint32_t* __restrict__ spos, uint32_t* __restrict__ seg_start,
                                 uint64_t* __restrict__ seg_key) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t m = *cnt;
    if (i > m || i >= n + 1) return;
    if (i == m) {
        seg_start[spos[n]] = (uint32_t)m;
        return;
    }
    if (spos[i + 1] != spos[i]) {
        seg_start[spos[i]] = (uint32_t)i;
        seg_key[spos[i]] = k[i];
    }
}

// Per touched vertex.  A vertex is either updated in place (its block still
// fits: insert new_deg <= capacity, the reference's `new_degree > limit`
// test inverted, BatchUpdateKernels.cuh:180-186; erase keeps the size class)
// or relocated into a block from the device allocator.  In place, only the
// part of the old list from `p` on moves.
struct TouchInfo {
    int64_t old_off;
    int64_t new_off;   // == old_off in place
    int32_t old_deg;
    int32_t new_deg;
    int32_t old_cap;
    int32_t new_cap;
    int32_t p;         // first old slot that moves (0 when relocating)
    int32_t inplace;
};

// insert: per touched source j -- new degree, in place or a new block
// per touched source j: new degree, in place or a new block (no allocation)
__device__ __forceinline__ void insert_touch(int64_t j, const uint32_t* __restrict__ seg_start,
                                             const uint64_t* __restrict__ seg_key,
                                             const uint32_t* __restrict__ ubk, KeyFmt f,
                                             const hn_vrec* __restrict__ rec, TouchInfo& t) {
    const int32_t v = f.src(seg_key[j]);
    const hn_vrec r = rec[v];
    const int32_t kj = (int32_t)(seg_start[j + 1] - seg_start[j]);
    t.old_off = r.off;
    t.old_deg = r.deg;
    t.old_cap = r.cap_log2;
    t.new_deg = r.deg + kj;
    t.inplace = (r.cap_log2 >= 0 && (int64_t)t.new_deg <= (1ll << r.cap_log2)) ? 1 : 0;
    if (t.inplace) {
        t.new_cap = r.cap_log2;
        t.new_off = r.off;
        // old entries up to the smallest batch value keep their slots
        t.p = (int32_t)ubk[seg_start[j]];
    } else {
        t.new_cap = class_of(t.new_deg);
        t.new_off = 0;
        t.p = 0;
    }
}

// allocation kernels run 1024-lane workgroups: one global atomic per size class
// and workgroup (alloc_block_group) instead of per wave
constexpr int ALLOC_BLOCK = 1024;

__global__ __launch_bounds__(ALLOC_BLOCK) void insert_alloc_kernel(
    const uint32_t* __restrict__ ucnt, const uint32_t* __restrict__ seg_start, const uint64_t* __restrict__ seg_key,
    const uint32_t* __restrict__ ubk, KeyFmt f, const hn_vrec* __restrict__ rec, AllocView A,
    TouchInfo* __restrict__ ti, uint32_t* __restrict__ stamp, const UpdScal* sc) {
    __shared__ AllocShared ash;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = j < (int64_t)*ucnt && !sc->err;   // every lane reaches the allocation below
    TouchInfo t{};
    if (active) insert_touch(j, seg_start, seg_key, ubk, f, rec, t);
    if (active && stamp) stamp[f.src(seg_key[j])] = seg_start[j];
    const bool need = active && !t.inplace;
    const int64_t blk = alloc_block_group(A, need ? t.new_cap : 0, need, ash);
    if (!active) return;
    if (need) t.new_off = blk;
    ti[j] = t;
}


// commit or roll back the tickets of this batch
__global__ void alloc_fix_kernel(AllocState* st, UpdScal* sc) {
    if (st->error) {
        st->want_units = st->bump_units - st->saved_bump;   // what the batch asked for (the arena grows by it)
        st->bump_units = st->saved_bump;
        for (int k = 0; k < NCLASS; k++) st->take[k] = 0;
        st->error = 0;
        sc->err |= ERR_OOM;
        return;
    }
    if (sc->err) {   // bad input: nothing was allocated
        for (int k = 0; k < NCLASS; k++) st->take[k] = 0;
        st->bump_units = st->saved_bump;
        return;
    }
    for (int k = 0; k < NCLASS; k++) {
        unsigned fc = st->free_count[k], t = st->take[k];
        st->free_count[k] = t >= fc ? 0u : fc - t;
        st->take[k] = 0;
    }
    st->saved_bump = st->bump_units;
}

// Scan output of the segmented copies' run lengths: P[i] = first item of run
// i, and tile_lo[t] = the run holding item t * TILE, so a copy tile finds its
// runs without a search (a run spanning several tiles writes them all).
template <int TILE>
struct RunOutT {
    int64_t* P;
    uint32_t* tile_lo;
    __device__ __forceinline__ void operator()(int64_t i, int64_t ex, int64_t x) const {
        P[i] = ex;
        if (x) {
            const int64_t t0 = (ex + TILE - 1) / TILE, t1 = (ex + x + TILE - 1) / TILE;
            for (int64_t t = t0; t < t1; t++) tile_lo[t] = (uint32_t)i;
        }
    }
};
using RunOut = RunOutT<COPY_TILE>;

// Segmented copies: run s moves P[s+1] - P[s] entries from slot src[s] on to
// slot dst[s] on.  A lane's items are located first, then all loads issue
// (unconditionally: a lane past the end reads slot 0), then all stores.
constexpr int COPY_CAP = LbShared<COPY_CAP_IPT, int64_t>::CAP;   // runs staged per copy tile

struct CopyShared {
    LbShared<COPY_CAP_IPT, int64_t> lb;   // base = src[s] - P[s]
    int64_t dst[COPY_CAP];
};

// a tile's run range [lo, hi] (hi at or past the run holding its last item)
// and its staged run count (> 0), or -(span) when the span overflows the stage
struct CopyRange {
    int64_t lo;
    int count;
};

__device__ __forceinline__ CopyRange copy_range(const int64_t* __restrict__ P, int64_t nruns, int64_t W,
                                                int64_t ntiles, int64_t tile, int64_t lo, int64_t next_lo) {
    // the run holding the next tile's first item bounds this tile's runs; the
    // last tile searches (the whole workgroup takes this branch together)
    const int64_t hi = tile + 1 < ntiles ? next_lo : block_upper_bound_m1(P, lo, nruns + 1, W - 1);
    const int64_t span = hi - lo + 1;
    return {lo, span <= COPY_CAP ? (int)span : -(int)(span < 0x7fffffff ? span : 0x7fffffff)};
}

// publish the new records and release the blocks that were left
__global__ __launch_bounds__(ALLOC_BLOCK) void commit_kernel(const uint32_t* __restrict__ ucnt,
                                                             const uint64_t* __restrict__ seg_key, KeyFmt f,
                                                             const TouchInfo* __restrict__ ti,
                                                             const uint8_t* __restrict__ skip,
                                                             hn_vrec* __restrict__ rec, AllocView A,
                                                             const UpdScal* sc) {
    __shared__ AllocShared ash;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = j < (int64_t)*ucnt && !sc->err;   // every thread reaches the release below
    bool rel = false;
    TouchInfo t{};
    if (active) {
        const bool live = !(skip && skip[j]);
        t = ti[j];
        if (live) {
            hn_vrec r;
            r.off = t.new_cap >= 0 ? t.new_off : 0;
            r.deg = t.new_deg;
            r.cap_log2 = t.new_cap;
            rec[f.src(seg_key[j])] = r;
        }
        rel = live && !t.inplace && t.old_cap >= 0;
    }
    free_block_group(A, rel ? t.old_cap : 0, t.old_off, rel, ash);
}

// Undo the allocation of a batch that fails after alloc_fix_kernel committed
// its tickets and before commit_kernel: the blocks handed to relocating
// vertices go back to the free stacks (a no-op when the batch already failed
// on the device, where the allocation was rolled back).
__global__ __launch_bounds__(ALLOC_BLOCK) void release_new_blocks_kernel(const uint32_t* __restrict__ ucnt,
                                                                         const TouchInfo* __restrict__ ti,
                                                                         const uint8_t* __restrict__ skip,
                                                                         AllocView A, const UpdScal* sc) {
    __shared__ AllocShared ash;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = j < (int64_t)*ucnt && !sc->err;   // every thread reaches the release below
    bool rel = false;
    TouchInfo t{};
    if (active) {
        t = ti[j];
        rel = !(skip && skip[j]) && !t.inplace && t.new_cap >= 0;
    }
    free_block_group(A, rel ? t.new_cap : 0, t.new_off, rel, ash);
}

// ---------------------------------------------------------------- erase
// per distinct (u,v,c): stored copies from lb = lower_bound(v), deleted count
// min(c, stored) and whether a copy survives (more than c stored).  The copies
// are counted forward from lb, at most c + 1 of them: c is bounded by the batch.
__global__ void erase_locate_kernel(const uint32_t* __restrict__ dcnt, const uint32_t* __restrict__ dstart,
                                    const uint64_t* __restrict__ dkey, KeyFmt f, const hn_vrec* __restrict__ rec,
                                    const int32_t* __restrict__ arena, uint32_t* __restrict__ dpos,
                                    uint32_t* __restrict__ dlen, uint8_t* __restrict__ live, const UpdScal* sc) {
    int64_t jj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= (int64_t)*dcnt) return;
    if (sc->err) {
        dlen[jj] = 0;
        dpos[jj] = 0;
        live[jj] = 0;
        return;
    }
    const uint64_t x = dkey[jj];
    const hn_vrec r = rec[f.src(x)];
    const int32_t d = f.dst(x);
    const int32_t* row = arena + r.off;
    const int64_t lb = lower_bound_i32(row, r.deg, d);
    const int64_t c = dstart[jj + 1] - dstart[jj];
    const int64_t lim = c + 1 < r.deg - lb ? c + 1 : r.deg - lb;
    int64_t e = 0;
    while (e < lim && row[lb + e] == d) e++;
    dpos[jj] = (uint32_t)lb;
    dlen[jj] = (uint32_t)(c < e ? c : e);
    live[jj] = e > c ? 1 : 0;
}

struct DLenIn {
    const uint32_t* dlen;
    const uint32_t* dcnt;
    __device__ __forceinline__ uint32_t operator()(int64_t i) const { return i < (int64_t)*dcnt ? dlen[i] : 0u; }
};

// per touched source u (segments of the distinct keys): deletions; in place
// unless the list drops to a smaller size class (the reference shrinks when
// new_degree <= limit/2, BatchUpdateKernels.cuh:180-186) or empties
// per touched source u: deletions, in place / smaller block / no block;
// returns whether a new block is needed
__device__ __forceinline__ bool erase_touch(int64_t u, const uint32_t* __restrict__ ustart,
                                            const uint64_t* __restrict__ ukey, KeyFmt f,
                                            const uint32_t* __restrict__ dpos, const uint32_t* __restrict__ dlen,
                                            const uint32_t* __restrict__ dlp, const hn_vrec* __restrict__ rec,
                                            TouchInfo& t, uint8_t* __restrict__ skip) {
    const int32_t v = f.src(ukey[u]);
    const hn_vrec r = rec[v];
    const uint32_t del = dlp[ustart[u + 1]] - dlp[ustart[u]];
    t.old_off = r.off;
    t.old_deg = r.deg;
    t.old_cap = r.cap_log2;
    t.new_deg = r.deg - (int32_t)del;
    t.p = 0;
    t.inplace = 0;
    t.new_off = 0;
    t.new_cap = class_of(t.new_deg);
    skip[u] = del == 0 ? 1 : 0;
    bool need = false;
    if (del > 0 && t.new_cap >= 0) {
        if (t.new_cap == r.cap_log2) {
            t.inplace = 1;
            t.new_off = r.off;
            // first deleted slot: the first distinct key of u that removes entries
            for (uint32_t jj = ustart[u]; jj < ustart[u + 1]; jj++) {
                if (dlen[jj]) {
                    t.p = (int32_t)dpos[jj];
                    break;
                }
            }
        } else {
            need = true;
        }
    }
    return need;
}

__global__ __launch_bounds__(ALLOC_BLOCK) void erase_alloc_kernel(
    const uint32_t* __restrict__ ucnt, const uint32_t* __restrict__ ustart, const uint64_t* __restrict__ ukey,
    KeyFmt f, const uint32_t* __restrict__ dpos, const uint32_t* __restrict__ dlen, const uint32_t* __restrict__ dlp,
    const hn_vrec* __restrict__ rec, AllocView A, TouchInfo* __restrict__ ti, uint8_t* __restrict__ skip,
    uint32_t* __restrict__ stamp, const UpdScal* sc) {
    __shared__ AllocShared ash;
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = u < (int64_t)*ucnt && !sc->err;   // every lane reaches the allocation below
    TouchInfo t{};
    bool need = false;
    if (active) need = erase_touch(u, ustart, ukey, f, dpos, dlen, dlp, rec, t, skip);
    if (active && stamp) stamp[f.src(ukey[u])] = ustart[u];
    const int64_t blk = alloc_block_group(A, need ? t.new_cap : 0, need, ash);
    if (!active) return;
    if (need) t.new_off = blk;
    ti[u] = t;
}


// Survivors move as runs: between two deleted ranges of a touched vertex u
// (and after the last one) the old entries keep their order and shift down by
// the number deleted before them.  The runs of u are consecutive: run jj + u
// (before distinct key jj), then u's tail run at ustart[u + 1] + u, so the
// items of every vertex are contiguous and in slot order.  A run copies
// [x0, x0 + len) of u's old block to x0 - shift in u's (new or same) block;
// runs of a vertex rewritten in place carry RUN_INPLACE.  rtail[run] names
// the tail run of the run's vertex.  One thread per distinct key and per
// touched vertex; key jj belongs to vertex tpos[jj + 1] - 1.
constexpr int64_t RUN_INPLACE = (int64_t)1 << 61;

__device__ __forceinline__ void erase_run(int64_t g, int64_t prev, int64_t end, int64_t shift, bool live,
                                          const TouchInfo& t, int64_t* __restrict__ glen, int64_t* __restrict__ gsrc,
                                          int64_t* __restrict__ gdst) {
    const int64_t p = t.p;
    const int64_t start = prev > p ? prev : p;
    const int64_t len = (live && end > start) ? end - start : 0;
    glen[g] = len;
    if (len) {
        gsrc[g] = (t.old_off + start) | (t.inplace ? RUN_INPLACE : 0);
        gdst[g] = t.new_off + start - shift;
    }
}

__global__ void erase_gaps_kernel(const uint32_t* __restrict__ dcnt, const uint32_t* __restrict__ ucnt,
                                  const uint32_t* __restrict__ ustart, const uint32_t* __restrict__ tpos,
                                  const uint32_t* __restrict__ dpos, const uint32_t* __restrict__ dlen,
                                  const uint32_t* __restrict__ dlp, const TouchInfo* __restrict__ ti,
                                  const uint8_t* __restrict__ skip, int64_t* __restrict__ glen,
                                  int64_t* __restrict__ gsrc, int64_t* __restrict__ gdst,
                                  uint32_t* __restrict__ rtail, const UpdScal* sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)*dcnt) {   // the run before distinct key i
        const int64_t u = (int64_t)tpos[i + 1] - 1;
        const uint32_t a = ustart[u];
        const int64_t prev = i > (int64_t)a ? (int64_t)dpos[i - 1] + dlen[i - 1] : 0;
        erase_run(i + u, prev, (int64_t)dpos[i], (int64_t)(dlp[i] - dlp[a]), !sc->err && !skip[u], ti[u], glen, gsrc,
                  gdst);
        rtail[i + u] = (uint32_t)(ustart[u + 1] + u);
    }
    if (i < (int64_t)*ucnt) {   // the tail run of touched vertex i
        const uint32_t a = ustart[i], b = ustart[i + 1];
        const TouchInfo t = ti[i];
        const int64_t prev = b > a ? (int64_t)dpos[b - 1] + dlen[b - 1] : 0;
        erase_run(b + i, prev, (int64_t)t.old_deg, (int64_t)(dlp[b] - dlp[a]), !sc->err && !skip[i], t, glen, gsrc,
                  gdst);
        rtail[b + i] = (uint32_t)(b + i);
    }
}

// run lengths: dcnt + ucnt runs
struct GapLenIn {
    const int64_t* glen;
    const uint32_t* dcnt;
    const uint32_t* ucnt;
    __device__ __forceinline__ int64_t operator()(int64_t g) const {
        return g < (int64_t)*dcnt + (int64_t)*ucnt ? glen[g] : 0;
    }
};

// The erase compaction runs in place, in one pass over the survivors, except
// for a few deferred entries.  Within a vertex the items are in slot order and
// every entry moves down (destination <= source), so an entry's destination
// can only be the source of an earlier item of its vertex.  A copy tile reads
// all its sources before it writes (a barrier), which settles its own items;
// the earlier items of its first vertex that lie in earlier tiles -- read in
// any order by other workgroups -- have sources up to s = source(i0 - 1).  The
// tile's first items whose destination is <= s are deferred: written to a
// side buffer with their destinations and placed by a second launch once
// every tile has read.  All other destinations lie above every source of an
// earlier tile and below every source of a later one.  Deferred entries are a
// prefix of the tile, within its first vertex, at most the number of entries
// deleted before them in that vertex; measured at 1-2 % of the moved entries
// on R-MAT graphs, where the staged alternative copies every moved entry
// twice.
__global__ void erase_defer_kernel(int64_t ntile_max, int64_t nruns, const int64_t* __restrict__ P,
                                   const uint32_t* __restrict__ tile_lo, const int64_t* __restrict__ src,
                                   const int64_t* __restrict__ dst, const uint32_t* __restrict__ rtail,
                                   uint32_t* __restrict__ tcnt, const UpdScal* sc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntile_max) return;
    const int64_t W = P[nruns];
    const int64_t ntiles = (W + ERASE_TILE - 1) / ERASE_TILE;
    int64_t c = 0;
    if (!sc->err && t > 0 && t < ntiles) {
        const int64_t i0 = t * ERASE_TILE;
        const int64_t r = tile_lo[t];                                   // holds i0
        const int64_t rp = upper_bound_m1<int64_t>(P, 0, r + 1, i0 - 1);  // holds i0 - 1
        if ((src[r] & RUN_INPLACE) && rtail[r] == rtail[rp]) {
            const int64_t d0 = dst[r] - P[r] + i0;
            const int64_t sprev = (src[rp] & ~RUN_INPLACE) - P[rp] + (i0 - 1);
            if (sprev >= d0) {
                c = sprev - d0 + 1;
                const int64_t vend = P[rtail[r] + 1];   // end of the vertex's items
                c = c < vend - i0 ? c : vend - i0;
                c = c < (int64_t)ERASE_TILE ? c : (int64_t)ERASE_TILE;
            }
        }
    }
    tcnt[t] = (uint32_t)c;
}

struct TileCountIn {
    const uint32_t* c;
    __device__ __forceinline__ int64_t operator()(int64_t t) const { return (int64_t)c[t]; }
};

// scan output of the deferred counts: each tile's offset in the side buffer;
// the last tile also records the total in the batch's status word, read back
// with it
struct SideOut {
    int64_t* P;
    int64_t last;
    UpdScal* sc;
    __device__ __forceinline__ void operator()(int64_t i, int64_t ex, int64_t x) const {
        P[i] = ex;
        if (i == last) sc->side = ex + x;
    }
};

// s_waitcnt vmcnt(0): every vector memory load of the wave has returned
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One pass of the in-place compaction: tile t's first tcnt[t] items go to the
// side buffer at doff[t] (entry and destination), every other item to its
// destination.  Sources are read (values, then each meta-data field) by the
// whole tile before any of its writes.  META = false is the variant for
// graphs without edge meta-data (am.n == 0): without the field loop it fits
// in 64 VGPRs, so twice as many waves per SIMD are resident.
template <bool META>
__global__ __launch_bounds__(BLOCK) void erase_copy_kernel(int64_t nruns, const int64_t* __restrict__ P,
                                                           const uint32_t* __restrict__ tile_lo,
                                                           const int64_t* __restrict__ src,
                                                           const int64_t* __restrict__ dst,
                                                           const uint32_t* __restrict__ tcnt,
                                                           const int64_t* __restrict__ doff,
                                                           int32_t* __restrict__ arena, int32_t* __restrict__ side,
                                                           int64_t* __restrict__ side_dst, const UpdScal* sc,
                                                           MetaSet am, MetaSet sm) {
    __shared__ CopyShared sh;
    if (sc->err) return;
    const int64_t W = P[nruns];
    const int64_t ntiles = (W + ERASE_TILE - 1) / ERASE_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t w0 = tile * ERASE_TILE;
        const int64_t wend = (w0 + (int64_t)ERASE_TILE) < W ? (w0 + (int64_t)ERASE_TILE) : W;
        const CopyRange rg = copy_range(P, nruns, W, ntiles, tile, tile_lo[tile],
                                        tile + 1 < ntiles ? (int64_t)tile_lo[tile + 1] : 0);
        if (rg.count > 0)
            for (int i = threadIdx.x; i < rg.count; i += BLOCK) {
                const int64_t s = rg.lo + i;
                sh.lb.pref[i] = P[s];
                sh.lb.base[i] = (src[s] & ~RUN_INPLACE) - P[s];
                sh.dst[i] = dst[s] - P[s];
            }
        const int64_t ndef = tcnt[tile];
        const int64_t o = doff[tile];
        __syncthreads();
        int64_t si[ERASE_IPT], di[ERASE_IPT];
#pragma unroll
        for (int j = 0; j < ERASE_IPT; j++) {
            const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            if (rg.count > 0) {
                const int s = lb_locate<COPY_CAP_IPT>(sh.lb, rg.count, w);
                si[j] = sh.lb.base[s] + w;
                di[j] = sh.dst[s] + w;
            } else {
                const int64_t s = upper_bound_m1<int64_t>(P, rg.lo, rg.lo + (int64_t)(-rg.count), w);
                si[j] = (src[s] & ~RUN_INPLACE) - P[s] + w;
                di[j] = dst[s] - P[s] + w;
            }
            if (w >= wend) si[j] = 0, di[j] = -1;
        }
        int32_t v[ERASE_IPT];
#pragma unroll
        for (int j = 0; j < ERASE_IPT; j++) v[j] = arena[si[j]];
        // the tile's sources are all read before any slot is rewritten: the
        // loads have returned (the barrier alone does not wait for them)
        wait_loads();
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ERASE_IPT; j++) {
            if (di[j] < 0) continue;
            const int64_t k = (int64_t)j * BLOCK + threadIdx.x;   // item within the tile
            if (k < ndef) {
                side[o + k] = v[j];
                side_dst[o + k] = di[j];
            } else {
                arena[di[j]] = v[j];
            }
        }
        for (int f = 0; META && f < am.n; f++) {
            const bool wide = am.bytes[f] == 8;
            uint64_t mv[ERASE_IPT];
#pragma unroll
            for (int j = 0; j < ERASE_IPT; j++)
                mv[j] = wide ? reinterpret_cast<const uint64_t*>(am.p[f])[si[j]]
                             : reinterpret_cast<const uint32_t*>(am.p[f])[si[j]];
            wait_loads();
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ERASE_IPT; j++) {
                if (di[j] < 0) continue;
                const int64_t k = (int64_t)j * BLOCK + threadIdx.x;   // item within the tile
                char* b = k < ndef ? sm.p[f] : am.p[f];
                const int64_t x = k < ndef ? o + k : di[j];
                if (wide) reinterpret_cast<uint64_t*>(b)[x] = mv[j];
                else reinterpret_cast<uint32_t*>(b)[x] = (uint32_t)mv[j];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- insert, moves
// The old entries that move: A_j[p, old_deg) of touched vertex j, items
// aP[j] .. aP[j + 1] in slot order.  Entry r goes to slot r + #(B_j < A_j[r])
// of j's block (the same block in place; the new one when relocating, where
// p = 0); the batch entries fill the slots left between them.
struct MoveLenIn {
    const TouchInfo* ti;
    const uint32_t* ucnt;
    const UpdScal* sc;
    __device__ __forceinline__ int64_t operator()(int64_t j) const {
        if (j >= (int64_t)*ucnt || sc->err) return 0;
        return (int64_t)(ti[j].old_deg - ti[j].p);
    }
};

// #(B < x) over nb sorted batch keys
__device__ __forceinline__ int64_t count_below(const uint64_t* __restrict__ B, int64_t nb, KeyFmt f, int32_t x) {
    int64_t lo = 0, hi = nb;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (f.dst(B[mid]) < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The move mirrors the erase compaction: every moving entry goes up
// (destination > source), so its destination can only be the source of a
// later item of its vertex.  A copy tile reads all its sources before it
// writes; the items of its last vertex in later tiles have sources from
// source(i1) on (i1 = the next tile's first item).  The tile's last items
// whose destination reaches source(i1) are deferred: at most
// c = #(B < A[r_last]) of them, since r + #(B < A[r]) > r_last needs
// r > r_last - c.  They are placed once every tile has read.
__global__ void insert_defer_kernel(int64_t ntile_max, int64_t nseg, const int64_t* __restrict__ aP,
                                    const uint32_t* __restrict__ tile_lo, const TouchInfo* __restrict__ ti,
                                    const uint32_t* __restrict__ seg_start, const uint64_t* __restrict__ keys,
                                    KeyFmt f, const int32_t* __restrict__ arena, uint32_t* __restrict__ tcnt,
                                    const UpdScal* sc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntile_max) return;
    const int64_t W = aP[nseg];
    const int64_t ntiles = (W + COPY_TILE - 1) / COPY_TILE;
    int64_t c = 0;
    if (!sc->err && t + 1 < ntiles) {
        const int64_t i1 = (t + 1) * COPY_TILE;
        const int64_t j = tile_lo[t + 1];   // holds i1
        if (aP[j] < i1) {                   // and i1 - 1, the tile's last item
            const TouchInfo tj = ti[j];
            if (tj.inplace) {
                const int64_t r_last = tj.p + (i1 - 1 - aP[j]);
                c = count_below(keys + seg_start[j], tj.new_deg - tj.old_deg, f, arena[tj.old_off + r_last]);
                const int64_t first = aP[j] > t * COPY_TILE ? aP[j] : t * COPY_TILE;
                c = c < i1 - first ? c : i1 - first;
            }
        }
    }
    tcnt[t] = (uint32_t)c;
}

struct MoveShared {
    LbShared<COPY_CAP_IPT, int64_t> lb;   // pref = aP[j], base = old_off + p - aP[j]
    int64_t dbase[COPY_CAP];               // new_off + p - aP[j]
    uint32_t bstart[COPY_CAP];             // j's batch segment
    uint32_t bn[COPY_CAP];
};

// One pass of the move: tile t's last tcnt[t] items go to the side buffer at
// doff[t] (entry and destination), every other item to its destination.
// Sources are read (values, then each meta-data field) by the whole tile
// before any of its writes.
__global__ __launch_bounds__(BLOCK) void insert_move_kernel(int64_t nseg, const int64_t* __restrict__ aP,
                                                            const uint32_t* __restrict__ tile_lo,
                                                            const TouchInfo* __restrict__ ti,
                                                            const uint32_t* __restrict__ seg_start,
                                                            const uint64_t* __restrict__ keys, KeyFmt f,
                                                            const uint32_t* __restrict__ tcnt,
                                                            const int64_t* __restrict__ doff,
                                                            int32_t* __restrict__ arena, int32_t* __restrict__ side,
                                                            int64_t* __restrict__ side_dst, const UpdScal* sc,
                                                            MetaSet am, MetaSet sm) {
    __shared__ MoveShared sh;
    if (sc->err) return;
    const int64_t W = aP[nseg];
    const int64_t ntiles = (W + COPY_TILE - 1) / COPY_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t w0 = tile * COPY_TILE;
        const int64_t wend = (w0 + (int64_t)COPY_TILE) < W ? (w0 + (int64_t)COPY_TILE) : W;
        const CopyRange rg = copy_range(aP, nseg, W, ntiles, tile, tile_lo[tile],
                                        tile + 1 < ntiles ? (int64_t)tile_lo[tile + 1] : 0);
        if (rg.count > 0)
            for (int i = threadIdx.x; i < rg.count; i += BLOCK) {
                const int64_t j = rg.lo + i;
                const TouchInfo t = ti[j];
                sh.lb.pref[i] = aP[j];
                sh.lb.base[i] = t.old_off + t.p - aP[j];
                sh.dbase[i] = t.new_off + t.p - aP[j];
                sh.bstart[i] = seg_start[j];
                sh.bn[i] = (uint32_t)(t.new_deg - t.old_deg);
            }
        const int64_t first_def = wend - (int64_t)tcnt[tile];   // items from here on are deferred
        const int64_t o = doff[tile];
        __syncthreads();
        int64_t si[COPY_IPT], di[COPY_IPT];
        uint32_t bs[COPY_IPT], bn[COPY_IPT];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) {
            const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            si[j] = 0;
            di[j] = -1;
            bs[j] = 0;
            bn[j] = 0;
            if (w < wend) {
                if (rg.count > 0) {
                    const int s = lb_locate<COPY_CAP_IPT>(sh.lb, rg.count, w);
                    si[j] = sh.lb.base[s] + w;
                    di[j] = sh.dbase[s] + w;
                    bs[j] = sh.bstart[s];
                    bn[j] = sh.bn[s];
                } else {
                    const int64_t s = upper_bound_m1<int64_t>(aP, rg.lo, rg.lo + (int64_t)(-rg.count), w);
                    const TouchInfo t = ti[s];
                    si[j] = t.old_off + t.p - aP[s] + w;
                    di[j] = t.new_off + t.p - aP[s] + w;
                    bs[j] = seg_start[s];
                    bn[j] = (uint32_t)(t.new_deg - t.old_deg);
                }
            }
        }
        int32_t v[COPY_IPT];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) v[j] = arena[si[j]];
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++)
            if (di[j] >= 0) di[j] += count_below(keys + bs[j], bn[j], f, v[j]);
        // the tile's sources are all read before any slot is rewritten
        wait_loads();
        __syncthreads();
#pragma unroll
        for (int j = 0; j < COPY_IPT; j++) {
            if (di[j] < 0) continue;
            const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
            if (w >= first_def) {
                side[o + (w - first_def)] = v[j];
                side_dst[o + (w - first_def)] = di[j];
            } else {
                arena[di[j]] = v[j];
            }
        }
        for (int fl = 0; fl < am.n; fl++) {
            const bool wide = am.bytes[fl] == 8;
            uint64_t mv[COPY_IPT];
#pragma unroll
            for (int j = 0; j < COPY_IPT; j++)
                mv[j] = wide ? reinterpret_cast<const uint64_t*>(am.p[fl])[si[j]]
                             : reinterpret_cast<const uint32_t*>(am.p[fl])[si[j]];
            wait_loads();
            __syncthreads();
#pragma unroll
            for (int j = 0; j < COPY_IPT; j++) {
                if (di[j] < 0) continue;
                const int64_t w = w0 + (int64_t)j * BLOCK + threadIdx.x;
                char* b = w >= first_def ? sm.p[fl] : am.p[fl];
                const int64_t x = w >= first_def ? o + (w - first_def) : di[j];
                if (wide) reinterpret_cast<uint64_t*>(b)[x] = mv[j];
                else reinterpret_cast<uint32_t*>(b)[x] = (uint32_t)mv[j];
            }
        }
        __syncthreads();
    }
}

// the deferred entries of either in-place pass (and their fields) to their
// destinations, once every copy tile has read its sources
__global__ void place_side_kernel(int64_t m, const int32_t* __restrict__ side, const int64_t* __restrict__ side_dst,
                                    int32_t* __restrict__ arena, MetaSet am, MetaSet sm) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int64_t d = side_dst[k];
    arena[d] = side[k];
    if (am.any()) meta_put(am, d, sm, k);
}

// kept batch entry i (the q-th of its source segment j) to slot
// q + #(A_j <= value) of j's block (stored copies first on ties), with its
// fields from batch position bidx[i]
__global__ void insert_batch_kernel(const uint32_t* __restrict__ nb, const uint64_t* __restrict__ keys,
                                    const uint32_t* __restrict__ spos, const uint32_t* __restrict__ seg_start,
                                    const uint32_t* __restrict__ ubk, const TouchInfo* __restrict__ ti, KeyFmt f,
                                    int32_t* __restrict__ arena, const UpdScal* sc, MetaSet am, MetaSet bm,
                                    const uint32_t* __restrict__ bidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)*nb || sc->err) return;
    const int64_t j = (int64_t)spos[i + 1] - 1;
    const int64_t dest = ti[j].new_off + (i - (int64_t)seg_start[j]) + ubk[i];
    arena[dest] = f.dst(keys[i]);
    if (am.any()) meta_put(am, dest, bm, bidx[i]);
}

}  // namespace

// Resolves this file's code object (see preload_code_objects in graph.hip).
void preload_update_code() {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&pack_batch_kernel)) != hipSuccess) (void)hipGetLastError();
}

// scratch for a batch of n entries
struct UpdBufs {
    uint64_t* ka;
    uint64_t* kb;
    uint32_t* pos;       // n+1
    uint32_t* spos;      // n+1
    uint32_t* seg_start; // n+1
    uint64_t* seg_key;   // n+1
    uint32_t* seg2_start;
    uint64_t* seg2_key;
    uint32_t* tpos;      // n+1
    uint32_t* dpos;      // n
    uint32_t* dlen;      // n
    uint32_t* dlp;       // n+1
    int64_t* wpos;       // n+1
    int64_t* stpos;      // n+1
    TouchInfo* ti;       // n
    uint8_t* skip;       // n
    uint8_t* live;       // n: erase, a copy of the distinct key survives
    int64_t* glen;       // 2n: erase runs (distinct keys, then tail runs)
    int64_t* gsrc;       // 2n
    int64_t* gdst;       // 2n
    int64_t* gP;         // 2n+1
    uint32_t* tile_lo;   // first run of each copy tile
    uint32_t* rtail;     // 2n: erase, the tail run of each run's vertex
    uint32_t* tcnt;      // per copy tile: entries deferred
    int64_t* doff;       // per copy tile + 1: side-buffer offset of the deferred entries
    UpdScal* sc;
    void* tmp;
    uint32_t* ia;        // n+1: batch positions (edge meta-data only)
    uint32_t* ib;        // n+1
    MetaSet sm;          // side buffer of the edge meta-data (deferred entries)
    int64_t ntile_max;   // copy tiles the buffers above are sized for
};

static hn_status upd_buffers(hn_graph* g, int64_t n, UpdBufs* b) {
    auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
    const int64_t n1 = n + 1;
    int64_t o = 0;
    int64_t o_ka = o; o = al(o + n1 * 8);
    int64_t o_kb = o; o = al(o + n1 * 8);
    int64_t o_pos = o; o = al(o + n1 * 4);
    int64_t o_spos = o; o = al(o + n1 * 4);
    int64_t o_ss = o; o = al(o + n1 * 4);
    int64_t o_sk = o; o = al(o + n1 * 8);
    int64_t o_ss2 = o; o = al(o + n1 * 4);
    int64_t o_sk2 = o; o = al(o + n1 * 8);
    int64_t o_tpos = o; o = al(o + n1 * 4);
    int64_t o_dpos = o; o = al(o + n1 * 4);
    int64_t o_dlen = o; o = al(o + n1 * 4);
    int64_t o_dlp = o; o = al(o + n1 * 4);
    int64_t o_wpos = o; o = al(o + n1 * 8);
    int64_t o_stpos = o; o = al(o + n1 * 8);
    int64_t o_ti = o; o = al(o + n1 * (int64_t)sizeof(TouchInfo));
    int64_t o_skip = o; o = al(o + n1);
    int64_t o_live = o; o = al(o + n1);
    int64_t o_gl = o; o = al(o + 2 * n1 * 8);
    int64_t o_gs = o; o = al(o + 2 * n1 * 8);
    int64_t o_gd = o; o = al(o + 2 * n1 * 8);
    int64_t o_gp = o; o = al(o + (2 * n1 + 1) * 8);
    // staged / moved entries never exceed the stored ones
    const int64_t ntile_max = (g->ne + n) / COPY_TILE + 2;
    int64_t o_tl = o; o = al(o + ntile_max * 4);
    int64_t o_rt = o; o = al(o + 2 * n1 * 4);
    int64_t o_tc = o; o = al(o + ntile_max * 4);
    int64_t o_do = o; o = al(o + (ntile_max + 1) * 8);
    int64_t o_sc = o; o = al(o + 64);
    const bool meta = g->meta.any();
    int64_t o_ia = o; o = al(o + (meta ? n1 * 4 : 0));
    int64_t o_ib = o; o = al(o + (meta ? n1 * 4 : 0));
    int64_t o_tmp = o;
    o = al(o + (int64_t)std::max(sort_temp_bytes(n1), scan_temp_bytes(std::max(2 * n1, ntile_max))));
    HN_TRY(g->s_upd.reserve((size_t)o));
    char* base = g->s_upd.as<char>();
    b->ka = reinterpret_cast<uint64_t*>(base + o_ka);
    b->kb = reinterpret_cast<uint64_t*>(base + o_kb);
    b->pos = reinterpret_cast<uint32_t*>(base + o_pos);
    b->spos = reinterpret_cast<uint32_t*>(base + o_spos);
    b->seg_start = reinterpret_cast<uint32_t*>(base + o_ss);
    b->seg_key = reinterpret_cast<uint64_t*>(base + o_sk);
    b->seg2_start = reinterpret_cast<uint32_t*>(base + o_ss2);
    b->seg2_key = reinterpret_cast<uint64_t*>(base + o_sk2);
    b->tpos = reinterpret_cast<uint32_t*>(base + o_tpos);
    b->dpos = reinterpret_cast<uint32_t*>(base + o_dpos);
    b->dlen = reinterpret_cast<uint32_t*>(base + o_dlen);
    b->dlp = reinterpret_cast<uint32_t*>(base + o_dlp);
    b->wpos = reinterpret_cast<int64_t*>(base + o_wpos);
    b->stpos = reinterpret_cast<int64_t*>(base + o_stpos);
    b->ti = reinterpret_cast<TouchInfo*>(base + o_ti);
    b->skip = reinterpret_cast<uint8_t*>(base + o_skip);
    b->live = reinterpret_cast<uint8_t*>(base + o_live);
    b->glen = reinterpret_cast<int64_t*>(base + o_gl);
    b->gsrc = reinterpret_cast<int64_t*>(base + o_gs);
    b->gdst = reinterpret_cast<int64_t*>(base + o_gd);
    b->gP = reinterpret_cast<int64_t*>(base + o_gp);
    b->tile_lo = reinterpret_cast<uint32_t*>(base + o_tl);
    b->rtail = reinterpret_cast<uint32_t*>(base + o_rt);
    b->tcnt = reinterpret_cast<uint32_t*>(base + o_tc);
    b->doff = reinterpret_cast<int64_t*>(base + o_do);
    b->sc = reinterpret_cast<UpdScal*>(base + o_sc);
    b->tmp = base + o_tmp;
    b->ia = meta ? reinterpret_cast<uint32_t*>(base + o_ia) : nullptr;
    b->ib = meta ? reinterpret_cast<uint32_t*>(base + o_ib) : nullptr;
    b->sm = meta_like(g->meta, nullptr);
    b->ntile_max = ntile_max;
    return HN_OK;
}

// copy a host or device batch into the key buffer (packed, validated) and sort it
static hn_status load_and_sort(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n, uint32_t flags,
                               UpdBufs& b, KeyFmt f, hipStream_t st, uint64_t** sorted) {
    const int32_t* ds = src;
    const int32_t* dd = dst;
    int32_t* staged = nullptr;
    if (!(flags & HN_DEVICE_PTRS)) {
        HN_TRY(dalloc(&staged, (size_t)(2 * n)));
        HN_HIP(hipMemcpyAsync(staged, src, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
        HN_HIP(hipMemcpyAsync(staged + n, dst, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
        ds = staged;
        dd = staged + n;
    }
    HN_HIP(hipMemsetAsync(b.sc, 0, sizeof(UpdScal), st));
    pack_batch_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(n, ds, dd, g->nv, f.sbits, b.ka, b.ia, b.sc);
    HN_CHECK_LAUNCH();
    if (staged) {
        HN_HIP(hipStreamSynchronize(st));
        dfree(staged);
    }
    uint64_t* keys = b.ka;
    uint64_t* alt = b.kb;
    uint32_t* vals = b.ia;
    uint32_t* valt = b.ib;
    if (vals) HN_TRY(radix_sort_u64_kv(&keys, &alt, &vals, &valt, n, 2 * f.sbits, b.tmp, st));
    else HN_TRY(radix_sort_u64(&keys, &alt, n, 2 * f.sbits, b.tmp, st));
    if (keys != b.ka) {
        // keep the sorted keys (and positions) in ka / ia so kb / ib are free for the compaction
        HN_HIP(hipMemcpyAsync(b.ka, keys, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
        if (vals) HN_HIP(hipMemcpyAsync(b.ia, vals, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
    }
    *sorted = b.ka;
    return HN_OK;
}

static hn_status check_scal(hn_graph* g, UpdBufs& b, hipStream_t st, UpdScal* h) {
    HN_TRY(read_back(g, b.sc, sizeof(UpdScal), h, st));
    if (h->err & ERR_BAD_ID) {
        set_error("batch holds a vertex id outside [0, nv)");
        return HN_E_INVALID_ARG;
    }
    if (h->err & ERR_OOM) {
        set_error("edge arena exhausted (graph unchanged)");
        return HN_E_ARENA_FULL;   // the entry point grows the arena and runs the batch again
    }
    return HN_OK;
}

__global__ void store_changed_kernel(const uint32_t* __restrict__ total, UpdScal* sc) {
    if (!sc->err) sc->changed = *total;
}

// After a batch applied to a symmetric graph: pairs outside the batch kept
// their membership, so the adjacency stays symmetric iff every batch pair
// (u,w) has (u,w) and (w,u) both stored or both absent.
//
// The batch entries with source w are found through stamp[w], the position
// of w's first key in the sorted keys, written by this batch's allocation
// kernel.  A stamp is valid when the key there has source w and the key before
// it does not (stale values from earlier batches fail the test, so the array
// is never cleared); the search for (w, u) then gallops from there, within
// w's keys.
__device__ __forceinline__ bool batch_has(const uint32_t* __restrict__ stamp, const uint64_t* __restrict__ keys,
                                          int64_t n, KeyFmt f, int32_t w, int32_t u, int64_t* at) {
    const int64_t s = stamp[w];
    if (s >= n) return false;
    const uint64_t k0 = keys[s];
    if (f.src(k0) != w || (s > 0 && f.src(keys[s - 1]) == w)) return false;
    const uint64_t x = ((uint64_t)w << f.sbits) | (uint64_t)u;
    if (k0 >= x) {
        *at = s;
        return k0 == x;
    }
    // keys[lo] < x; gallop until keys[hi] >= x (keys past w's are larger)
    int64_t lo = s, hi = s + 1, step = 1;
    while (hi < n && keys[hi] < x) {
        lo = hi;
        step <<= 1;
        hi = s + step;
    }
    if (hi > n) hi = n;
    while (hi - lo > 1) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < x) lo = mid; else hi = mid;
    }
    *at = hi;
    return hi < n && keys[hi] == x;
}

// Insert: every kept pair (u,w) is stored.  (w,u) is stored too when the kept
// batch holds it; only otherwise is it looked up in w's list.
__global__ void insert_symmetry_kernel(const uint32_t* __restrict__ count, const uint64_t* __restrict__ keys,
                                       KeyFmt f, const uint32_t* __restrict__ stamp,
                                       const hn_vrec* __restrict__ rec, const int32_t* __restrict__ arena,
                                       UpdScal* sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)*count;
    if (i >= n || sc->err) return;
    const int32_t u = f.src(keys[i]), w = f.dst(keys[i]);
    int64_t at;
    if (batch_has(stamp, keys, n, f, w, u, &at)) return;
    const hn_vrec rw = rec[w];
    const int64_t pw = lower_bound_i32(arena + rw.off, rw.deg, u);
    if (!(pw < rw.deg && arena[rw.off + pw] == u)) sc->asym = 1u;   // every writer stores the same value
}

// Erase: (u,w) is stored afterwards iff a copy survived its deletions.  (w,u)
// is stored afterwards iff a copy of it survived, when the batch holds it, and
// otherwise iff it was stored before, which (the graph being symmetric) is iff
// (u,w) was stored before, i.e. iff (u,w) deleted something.  No list is read.
__global__ void erase_symmetry_kernel(const uint32_t* __restrict__ dcnt, const uint64_t* __restrict__ dkey, KeyFmt f,
                                      const uint32_t* __restrict__ dlen, const uint8_t* __restrict__ live,
                                      const uint32_t* __restrict__ stamp, UpdScal* sc) {
    const int64_t jj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)*dcnt;
    if (jj >= n || sc->err) return;
    const int32_t u = f.src(dkey[jj]), w = f.dst(dkey[jj]);
    int64_t at;
    const bool b = batch_has(stamp, dkey, n, f, w, u, &at) ? live[at] != 0 : dlen[jj] > 0;
    if ((live[jj] != 0) != b) sc->asym = 1u;
}

// the per-vertex segment stamps, when the graph is known symmetric (the only
// case whose batches check symmetry); null otherwise
static hn_status symmetry_stamp(hn_graph* g, uint32_t** stamp) {
    *stamp = nullptr;
    if (g->sym != 1) return HN_OK;
    HN_TRY(g->s_stamp.reserve(sizeof(uint32_t) * (size_t)(g->nv > 0 ? g->nv : 1)));
    *stamp = g->s_stamp.as<uint32_t>();
    return HN_OK;
}

// symmetry bookkeeping after a successful batch (was_sym: known symmetric before)
static void note_symmetry(hn_graph* g, bool was_sym, const UpdScal& h) {
    g->sym = was_sym ? (h.asym ? -1 : 1) : 0;
    g->sym_version = g->version;
}

// the side buffer for the batch's deferred entries (their count is read back
// with the error word, which is checked first), one per meta-data field in
// b.sm, and when asked one 64-bit destination per entry
static hn_status reserve_side(hn_graph* g, UpdBufs& b, hipStream_t st, int64_t* total, int32_t** stage,
                              int64_t** dst = nullptr) {
    UpdScal h;
    HN_TRY(check_scal(g, b, st, &h));
    // test hook: HORNET_HIP_FAIL_SIDE_RESERVE=1 makes this reservation fail as
    // a device allocation would, after the batch's blocks were allocated
    if (const char* e = std::getenv("HORNET_HIP_FAIL_SIDE_RESERVE"))
        if (e[0] == '1') {
            set_error("side buffer: allocation failure forced by HORNET_HIP_FAIL_SIDE_RESERVE");
            return HN_E_OOM;
        }
    *total = (int64_t)h.side;
    const size_t cnt = (size_t)(*total > 0 ? *total : 1);
    // entries, then (when asked for) one 64-bit destination per entry
    const size_t vbytes = (sizeof(int32_t) * cnt + 255) & ~size_t(255);
    HN_TRY(g->s_misc.reserve(vbytes + (dst ? sizeof(int64_t) * cnt : 0)));
    *stage = g->s_misc.as<int32_t>();
    if (dst) *dst = reinterpret_cast<int64_t*>(g->s_misc.as<char>() + vbytes);
    if (g->meta.any()) {
        // one side array per meta-data field, 256-byte aligned, in s_mstage
        size_t bytes = 0;
        for (int f = 0; f < g->meta.n; f++) bytes += (cnt * g->meta.bytes[f] + 255) & ~size_t(255);
        HN_TRY(g->s_mstage.reserve(bytes));
        size_t at = 0;
        for (int f = 0; f < g->meta.n; f++) {
            b.sm.p[f] = g->s_mstage.as<char>() + at;
            at += (cnt * g->meta.bytes[f] + 255) & ~size_t(255);
        }
    }
    return HN_OK;
}

static void record_last(hn_graph* g, int32_t kind, int64_t submitted, int64_t applied, const TouchInfo* ti,
                        const uint32_t* ucnt, const uint8_t* skip, const int64_t* moved) {
    g->last_kind = kind;
    g->last_submitted = submitted;
    g->last_applied = applied;
    g->last_ti = ti;
    g->last_ucnt = ucnt;
    g->last_skip = skip;
    g->last_moved = moved;
}

// [0] touched, [1] relocated, [2] sum of old degrees over the touched, [3] sum
// over the relocated of the old (insert) or new (erase) degree
__global__ void update_stats_kernel(const uint32_t* __restrict__ ucnt, const TouchInfo* __restrict__ ti,
                                    const uint8_t* __restrict__ skip, int erase,
                                    unsigned long long* __restrict__ acc) {
    unsigned long long rel = 0, td = 0, rd = 0;
    const int64_t n = *ucnt;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const TouchInfo t = ti[j];
        td += (unsigned long long)t.old_deg;
        if (!(skip && skip[j]) && !t.inplace && t.new_cap >= 0) {
            rel++;
            rd += (unsigned long long)(erase ? t.new_deg : t.old_deg);
        }
    }
    rel = wave_reduce_sum(rel);
    td = wave_reduce_sum(td);
    rd = wave_reduce_sum(rd);
    if (__lane_id() == 0) {
        if (rel) atomicAdd(acc + 1, rel);
        if (td) atomicAdd(acc + 2, td);
        if (rd) atomicAdd(acc + 3, rd);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(acc, (unsigned long long)n);
}

// A batch that finds the edge arena full changes nothing; the arena then
// grows (BlockArrayManager adds a BlockArray when a size class is full,
// hornet/include/Core/MemoryManager/BlockArray/BlockArray.i.cuh:138-161) and
// the batch runs again.  HN_E_OOM only when device memory is exhausted.
template <typename Fn>
static hn_status with_growth(hn_graph* g, void* stream, Fn once) {
    for (int attempt = 0;; attempt++) {
        const hn_status s = once();
        if (s != HN_E_ARENA_FULL) return s;
        if (attempt >= 16) {
            set_error("edge arena: batch still does not fit after growing");
            return HN_E_OOM;
        }
        HN_TRY(grow_arena(g, 0, as_stream(stream)));
    }
}

}  // namespace hn

using namespace hn;

extern "C" {


// device copies of a host batch's meta-data fields, released with the object
struct BatchMeta {
    MetaSet m;
    bool owned = false;
    ~BatchMeta() {
        if (owned)
            for (int f = 0; f < m.n; f++) dfree(m.p[f]);
    }
};

static hn_status insert_once(hn_graph* g, const int32_t* src, const int32_t* dst, const void* const* meta,
                             int64_t n, uint32_t flags, int64_t* inserted, void* stream) {
    HN_ARG(g, "null graph");
    HN_ARG(n >= 0, "negative batch size");
    if (inserted) *inserted = 0;
    HN_ARG(meta == nullptr || g->meta.any(), "the graph has no edge meta-data");
    if (n == 0) return HN_OK;
    HN_ARG(src && dst, "null batch");
    HN_ARG(n < (int64_t)0x7FFFFFFF, "batch larger than 2^31-1 entries");
    HN_ARG(g->ne + n < (int64_t)0xFFFFFFFF, "edge count would exceed 2^32-1");
    hipStream_t st = as_stream(stream);
    BatchMeta bm;
    bm.m = meta_like(g->meta, nullptr);   // null fields insert zeros
    if (meta) {
        for (int f = 0; f < g->meta.n; f++) HN_ARG(meta[f] != nullptr, "a meta-data array is null");
        if (flags & HN_DEVICE_PTRS) {
            for (int f = 0; f < g->meta.n; f++) bm.m.p[f] = const_cast<char*>(static_cast<const char*>(meta[f]));
        } else {
            bm.owned = true;
            for (int f = 0; f < g->meta.n; f++) {
                const size_t bytes = (size_t)n * (size_t)g->meta.bytes[f];
                HN_TRY(dmalloc(reinterpret_cast<void**>(&bm.m.p[f]), bytes));
                HN_HIP(hipMemcpyAsync(bm.m.p[f], meta[f], bytes, hipMemcpyHostToDevice, st));
            }
        }
    }
    UpdBufs b;
    HN_TRY(upd_buffers(g, n, &b));
    KeyFmt f{key_bits_for(g->nv)};
    uint64_t* keys = nullptr;
    HN_TRY(load_and_sort(g, src, dst, n, flags, b, f, st, &keys));
    const unsigned gridn = (unsigned)ceil_div(n + 1, BLOCK);
    // 1. place in the stored list and keep flag (batch / graph dedup) of every
    // entry -> compact into kb (keep flags in dpos, places in dlen, kept places
    // in dlp: the erase arrays are free in insert)
    insert_locate_kernel<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, st>>>(
        n, keys, f, (flags & HN_DEDUP_BATCH) ? 1 : 0, (flags & HN_DEDUP_GRAPH) ? 1 : 0, g->rec, g->arena, b.dpos,
        b.dlen, b.sc);
    HN_CHECK_LAUNCH();
    HN_TRY(exclusive_scan<uint32_t>(ArrayIn<uint32_t>{b.dpos}, b.pos, n, b.tmp, st));
    uint32_t* ubk = b.dlp;
    compact_keys_kernel<<<gridn, BLOCK, 0, st>>>(n, keys, b.pos, b.kb, b.ia, b.ib, b.dlen, ubk);
    HN_CHECK_LAUNCH();
    uint32_t* nb = b.pos + n;   // entries kept
    // 2. segments by source
    HN_TRY(exclusive_scan<uint32_t>(SegStartIn{b.kb, nb, f.sbits}, b.spos, n, b.tmp, st));
    seg_write_kernel<<<gridn, BLOCK, 0, st>>>(n, b.kb, nb, b.spos, b.seg_start, b.seg_key);
    HN_CHECK_LAUNCH();
    uint32_t* ucnt = b.spos + n;
    // 3. in place or new blocks
    uint32_t* stamp = nullptr;
    HN_TRY(symmetry_stamp(g, &stamp));
    insert_alloc_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(
        ucnt, b.seg_start, b.seg_key, ubk, f, g->rec, g->alloc_view(), b.ti, stamp, b.sc);
    HN_CHECK_LAUNCH();
    alloc_fix_kernel<<<1, 1, 0, st>>>(g->alloc, b.sc);
    HN_CHECK_LAUNCH();
    // from here to the commit, a failure hands the new blocks back
    auto undo = [&](hn_status e) {
        release_new_blocks_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(
            ucnt, b.ti, nullptr, g->alloc_view(), b.sc);
        (void)hipStreamSynchronize(st);
        return e;
    };
#define HN_UNDO_TRY(expr)                  \
    do {                                   \
        hn_status s_ = (expr);             \
        if (s_ != HN_OK) return undo(s_);  \
    } while (0)
#define HN_UNDO_LAUNCH() HN_UNDO_TRY(hipGetLastError() == hipSuccess ? HN_OK : HN_E_HIP)
    // 4. the old entries that move, in place in one pass (the few that could
    // overwrite a slot a later copy tile has not read go through a side
    // buffer), then the batch entries into the slots left between them
    HN_UNDO_TRY(exclusive_scan_fn<int64_t>(MoveLenIn{b.ti, ucnt, b.sc}, RunOut{b.wpos, b.tile_lo}, b.wpos + n, n,
                                           b.tmp, st));
    const int64_t nt = b.ntile_max;
    insert_defer_kernel<<<(unsigned)ceil_div(nt, BLOCK), BLOCK, 0, st>>>(nt, n, b.wpos, b.tile_lo, b.ti, b.seg_start,
                                                                        b.kb, f, g->arena, b.tcnt, b.sc);
    HN_UNDO_LAUNCH();
    HN_UNDO_TRY(exclusive_scan_fn<int64_t>(TileCountIn{b.tcnt}, SideOut{b.doff, nt - 1, b.sc}, b.doff + nt, nt,
                                           b.tmp, st));
    int64_t ndefer = 0;
    int32_t* side = nullptr;
    int64_t* side_dst = nullptr;
    HN_UNDO_TRY(reserve_side(g, b, st, &ndefer, &side, &side_dst));
    const int64_t grid = (int64_t)num_cus() * COPY_GRID_PER_CU;
    insert_move_kernel<<<(unsigned)grid, BLOCK, 0, st>>>(n, b.wpos, b.tile_lo, b.ti, b.seg_start, b.kb, f, b.tcnt,
                                                         b.doff, g->arena, side, side_dst, b.sc, g->meta, b.sm);
    HN_UNDO_LAUNCH();
    if (ndefer > 0) {
        place_side_kernel<<<(unsigned)ceil_div(ndefer, BLOCK), BLOCK, 0, st>>>(ndefer, side, side_dst, g->arena,
                                                                                 g->meta, b.sm);
        HN_UNDO_LAUNCH();
    }
    insert_batch_kernel<<<gridn, BLOCK, 0, st>>>(nb, b.kb, b.spos, b.seg_start, ubk, b.ti, f, g->arena, b.sc, g->meta,
                                                 bm.m, b.ib);
    HN_UNDO_LAUNCH();
#undef HN_UNDO_LAUNCH
#undef HN_UNDO_TRY
    // 5. publish
    commit_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(ucnt, b.seg_key, f, b.ti, nullptr, g->rec, g->alloc_view(), b.sc);
    HN_CHECK_LAUNCH();
    store_changed_kernel<<<1, 1, 0, st>>>(nb, b.sc);
    HN_CHECK_LAUNCH();
    const bool was_sym = g->sym == 1;
    if (was_sym) {
        insert_symmetry_kernel<<<gridn, BLOCK, 0, st>>>(nb, b.kb, f, stamp, g->rec, g->arena, b.sc);
        HN_CHECK_LAUNCH();
    }
    UpdScal h;
    HN_TRY(check_scal(g, b, st, &h));
    g->ne += (int64_t)h.changed;
    g->version++;
    note_symmetry(g, was_sym, h);
    record_last(g, 1, n, (int64_t)h.changed, b.ti, ucnt, nullptr, b.wpos + n);
    if (inserted) *inserted = (int64_t)h.changed;
    return HN_OK;
}

static hn_status erase_once(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n, uint32_t flags,
                            int64_t* erased, void* stream) {
    HN_ARG(g, "null graph");
    HN_ARG(n >= 0, "negative batch size");
    if (erased) *erased = 0;
    if (n == 0) return HN_OK;
    HN_ARG(src && dst, "null batch");
    HN_ARG(n < (int64_t)0x7FFFFFFF, "batch larger than 2^31-1 entries");
    hipStream_t st = as_stream(stream);
    UpdBufs b;
    HN_TRY(upd_buffers(g, n, &b));
    KeyFmt f{key_bits_for(g->nv)};
    uint64_t* keys = nullptr;
    HN_TRY(load_and_sort(g, src, dst, n, flags, b, f, st, &keys));
    const unsigned gridn = (unsigned)ceil_div(n + 1, BLOCK);
    // 1. optional in-batch dedup -> kb
    HN_TRY(exclusive_scan<uint32_t>(FirstKeepIn{keys, (flags & HN_DEDUP_BATCH) ? 1 : 0}, b.pos, n, b.tmp, st));
    compact_keys_kernel<<<gridn, BLOCK, 0, st>>>(n, keys, b.pos, b.kb, nullptr, nullptr, nullptr, nullptr);
    HN_CHECK_LAUNCH();
    uint32_t* nb = b.pos + n;
    // 2. distinct keys with multiplicities
    HN_TRY(exclusive_scan<uint32_t>(SegStartIn{b.kb, nb, 64}, b.spos, n, b.tmp, st));
    seg_write_kernel<<<gridn, BLOCK, 0, st>>>(n, b.kb, nb, b.spos, b.seg_start, b.seg_key);
    HN_CHECK_LAUNCH();
    uint32_t* dcnt = b.spos + n;
    // 3. stored ranges and deletion counts per distinct key
    erase_locate_kernel<<<gridn, BLOCK, 0, st>>>(dcnt, b.seg_start, b.seg_key, f, g->rec, g->arena, b.dpos, b.dlen,
                                                 b.live, b.sc);
    HN_CHECK_LAUNCH();
    HN_TRY(exclusive_scan<uint32_t>(DLenIn{b.dlen, dcnt}, b.dlp, n, b.tmp, st));
    // 4. group the distinct keys by source
    HN_TRY(exclusive_scan<uint32_t>(SegStartIn{b.seg_key, dcnt, f.sbits}, b.tpos, n, b.tmp, st));
    seg_write_kernel<<<gridn, BLOCK, 0, st>>>(n, b.seg_key, dcnt, b.tpos, b.seg2_start, b.seg2_key);
    HN_CHECK_LAUNCH();
    uint32_t* ucnt = b.tpos + n;
    // 5. in place, smaller block, or no block
    uint32_t* stamp = nullptr;
    HN_TRY(symmetry_stamp(g, &stamp));
    erase_alloc_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(
        ucnt, b.seg2_start, b.seg2_key, f, b.dpos, b.dlen, b.dlp, g->rec, g->alloc_view(), b.ti, b.skip, stamp, b.sc);
    HN_CHECK_LAUNCH();
    alloc_fix_kernel<<<1, 1, 0, st>>>(g->alloc, b.sc);
    HN_CHECK_LAUNCH();
    // from here to the commit, a failure hands the new blocks back
    auto undo = [&](hn_status e) {
        release_new_blocks_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(
            ucnt, b.ti, b.skip, g->alloc_view(), b.sc);
        (void)hipStreamSynchronize(st);
        return e;
    };
#define HN_UNDO_TRY(expr)                  \
    do {                                   \
        hn_status s_ = (expr);             \
        if (s_ != HN_OK) return undo(s_);  \
    } while (0)
#define HN_UNDO_LAUNCH() HN_UNDO_TRY(hipGetLastError() == hipSuccess ? HN_OK : HN_E_HIP)
    // 6. the survivors as runs between deleted ranges, grouped by vertex
    erase_gaps_kernel<<<gridn, BLOCK, 0, st>>>(dcnt, ucnt, b.seg2_start, b.tpos, b.dpos, b.dlen, b.dlp, b.ti, b.skip,
                                               b.glen, b.gsrc, b.gdst, b.rtail, b.sc);
    HN_UNDO_LAUNCH();
    HN_UNDO_TRY(exclusive_scan_fn<int64_t>(GapLenIn{b.glen, dcnt, ucnt}, RunOutT<ERASE_TILE>{b.gP, b.tile_lo},
                                           b.gP + 2 * n, 2 * n, b.tmp, st));
    // 7. in-place compaction in one pass; the few entries that would overwrite
    // a source of an earlier tile go through a side buffer
    const int64_t nt = b.ntile_max;
    erase_defer_kernel<<<(unsigned)ceil_div(nt, BLOCK), BLOCK, 0, st>>>(nt, 2 * n, b.gP, b.tile_lo, b.gsrc, b.gdst,
                                                                       b.rtail, b.tcnt, b.sc);
    HN_UNDO_LAUNCH();
    HN_UNDO_TRY(exclusive_scan_fn<int64_t>(TileCountIn{b.tcnt}, SideOut{b.doff, nt - 1, b.sc}, b.doff + nt, nt,
                                           b.tmp, st));
    int64_t ndefer = 0;
    int32_t* side = nullptr;
    int64_t* side_dst = nullptr;
    HN_UNDO_TRY(reserve_side(g, b, st, &ndefer, &side, &side_dst));
    const int64_t grid = (int64_t)num_cus() * COPY_GRID_PER_CU;
    if (g->meta.n > 0)
        erase_copy_kernel<true><<<(unsigned)grid, BLOCK, 0, st>>>(2 * n, b.gP, b.tile_lo, b.gsrc, b.gdst, b.tcnt,
                                                                  b.doff, g->arena, side, side_dst, b.sc, g->meta,
                                                                  b.sm);
    else
        erase_copy_kernel<false><<<(unsigned)grid, BLOCK, 0, st>>>(2 * n, b.gP, b.tile_lo, b.gsrc, b.gdst, b.tcnt,
                                                                   b.doff, g->arena, side, side_dst, b.sc, g->meta,
                                                                   b.sm);
    HN_UNDO_LAUNCH();
    if (ndefer > 0) {
        place_side_kernel<<<(unsigned)ceil_div(ndefer, BLOCK), BLOCK, 0, st>>>(ndefer, side, side_dst, g->arena,
                                                                               g->meta, b.sm);
        HN_UNDO_LAUNCH();
    }
#undef HN_UNDO_LAUNCH
#undef HN_UNDO_TRY
    // 8. publish
    commit_kernel<<<(unsigned)ceil_div(n + 1, ALLOC_BLOCK), ALLOC_BLOCK, 0, st>>>(ucnt, b.seg2_key, f, b.ti, b.skip, g->rec, g->alloc_view(), b.sc);
    HN_CHECK_LAUNCH();
    store_changed_kernel<<<1, 1, 0, st>>>(b.dlp + n, b.sc);
    HN_CHECK_LAUNCH();
    const bool was_sym = g->sym == 1;
    if (was_sym) {
        erase_symmetry_kernel<<<gridn, BLOCK, 0, st>>>(dcnt, b.seg_key, f, b.dlen, b.live, stamp, b.sc);
        HN_CHECK_LAUNCH();
    }
    UpdScal h;
    HN_TRY(check_scal(g, b, st, &h));
    g->ne -= (int64_t)h.changed;
    g->version++;
    note_symmetry(g, was_sym, h);
    record_last(g, 2, n, (int64_t)h.changed, b.ti, ucnt, b.skip, b.gP + 2 * n);
    if (erased) *erased = (int64_t)h.changed;
    return HN_OK;
}

hn_status hn_graph_insert(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n, uint32_t flags,
                          int64_t* inserted, void* stream) {
    return hn_graph_insert_meta(g, src, dst, nullptr, n, flags, inserted, stream);
}

hn_status hn_graph_insert_meta(hn_graph* g, const int32_t* src, const int32_t* dst, const void* const* meta,
                               int64_t n, uint32_t flags, int64_t* inserted, void* stream) {
    return with_growth(g, stream, [&]() { return insert_once(g, src, dst, meta, n, flags, inserted, stream); });
}

hn_status hn_graph_erase(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n, uint32_t flags,
                         int64_t* erased, void* stream) {
    return with_growth(g, stream, [&]() { return erase_once(g, src, dst, n, flags, erased, stream); });
}

hn_status hn_graph_update_stats(hn_graph* g, hn_update_stats* out, void* stream) {
    HN_ARG(g && out, "null argument");
    *out = hn_update_stats{};
    out->kind = g->last_kind;
    if (!g->last_kind) return HN_OK;
    hipStream_t st = as_stream(stream);
    HN_TRY(g->s_misc.reserve(64));
    unsigned long long* acc = g->s_misc.as<unsigned long long>();
    HN_HIP(hipMemsetAsync(acc, 0, 32, st));
    update_stats_kernel<<<(unsigned)num_cus(), BLOCK, 0, st>>>(
        g->last_ucnt, static_cast<const TouchInfo*>(g->last_ti), g->last_skip, g->last_kind == 2 ? 1 : 0, acc);
    HN_CHECK_LAUNCH();
    unsigned long long h[4];
    int64_t moved = 0;
    HN_TRY(read_back(g, acc, sizeof(h), h, st));
    HN_TRY(read_back(g, g->last_moved, sizeof(moved), &moved, st));
    out->submitted = g->last_submitted;
    out->applied = g->last_applied;
    out->touched = (int64_t)h[0];
    out->relocated = (int64_t)h[1];
    out->touched_deg = (int64_t)h[2];
    out->relocated_deg = (int64_t)h[3];
    out->moved = moved;
    return HN_OK;
}

}  // extern "C"
