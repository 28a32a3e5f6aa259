// graph.hpp -- internal layout of an hn_graph (vertex records + edge arena +
// device-side block allocator).
//
// Reference: hornet::gpu::Hornet (hornet/include/Core/Hornet.cuh:54-131) keeps
// a vertex SoA {degree, edge_block_ptr, vertex_offset, edges_per_block} and a
// host-side BlockArrayManager (hornet/include/Core/MemoryManager/BlockArray/
// BlockArray.i.cuh:111-173) that finds free power-of-two blocks with a host
// bit tree and one cudaMalloc per 2^23-slot BlockArray.
//
// Here the edge storage is ONE device arena of int32 slots, carved into
// power-of-two blocks of at least 32 slots (128 B, so every block start is
// 128-byte aligned).  Block offsets are kept in 128-byte units (uint32), so an
// arena may span up to 2^32 * 128 B = 512 GiB -- more than the 288 GB of an
// MI355X.  Allocation and release happen on the device:
//   * one free stack (LIFO of unit offsets) per size class k (block = 2^k slots);
//   * a bump pointer for fresh blocks.
// A batch allocates in one kernel (every requester takes a ticket per class;
// tickets below the stack depth pop, the rest bump), then a fixup kernel
// commits the stack depths, and frees are pushed after the data has moved.
// No host round trip per vertex.
#pragma once
#include <atomic>
#include <mutex>

#include "common.hpp"

namespace hn {

constexpr int MIN_CLASS = 5;          // 32 slots = 128 bytes
constexpr int NCLASS = 32;
constexpr int64_t UNIT_SLOTS = 32;    // slots per 128-byte unit

struct AllocState {
    unsigned long long bump_units;    // next never-used unit
    unsigned long long cap_units;     // arena capacity in units
    unsigned long long saved_bump;    // bump at the start of the current batch
    unsigned int free_count[NCLASS];  // stack depth per class
    unsigned int take[NCLASS];        // tickets handed out in the current batch
    unsigned int error;               // bit 0: arena exhausted
    unsigned int pad;
    unsigned long long want_units;    // fresh units the last batch that exhausted the arena asked for
};

struct AllocView {
    AllocState* st;
    uint32_t* stacks;                 // all class stacks, concatenated
    const int64_t* stack_off;         // start of class k's stack (entries)
};

__device__ __forceinline__ int class_of(int32_t deg) {
    if (deg <= 0) return -1;
    int k = 32 - __clz(deg - 1);      // ceil(log2(deg)) for deg >= 1 (deg=1 -> 0)
    return k < MIN_CLASS ? MIN_CLASS : k;
}

// Block allocation for kernels where many lanes allocate at once: tickets
// and fresh units are first counted with LDS atomics, then each workgroup
// takes one global atomic per size class it uses
// and one on the bump pointer.  Every thread of the workgroup must call it
// (it synchronises the workgroup); lanes with want == false pass through.
// Class k's tickets below its stack depth pop blocks, the rest bump fresh
// units; exhaustion returns -1 and sets the error bit.
struct AllocShared {
    unsigned int cnt[NCLASS];
    unsigned int base[NCLASS];
    unsigned long long units;
    unsigned long long b0;
};

__device__ __forceinline__ int64_t alloc_block_group(const AllocView& A, int k, bool want, AllocShared& sh) {
    for (int i = threadIdx.x; i < NCLASS; i += blockDim.x) sh.cnt[i] = 0;
    if (threadIdx.x == 0) sh.units = 0;
    __syncthreads();
    unsigned int lt = 0;
    if (want) lt = atomicAdd(&sh.cnt[k], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < NCLASS; i += blockDim.x)
        sh.base[i] = sh.cnt[i] ? atomicAdd(&A.st->take[i], sh.cnt[i]) : 0u;
    __syncthreads();
    int64_t res = -2;
    bool bump = false;
    unsigned long long units = 0, uoff = 0;
    if (want) {
        const unsigned ticket = sh.base[k] + lt;
        const unsigned fc = A.st->free_count[k];
        if (ticket < fc) {
            res = (int64_t)A.stacks[A.stack_off[k] + (fc - 1 - ticket)] * UNIT_SLOTS;
        } else {
            bump = true;
            units = 1ull << (k - MIN_CLASS);
            uoff = atomicAdd(&sh.units, units);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) sh.b0 = sh.units ? atomicAdd(&A.st->bump_units, sh.units) : 0ull;
    __syncthreads();
    if (bump) {
        const unsigned long long u = sh.b0 + uoff;
        if (u + units > A.st->cap_units) {
            atomicOr(&A.st->error, 1u);
            res = -1;
        } else {
            res = (int64_t)u * UNIT_SLOTS;
        }
    }
    return res;
}

// Push freed blocks onto their class stacks with one atomicAdd per class for
// the whole workgroup (counts in LDS), as alloc_block_group takes them.  Every
// thread of the workgroup calls it (threads that free nothing pass want =
// false).
__device__ __forceinline__ void free_block_group(const AllocView& A, int k, int64_t slot_off, bool want,
                                                 AllocShared& sh) {
    for (int i = threadIdx.x; i < NCLASS; i += blockDim.x) sh.cnt[i] = 0;
    __syncthreads();
    unsigned int lt = 0;
    if (want) lt = atomicAdd(&sh.cnt[k], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < NCLASS; i += blockDim.x)
        sh.base[i] = sh.cnt[i] ? atomicAdd(&A.st->free_count[i], sh.cnt[i]) : 0u;
    __syncthreads();
    if (want) A.stacks[A.stack_off[k] + sh.base[k] + lt] = (uint32_t)(slot_off / UNIT_SLOTS);
}

}  // namespace hn

namespace hn {
// Edge meta-data arrays (parallel to the arena, a staging area, a batch or an
// export): field f of entry i is p[f][i], bytes[f] wide.  p[f] == nullptr
// reads as zero.
struct MetaSet {
    int32_t n = 0;
    int32_t bytes[HN_MAX_EDGE_META] = {};
    char* p[HN_MAX_EDGE_META] = {};
    __host__ __device__ bool any() const { return n > 0; }
};
// the same layout over other arrays
inline MetaSet meta_like(const MetaSet& m, char* const* ptrs) {
    MetaSet r = m;
    for (int f = 0; f < HN_MAX_EDGE_META; f++) r.p[f] = f < m.n && ptrs ? ptrs[f] : nullptr;
    return r;
}
// d[di] = s[si] for every field
__device__ __forceinline__ void meta_put(const MetaSet& d, int64_t di, const MetaSet& s, int64_t si) {
    for (int f = 0; f < d.n; f++) {
        if (!d.p[f]) continue;
        if (d.bytes[f] == 8) {
            const uint64_t v = s.p[f] ? reinterpret_cast<const uint64_t*>(s.p[f])[si] : 0ull;
            reinterpret_cast<uint64_t*>(d.p[f])[di] = v;
        } else {
            const uint32_t v = s.p[f] ? reinterpret_cast<const uint32_t*>(s.p[f])[si] : 0u;
            reinterpret_cast<uint32_t*>(d.p[f])[di] = v;
        }
    }
}
}  // namespace hn

struct hn_graph {
    int32_t nv = 0;
    int64_t ne = 0;
    uint64_t version = 0;             // bumped by every applied insert / erase
    // adjacency symmetry ((u,w) stored iff (w,u) stored): 1 known symmetric,
    // -1 known not (for sym_version), 0 unknown
    int32_t sym = 0;
    uint64_t sym_version = 0;
    std::mutex sym_mu;                // serialises the symmetry check (serving slots call it concurrently)
    hn_vrec* rec = nullptr;           // device, nv records
    int32_t* arena = nullptr;         // device, cap_units * 32 slots
    hn::MetaSet meta;                 // device, edge meta-data parallel to the arena
    int64_t cap_units = 0;
    hn::AllocState* alloc = nullptr;  // device
    uint32_t* stacks = nullptr;       // device
    int64_t* stack_off = nullptr;     // device, NCLASS entries
    int64_t h_stack_off[hn::NCLASS] = {};
    // scratch owned by the handle (grown on demand)
    hn::Scratch s_bfs;                // BFS queues, bitmap, work prefix
    hn::Scratch s_scan;               // scan partials
    hn::Scratch s_upd;                // batch update buffers
    hn::Scratch s_tc;                 // triangle counting
    hn::Scratch s_misc;
    hn::Scratch s_mstage;             // staging of edge meta-data during updates
    hn::Scratch s_stamp;              // per vertex: its source segment in the current batch (unchecked)
    uint64_t* h_pinned = nullptr;     // small pinned readback area
    // levels + 1 of the last single-source BFS: when its level loop first
    // polls for the end (serving slots run BFS concurrently, hence atomic)
    std::atomic<int> bfs_last_levels{0};
    // timing events of the BFS expansion (one pair per queued level), created
    // on the first timed traversal
    static constexpr int BFS_EVENTS = 64;
    hipEvent_t bfs_ev[BFS_EVENTS] = {};
    // the last applied batch, for hn_graph_update_stats (pointers into s_upd,
    // valid until the next batch)
    int32_t last_kind = 0;            // 0 none, 1 insert, 2 erase
    int64_t last_submitted = 0;
    int64_t last_applied = 0;
    const void* last_ti = nullptr;    // per touched vertex (update.hip TouchInfo)
    const uint32_t* last_ucnt = nullptr;
    const uint8_t* last_skip = nullptr;
    const int64_t* last_moved = nullptr;
    void* ms = nullptr;               // multi-source BFS engine (msbfs.hip)
    void (*ms_release)(void*) = nullptr;

    hn::AllocView alloc_view() const { return hn::AllocView{alloc, stacks, stack_off}; }
    hn_dview view() const {
        hn_dview v;
        v.nv = nv;
        v.nmeta = meta.n;
        v.ne = ne;
        v.rec = rec;
        v.arena = arena;
        for (int f = 0; f < HN_MAX_EDGE_META; f++) {
            v.meta[f] = f < meta.n ? meta.p[f] : nullptr;
            v.meta_bytes[f] = f < meta.n ? meta.bytes[f] : 0;
        }
        return v;
    }
};

namespace hn {
// build a graph from device CSR (offsets int64[nv+1], adj int32[ne]); adjacency
// must already be sorted per vertex.  Does not take ownership of the inputs.
hn_status graph_from_device_csr(int32_t nv, int64_t ne, const int64_t* d_off, const int32_t* d_adj,
                                double slack, hipStream_t st, hn_graph** out,
                                const MetaSet& d_meta = MetaSet());
// device CSR from device COO: sort + dedup (+ symmetrize / drop self loops);
// returns buffers allocated with dalloc (caller frees).
hn_status csr_from_device_coo(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                              uint32_t flags, hipStream_t st, int64_t** d_off, int32_t** d_adj,
                              int64_t* ne_out);
// exclusive prefix of the record degrees into d_eoff[nv+1] (int64)
hn_status degree_prefix(hn_graph* g, int64_t* d_eoff, hipStream_t st);
// Internal status (never returned through the C ABI): a batch found the edge
// arena full and changed nothing; the caller grows the arena and retries.
constexpr hn_status HN_E_ARENA_FULL = 100;
// Grow the edge arena (and the meta-data arrays and free stacks) so that at
// least `need_units` fresh units (and at least what the last batch that
// exhausted the arena asked for) fit beyond the bump pointer; capacity at
// least doubles.  Blocks keep their unit offsets, so the vertex records stay
// valid; the arena (and every pointer into it, such as a device view) moves.
// HN_E_OOM when device memory or the 2^32-unit offset range is exhausted
// (the graph is then unchanged).  Synchronises.
hn_status grow_arena(hn_graph* g, uint64_t need_units, hipStream_t st);
// synchronous readback of a few bytes via the handle's pinned buffer
hn_status read_back(hn_graph* g, const void* dptr, size_t bytes, void* host, hipStream_t st);
// batched BFS of k <= 64 sources (host array), optional per-source depth
// limits (0 = none); dist is device int32[k * nv], source-major
hn_status ms_run(hn_graph* g, const int32_t* sources, int32_t k, const int32_t* depth_limit, int32_t* dist,
                 hn_msbfs_stats* stats, hipStream_t caller);
// standalone engines (one per concurrent caller thread)
struct MsEngine;
hn_status ms_engine_create(hn_graph* g, MsEngine** out);
void ms_engine_destroy(MsEngine* e);
hn_status ms_run_engine(MsEngine* e, const int32_t* sources, int32_t k, const int32_t* depth_limit, int32_t* dist,
                        hn_msbfs_stats* stats, hipStream_t caller);
// whether the adjacency is symmetric (checked once per graph version when
// not known; the multi-source BFS pulls only over symmetric graphs)
hn_status graph_is_symmetric(hn_graph* g, hipStream_t st, bool* yes);
// Multi-version windows (one launch serving queries of several graph
// versions).  begin: before the window's updates are applied, with every
// update entry on the device (d_tt[i] = update index 1..63 << 1 | 1 for an insert);
// end: after them (failed = bit t for an update that changed nothing; sv[t] =
// sources of window version t), builds the side records and synchronises;
// clear: after the launch.
// (*symmetric = false, and no window opened, when an entry lacks its reverse
// in the same update.)
hn_status ms_window_begin(MsEngine* e, const int32_t* d_src, const int32_t* d_dst, const uint8_t* d_tt, int64_t n,
                          hipStream_t st, bool* symmetric);
hn_status ms_window_end(MsEngine* e, int32_t m, uint64_t failed, const uint64_t* sv, hipStream_t st);
hn_status ms_window_clear(MsEngine* e, hipStream_t st);
// ms_run_engine with dist == nullptr runs in digest mode: no distance matrix;
// the device array returned here then holds {reached, distance sum,
// eccentricity} per source (k x 3 uint64), ordered on the caller's stream.
const unsigned long long* ms_engine_digest(const MsEngine* e);
// device bytes one engine holds for a graph of nv vertices and ne edges
int64_t ms_engine_bytes(int64_t nv, int64_t ne);

// One per source file: resolves that file's code object (preload_code_objects).
void preload_bfs_code();
void preload_gen_code();
void preload_graph_code();
void preload_msbfs_code();
void preload_prims_code();
void preload_tc_code();
void preload_update_code();
}  // namespace hn
