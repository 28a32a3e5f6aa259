/*
 * hornet_hip.h -- C ABI of libhornet_hip.so, the MI355X (gfx950) implementation
 * of Hornet's dynamic-graph hot path.
 *
 * The reference has no shared-library boundary: everything is header-only
 * C++/CUDA instantiated in the caller's translation unit
 * (hornetsnest/include/HornetAlg.hpp:44-53).  This header is the boundary that
 * the C++ facade (the headers under include/hornet/) and any FFI binding call; each entry
 * point names the reference interface it replaces.
 *
 * Conventions
 *   - Every call returns an hn_status; nothing here exits or throws (the
 *     reference prints and calls exit(EXIT_FAILURE) on any CUDA error,
 *     xlib/src/Device/Util/SafeCudaAPI.cpp:71-94).  hn_last_error() gives a
 *     thread-local message for the last failure.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Calls that return a host-visible scalar synchronise that stream; all
 *     others are asynchronous on it.
 *   - Input pointers are borrowed for the duration of the call.  Output
 *     buffers are caller-owned, with explicit capacity where sizes vary.
 *   - A graph handle is not internally locked: one host thread (or one
 *     external lock) per handle.  Different handles are independent.
 *   - Vertex ids and degrees are int32 (VID_T = DEGREE_T = int,
 *     hornet/include/Core/Conf/HornetConf.cuh:41-42); edge counts and offsets
 *     are int64.
 */
#ifndef HORNET_HIP_H
#define HORNET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hn_status;
enum {
    HN_OK = 0,
    HN_E_INVALID_ARG = 1, /* bad size, id out of range, null pointer        */
    HN_E_OOM = 2,         /* device memory or edge arena exhausted           */
    HN_E_HIP = 3,         /* a HIP runtime call failed                       */
    HN_E_NO_DEVICE = 4,   /* no usable gfx950 device                         */
    HN_E_CAPACITY = 5     /* caller's output buffer too small                */
};

/* flags */
#define HN_DEVICE_PTRS      0x1u  /* pointer arguments are device memory      */
#define HN_OFFSETS_INT32    0x2u  /* CSR offsets are int32 (HornetInit form)  */
#define HN_DEDUP_BATCH      0x4u  /* insert/erase: removeBatchDuplicates      */
#define HN_DEDUP_GRAPH      0x8u  /* insert: removeGraphDuplicates            */
#define HN_DROP_SELF_LOOPS  0x10u /* graph build: drop (v,v)                  */
#define HN_SYMMETRIZE       0x20u /* graph build: add (d,s) for every (s,d)   */

typedef struct hn_graph hn_graph;

/* Per-vertex record in HBM: 16 bytes, one 128-bit load.  Replaces the
 * reference's SoA {degree, edge_block_ptr, vertex_offset, edges_per_block}
 * (hornet/include/Core/HornetDevice/impl/Vertex.i.cuh:59-146).  The adjacency
 * of v is arena[off .. off+deg), sorted ascending; its block holds
 * 2^cap_log2 slots (cap_log2 = -1 for a vertex without a block). */
typedef struct hn_vrec {
    int64_t off;
    int32_t deg;
    int32_t cap_log2;
} hn_vrec;

/* Edge meta-data (TypeList<EdgeMetaTypes...>, hornet/include/Core/HornetInitialize/HornetInit.cuh:54-97):
 * up to HN_MAX_EDGE_META fields per edge, each 4 or 8 bytes, kept as arrays
 * parallel to the edge arena (field f of the edge in slot i is meta[f][i]). */
#define HN_MAX_EDGE_META 4
typedef struct hn_edge_meta {
    int32_t nfields;
    int32_t bytes[HN_MAX_EDGE_META];   /* 4 or 8 */
} hn_edge_meta;

/* POD device view passed by value into kernels (HornetDevice,
 * hornet/include/Core/HornetDevice/HornetDevice.cuh:52-90).  Valid until the
 * next mutating call on the graph. */
typedef struct hn_dview {
    int32_t nv;
    int32_t nmeta;                      /* edge meta-data fields (0 = none) */
    int64_t ne;
    hn_vrec* rec;
    int32_t* arena;
    void* meta[HN_MAX_EDGE_META];       /* field f of arena slot i: meta[f][i] */
    int32_t meta_bytes[HN_MAX_EDGE_META];
} hn_dview;

/* BFS statistics for one traversal. */
typedef struct hn_bfs_stats {
    int32_t levels;          /* number of non-empty frontiers                 */
    int32_t time_kernels;    /* in: nonzero -> time the expansion kernel      */
    int64_t visited;         /* reached vertices, source included             */
    int64_t edges_scanned;   /* sum of degrees of reached vertices (m_C)      */
    int64_t expand_ns;       /* out: summed expansion-kernel time (HIP events) */
    int32_t expand_launches; /* out: expansion-kernel launches                */
    int32_t pad_;
} hn_bfs_stats;

/* Multi-source BFS statistics for one launch. */
typedef struct hn_msbfs_stats {
    int32_t levels;       /* level steps run                                  */
    int32_t push_levels;  /* of which top-down (push)                         */
    int32_t pull_levels;  /* of which bottom-up (pull)                        */
    int32_t pad_;
    int64_t edges;        /* sum over sources of degrees of reached vertices  */
} hn_msbfs_stats;

/* ---------------------------------------------------------------- misc */
const char* hn_status_string(hn_status s);
const char* hn_last_error(void);
/* Library version string. */
const char* hn_version(void);
/* Device properties the kernels size themselves from. */
hn_status hn_device_info(int32_t* num_cus, int64_t* lds_per_cu, int64_t* hbm_bytes,
                         int32_t* wave_size);

/* ----------------------------------------------------------- storage */
/* Hornet(HornetInit&) -- hornet/include/Core/HornetInitialize/HornetInitialize.i.cuh:86-146
 * with HornetInit(nV, nE, offsets, edges) -- HornetInit.cuh:70-73.
 * offsets: nv+1 entries (int64, or int32 with HN_OFFSETS_INT32), host memory
 * unless HN_DEVICE_PTRS.  Each adjacency is stored sorted ascending.
 * `arena_slack` >= 1.0 sizes the edge arena as slack x the pow2 blocks the
 * input needs (0 = default 2.0). */
hn_status hn_graph_create_csr(int32_t nv, int64_t ne, const void* offsets,
                              const int32_t* edges, uint32_t flags, double arena_slack,
                              void* stream, hn_graph** out);
/* Hornet(HornetInit&) after HornetInit::insertEdgeData(meta...)
 * (HornetInit.cuh:75-97): as hn_graph_create_csr, with `layout->nfields`
 * meta-data arrays meta[f] of ne elements of layout->bytes[f] bytes, in the
 * memory space the flags give the edges (a NULL meta[f] reads as zeros); they
 * follow their edges through every update.  layout NULL or nfields 0 =
 * hn_graph_create_csr. */
hn_status hn_graph_create_csr_meta(int32_t nv, int64_t ne, const void* offsets, const int32_t* edges,
                                   const hn_edge_meta* layout, const void* const* meta, uint32_t flags,
                                   double arena_slack, void* stream, hn_graph** out);
/* The graph's edge meta-data layout (nfields 0 for a graph without). */
hn_status hn_graph_edge_meta(const hn_graph* g, hn_edge_meta* layout);
/* Graph from a device COO (src,dst) of m entries: optional symmetrize /
 * self-loop drop (GraphStd::COOtoCSR, xlib/src/Graph/GraphStd.cpp:173-341),
 * always sorted and deduplicated.  Device pointers only. */
hn_status hn_graph_create_coo(int32_t nv, int64_t m, const int32_t* src, const int32_t* dst,
                              uint32_t flags, double arena_slack, void* stream, hn_graph** out);
/* ~Hornet() */
hn_status hn_graph_destroy(hn_graph* g);
/* Hornet::nV() / nE() -- Hornet.cuh:113-117 */
hn_status hn_graph_nv(const hn_graph* g, int32_t* nv);
hn_status hn_graph_ne(const hn_graph* g, int64_t* ne);
/* Hornet::device() -- HornetInsert.i.cuh:4-10 */
hn_status hn_graph_device_view(const hn_graph* g, hn_dview* view);
/* Arena bookkeeping: slots in use by blocks, slots reserved, free-list blocks. */
hn_status hn_graph_memory(hn_graph* g, int64_t* used_slots, int64_t* capacity_slots,
                          void* stream);
/* Hornet::getCOO(sort=true) -- HornetQuery.i.cuh:111-151 (flattenCOOKernel
 * :46-65).  Writes ne (src,dst) pairs sorted by (src,dst) into device buffers
 * of capacity `cap`; *n_out = ne.  Synchronises. */
hn_status hn_graph_export_coo(hn_graph* g, int32_t* src, int32_t* dst, int64_t cap,
                              int64_t* n_out, void* stream);
/* getCOO with the edge meta-data (COO<DEVICE, vid_t, TypeList<EdgeMetaTypes...>>,
 * HornetQuery.i.cuh:111-151): as hn_graph_export_coo, plus meta[f] (device,
 * cap elements of the field's size) for each field; meta may be NULL. */
hn_status hn_graph_export_coo_meta(hn_graph* g, int32_t* src, int32_t* dst, void* const* meta, int64_t cap,
                                   int64_t* n_out, void* stream);
/* CSR export (the working replacement for the reference's broken getCSR,
 * hornet/include/Core/Static/Static.i.cuh:100-103): offsets[nv+1] int64,
 * adj[cap] int32, device buffers.  Synchronises. */
hn_status hn_graph_export_csr(hn_graph* g, int64_t* offsets, int32_t* adj, int64_t cap,
                              void* stream);
/* Hornet::getCSR(sortAdjacencyList) -- hornet/include/Core/HornetOperations/HornetQuery.i.cuh:69-107
 * (CSR<DEVICE, vid_t, TypeList<EdgeMetaTypes...>>, hornet/include/Core/Static/Static.cuh:92-129):
 * hn_graph_export_csr plus meta[f] (device, cap elements of field f's size)
 * for each edge meta-data field; meta may be NULL.  Rows are sorted, as the
 * storage keeps them.  Synchronises. */
hn_status hn_graph_export_csr_meta(hn_graph* g, int64_t* offsets, int32_t* adj, void* const* meta, int64_t cap,
                                   void* stream);
/* An independent device copy of g (no reference counterpart: the reference's
 * immutable graph, HornetStatic, hornet/include/Core/Static/HornetStatic.cuh:16-66,
 * is built from a HornetInit).  Updates to either graph do not reach the other,
 * so queries can read a fixed version while the original keeps changing.
 * `arena_slack` as in hn_graph_create_csr (values < 1.0 mean 1.0: no room for
 * growth).  Synchronises. */
hn_status hn_graph_snapshot(hn_graph* g, double arena_slack, void* stream, hn_graph** out);
/* Per-vertex degrees into a device int32[nv]. */
hn_status hn_graph_degrees(hn_graph* g, int32_t* deg, void* stream);
/* Hornet::max_degree_id / max_degree -- HornetQuery.i.cuh:10-36 (lowest id
 * on ties).  Synchronises. */
hn_status hn_graph_max_degree(hn_graph* g, int32_t* id, int32_t* deg, void* stream);

/* ------------------------------------------------------------ updates */
/* Hornet::insert(batch, removeBatchDuplicates, removeGraphDuplicates) --
 * hornet/include/Core/HornetOperations/HornetInsert.i.cuh:12-27 with
 * BatchUpdate::preprocess, BatchUpdate.i.cuh:485-502.  Flags: HN_DEDUP_BATCH,
 * HN_DEDUP_GRAPH, HN_DEVICE_PTRS.  Adjacency stays sorted.  *inserted (may
 * be NULL) receives the number of entries added.  Synchronises. */
hn_status hn_graph_insert(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n,
                          uint32_t flags, int64_t* inserted, void* stream);
/* Hornet::insert of a BatchUpdate carrying edge meta-data (BatchUpdate.cuh:95-99,
 * sorted with the batch as in SoAPtr.i.cuh:581-613): meta[f] holds n elements
 * of the graph's field f, in the memory space of src / dst.  Equal (src, dst)
 * entries keep their batch order, so HN_DEDUP_BATCH keeps the first one's
 * meta-data (BatchUpdate.i.cuh:250-335); HN_DEDUP_GRAPH drops the first batch
 * copy of a stored pair, whose stored meta-data stay.  Within a row, copies
 * of one destination are kept in insertion order.  meta NULL inserts zeros.  hn_graph_insert on a graph with
 * meta-data inserts zeros too. */
hn_status hn_graph_insert_meta(hn_graph* g, const int32_t* src, const int32_t* dst, const void* const* meta,
                               int64_t n, uint32_t flags, int64_t* inserted, void* stream);
/* Hornet::erase(batch, removeBatchDuplicates) -- HornetInsert.i.cuh:71-89
 * with BatchUpdate::preprocess_erase, BatchUpdate.i.cuh:504-524: multiset
 * difference, absent edges ignored.  Adjacency order is preserved; of c stored
 * copies of one (u, v) an erase of k removes the min(k, c) oldest (this
 * decides which meta-data go).
 * *erased (may be NULL) receives the number of entries removed.
 * Synchronises. */
hn_status hn_graph_erase(hn_graph* g, const int32_t* src, const int32_t* dst, int64_t n,
                         uint32_t flags, int64_t* erased, void* stream);

/* What the last insert / erase on g did (no reference counterpart; the
 * measurement hook for the batch-update roofline).  Valid until the next
 * update of g.  Runs one small reduction; synchronises. */
typedef struct hn_update_stats {
    int64_t submitted;      /* entries in the batch                                   */
    int64_t applied;        /* b': entries inserted / erased                          */
    int64_t touched;        /* |T|: distinct sources of the (kept) batch entries      */
    int64_t relocated;      /* |R|: touched vertices moved to a new block             */
    int64_t touched_deg;    /* sum over T of the degree before the batch              */
    int64_t relocated_deg;  /* sum over R of the degree before (insert) / after (erase) */
    int64_t moved;          /* stored entries rewritten in place or copied            */
    int32_t kind;           /* 0 none yet, 1 insert, 2 erase                          */
    int32_t pad_;
} hn_update_stats;
hn_status hn_graph_update_stats(hn_graph* g, hn_update_stats* out, void* stream);

/* --------------------------------------------------------- algorithms */
/* BfsTopDown{set_parameters(source); run()} --
 * hornetsnest/src/Static/BreadthFirstSearch/TopDown.cu:88-106.
 * dist: device int32[nv]; source gets 0, unreached INT32_MAX.  Level loop on
 * the host with one 4-byte readback per level.  stats may be NULL. */
hn_status hn_bfs_topdown(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats,
                         void* stream);
/* hn_bfs_topdown without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * traversal scratch is one static device array per GPU of
 * hn_bfs_topdown_noalloc_limit() bytes (64 MiB: about 13 bytes per vertex,
 * R-MAT up to about 5 M vertices), held until the process ends; a graph needing more gets
 * HN_E_CAPACITY.  All calls on one GPU share that array, so calls that overlap
 * on one GPU (any streams) race.  Same results as hn_bfs_topdown; no kernel
 * timing (stats->time_kernels is ignored). */
hn_status hn_bfs_topdown_noalloc(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats,
                                 void* stream);
hn_status hn_bfs_topdown_noalloc_limit(int64_t* scratch_bytes);
/* Batched BFS: k <= 64 sources (host array) in one traversal with 64-bit
 * per-vertex source masks; no reference counterpart (the reference runs one
 * BfsTopDown per source, TopDown.cu:94-106).  dist: device int32[k * nv],
 * source-major, same convention as hn_bfs_topdown.  The level step is a
 * captured hipGraph replayed until the frontier empties.  stats may be NULL. */
hn_status hn_bfs_multi(hn_graph* g, const int32_t* sources, int32_t k, int32_t* dist,
                       hn_msbfs_stats* stats, void* stream);
/* Same with a per-source depth limit (0 = none, else <= 63): vertices farther
 * than the limit stay INT32_MAX (depth 2 gives the 2-hop neighbourhood). */
hn_status hn_bfs_multi_depth(hn_graph* g, const int32_t* sources, const int32_t* depth_limit,
                             int32_t k, int32_t* dist, hn_msbfs_stats* stats, void* stream);
/* Load-balancing parameters of the batched BFS: pull-side degree threshold
 * (open vertices of degree > T take one wave each and stop early; the others
 * go through the edge-granular pull; default 2^30 = none), and the push/pull
 * switch (push while frontier degree * alpha < unvisited degree).  0 keeps a
 * value. */
hn_status hn_bfs_multi_params(hn_graph* g, int32_t degree_threshold, int32_t push_alpha);
/* Load-balancing tile of the batched BFS expansion: work units (16-byte
 * adjacency groups) per lane per tile, 2, 4 (default) or 6. */
hn_status hn_bfs_multi_lb_tile(hn_graph* g, int32_t units_per_lane);
/* The batched BFS parameters in use (NULL outputs are skipped). */
hn_status hn_bfs_multi_get_params(hn_graph* g, int32_t* degree_threshold, int32_t* push_alpha,
                                  int32_t* units_per_lane);
/* Autotuner: times the batched BFS of `sources` over a grid of thresholds and
 * switch points, then over the load-balancing tile sizes at the best pair,
 * and keeps the fastest for this graph (hn_bfs_multi_get_params reads the
 * tile chosen).  dist is scratch of k * nv int32.  Synchronises. */
hn_status hn_bfs_multi_tune(hn_graph* g, const int32_t* sources, int32_t k, int32_t* dist,
                            int32_t* best_threshold, int32_t* best_alpha, void* stream);

/* TriangleCounting2{init(); run(); countTriangles()} on the undirected simple
 * graph the symmetric storage represents --
 * hornetsnest/src/Static/TriangleCounting/triangle2.cu:130-160 applied to the
 * DIRECTED_BY_DEGREE orientation (hornetsnest/test/TriangleTest2.cu:86-87,
 * xlib/src/Graph/GraphStd.cpp:247-304).  per_vertex (device uint64[nv], may be
 * NULL) receives the count credited to the lower-ranked endpoint of each
 * oriented edge.  Synchronises. */
hn_status hn_tc_count(hn_graph* g, uint64_t* total, uint64_t* per_vertex, void* stream);
/* TriangleCounting2::run() on the graph exactly as stored (no orientation):
 * for every stored (s,d) with s != d and deg(s) >= 2, |N(s) ∩ N(d)| credited
 * to the smaller-degree endpoint (primitives/Operator++.i.cuh:97-246,
 * triangle2.cu:75-124).  On a graph built DIRECTED_BY_DEGREE this is the
 * triangle count.  per_vertex may be NULL.  Synchronises. */
hn_status hn_tc2_stored(hn_graph* g, uint64_t* total, uint64_t* per_vertex, void* stream);

/* ------------------------------------------------------------- serving */
/* A request loop that keeps the graph resident and serves a stream of BFS
 * queries, 2-hop neighbourhood-size queries and update batches (no reference
 * counterpart: the reference has no serving mode).  Requests are taken in
 * arrival order; consecutive queries (no update between them) are batched, up
 * to 64, into one multi-source launch (2-hop = depth limit 2), and up to
 * `slots` launches run at once on their own streams; an update waits for the
 * launches before it and blocks the queries after it, so every query sees
 * exactly the updates that arrived before it (snapshot-consistent order).
 * The number of slots admitted is bounded by the device-memory budget. */
typedef struct hn_server hn_server;
enum { HN_REQ_BFS = 0, HN_REQ_TWO_HOP = 1, HN_REQ_INSERT = 2, HN_REQ_ERASE = 3 };

typedef struct hn_request {
    int32_t kind;          /* HN_REQ_*                                         */
    int32_t source;        /* queries                                          */
    int64_t n;             /* updates: entries                                 */
    const int32_t* src;    /* updates: host arrays, copied at submission       */
    const int32_t* dst;
} hn_request;

typedef struct hn_result {
    int64_t id;
    int32_t kind;
    int32_t status;        /* hn_status of the request                         */
    int64_t t_arrive_ns;   /* steady-clock times                               */
    int64_t t_start_ns;
    int64_t t_done_ns;
    int64_t reached;       /* BFS: reached vertices; 2-hop: |{v} u N(v) u N(N(v))| */
    int64_t dist_sum;      /* BFS: sum of the distances of reached vertices    */
    int32_t max_depth;     /* BFS: largest finite distance                     */
    int32_t batch;         /* queries that shared the launch                   */
    int64_t applied;       /* updates: entries inserted / erased               */
    uint64_t version;      /* graph version the request saw (update: produced) */
} hn_result;

typedef struct hn_serve_config {
    int32_t max_batch;        /* queries per launch, 1..64                     */
    int32_t slots;            /* concurrent launches wanted                    */
    int64_t mem_budget_bytes; /* device memory for launches (0: 1/4 of free)   */
    int32_t plain;            /* 1: one query per launch, single-source BFS    */
    int32_t min_window_queries; /* queries a multi-version window must batch
                                   across updates; fewer -> the updates apply
                                   one by one (0: 8; > 64: never a window)   */
} hn_serve_config;

hn_status hn_server_create(hn_graph* g, const hn_serve_config* cfg, hn_server** out);
/* Thread-safe; *id (may be NULL) receives the request id (0, 1, ...). */
hn_status hn_server_submit(hn_server* s, const hn_request* req, int64_t* id);
/* Submit reqs[i] at arrival_ns[i] after the call starts (a client thread),
 * then wait until every request is done. */
hn_status hn_server_replay(hn_server* s, const hn_request* reqs, const int64_t* arrival_ns, int64_t n);
hn_status hn_server_drain(hn_server* s);
/* Results of all requests so far, indexed by id; *n = number of requests. */
hn_status hn_server_results(hn_server* s, hn_result* out, int64_t cap, int64_t* n);
/* Slots admitted after the memory budget. */
hn_status hn_server_slots(hn_server* s, int32_t* slots);
/* Device bytes one launch slot of a batching server on g takes from the
 * memory budget (the multi-source engine; batched launches keep per-source
 * digests only, so max_batch does not change it). */
hn_status hn_server_slot_bytes(const hn_graph* g, int32_t max_batch, int64_t* bytes);
hn_status hn_server_destroy(hn_server* s);

/* ---------------------------------------------------------- generators */
/* Seeded R-MAT(a=.57,b=.19,c=.19,d=.05) pairs on the device: pair i, level l
 * uses draw i*64+l of SplitMix64 stream `seed` (specified by this library; the
 * reference has no generator).  Writes pairs [first, first+count). */
hn_status hn_gen_rmat_pairs(int32_t scale, int64_t first, int64_t count, uint64_t seed,
                            int32_t* src, int32_t* dst, void* stream);
/* R-MAT graph: edge_factor * 2^scale pairs, self loops dropped, symmetrized,
 * deduplicated, sorted adjacency. */
hn_status hn_gen_rmat_graph(int32_t scale, int32_t edge_factor, uint64_t seed,
                            double arena_slack, void* stream, hn_graph** out);
/* Chung-Lu pairs over [0,n): endpoints drawn ∝ w_i = (i+1)^(-1/(gamma-1))
 * from a cumulative weight table (floor(2^24 w_i) per vertex, built on the
 * host); pair i uses draws 2i and 2i+1 of stream `seed`. */
hn_status hn_gen_chung_lu_pairs(int32_t n, double gamma, int64_t first, int64_t count, uint64_t seed,
                                int32_t* src, int32_t* dst, void* stream);
/* Chung-Lu graph: n*avg_degree/2 pairs, self loops dropped, symmetrized,
 * deduplicated, sorted adjacency. */
hn_status hn_gen_chung_lu_graph(int32_t n, int32_t avg_degree, double gamma, uint64_t seed,
                                double arena_slack, void* stream, hn_graph** out);
/* Uniform update pairs over [0,n) with ~1 % in-batch repeats (stream `seed`). */
hn_status hn_gen_uniform_pairs(int32_t n, int64_t count, uint64_t seed, int32_t* src,
                               int32_t* dst, void* stream);

/* ---------------------------------------------------------- graph files */
/* graph::GraphStd<int, int64> -- xlib/include/Graph/GraphStd.hpp: a host CSR
 * read from a file, handed to hn_graph_create_csr the way the reference
 * drivers hand graph.csr_out_offsets() / csr_out_edges() to HornetInit
 * (hornetsnest/test/BFSTest.cu:18-23).  Host memory only; no device call. */
typedef struct hn_gstd hn_gstd;
/* StructureProp -- xlib/include/Graph/GraphBase.hpp:92-93,115-123 */
#define HN_GS_DIRECTED        0x1u
#define HN_GS_UNDIRECTED      0x2u
#define HN_GS_ENABLE_INGOING  0x4u
/* ParsingProp -- GraphBase.hpp:46-48 (PRINT_INFO is accepted and prints nothing) */
#define HN_PARSE_RANDOMIZE           0x1u
#define HN_PARSE_SORT                0x2u
#define HN_PARSE_PRINT_INFO          0x4u
#define HN_PARSE_RM_SINGLETON        0x8u
#define HN_PARSE_DIRECTED_BY_DEGREE  0x10u
/* GraphStd(structure) + read(filename, parsing) -- xlib/src/Graph/GraphBase.cpp:130-206
 * (dispatch on .mtx "%%MatrixMarket" / .gr "c"|"p" / .bin), readers
 * GraphStdRead.cpp:45-105,206-236, COOtoCSR GraphStd.cpp:173-351.
 * structure 0 takes the file's direction (MatrixMarket "symmetric" and DIMACS9:
 * UNDIRECTED, else DIRECTED); a directed file read UNDIRECTED is mirrored and
 * deduplicated.  RANDOMIZE relabels by a Fisher-Yates permutation drawn from
 * SplitMix64 stream `seed` (the reference seeds from the clock).  Errors
 * (HN_E_INVALID_ARG, message in hn_last_error): unreadable or unrecognised
 * file, ids outside [1, n], entry count different from the header, malformed
 * binary, an undirected file read into a DIRECTED structure. */
hn_status hn_gstd_read(const char* path, uint32_t structure, uint32_t parse, uint64_t seed,
                       hn_gstd** out);
/* GraphStd(csr_offsets, nV, csr_edges, nE) -- xlib/src/Graph/GraphStd.cpp:50-59:
 * a host graph (structure UNDIRECTED) copied from a host CSR, e.g. a graph's
 * hn_graph_export_csr, to be written with hn_gstd_write_binary. */
hn_status hn_gstd_from_csr(int32_t nv, int64_t ne, const int64_t* offsets, const int32_t* edges,
                           hn_gstd** out);
/* ~GraphStd() */
hn_status hn_gstd_destroy(hn_gstd* g);
/* nV(), nE() -- GraphBase.hpp:137-138,192-199 -- and the resolved StructureProp bits */
hn_status hn_gstd_info(const hn_gstd* g, int32_t* nv, int64_t* ne, uint32_t* structure);
/* csr_out_offsets() / csr_out_edges() -- GraphStd.i.hpp:239-251: nv+1 int64
 * offsets, ne int32 edges, each row in file order (sorted with
 * HN_PARSE_SORT); valid until hn_gstd_destroy. */
hn_status hn_gstd_csr_out(const hn_gstd* g, const int64_t** offsets, const int32_t** edges);
/* csr_in_offsets() / csr_in_edges() -- GraphStd.i.hpp:244-256: DIRECTED graphs
 * read with ENABLE_INGOING; an UNDIRECTED graph returns its out-CSR. */
hn_status hn_gstd_csr_in(const hn_gstd* g, const int64_t** offsets, const int32_t** edges);
/* max_out_degree_id() -- GraphStd.i.hpp:296-299 (first vertex of largest degree) */
hn_status hn_gstd_max_out_degree_id(const hn_gstd* g, int32_t* id);
/* writeBinary(filename) -- GraphStd.cpp:387-421: class id "intint"
 * (offset_bytes 4, GraphStd<int, int>) or "intlong" (8, GraphStd<int, int64_t>),
 * nV, nE, the 16-byte StructureProp (zero vtable word, state DIRECTED or
 * DIRECTED|ENABLE_INGOING), offsets, edges. */
hn_status hn_gstd_write_binary(const hn_gstd* g, const char* path, int32_t offset_bytes);
/* writeMarket(filename) -- GraphStd.cpp:425-437 */
hn_status hn_gstd_write_market(const hn_gstd* g, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* HORNET_HIP_H */
