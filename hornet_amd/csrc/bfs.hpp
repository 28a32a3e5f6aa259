// bfs.hpp -- the single-source BFS level loop over caller-placed scratch
// (bfs.hip), shared by the entry points that own their scratch differently.
#pragma once
#include "graph.hpp"

namespace hn {

// Level-loop control on the device: the loop runs without a host round trip
// per level, and the host polls `done` every few levels.
struct BfsCtl {
    uint32_t F;        // frontier size of the level being expanded
    uint32_t nextF;    // next frontier size (collect's slot counter)
    uint32_t level;    // distance given to this level's discoveries
    uint32_t levels;   // non-empty frontiers so far
    uint32_t done;     // the frontier is empty
    uint32_t pad_;
    uint64_t visited;  // sum of the frontier sizes
    uint64_t scanned;  // sum of the frontiers' degrees (m_C)
};

// Scratch of one traversal, carved from one device buffer: the handle's
// s_bfs, or the static array of the no-allocation entry point.
// The frontier needs no ping-pong: collect writes each discovered vertex's
// id, degree and adjacency offset into (q, qdeg, qoff), which the next
// level's scan and expansion read before that level's collect overwrites them.
struct BfsBufs {
    uint32_t* bitmap;   // visited as of the start of the level
    uint32_t* full;     // bit w: bitmap word w is all ones (every one of its 32 vertices visited)
    int32_t* q;
    int32_t* qdeg;
    int64_t* qoff;
    uint32_t* P;
    uint32_t* tile_lo;
    uint8_t* state;     // one byte per vertex, padded to whole bitmap words
    uint64_t* tot;      // [0] level totals (edges << 32 | groups)
    BfsCtl* ctl;
    void* scan_tmp;     // scan partials
};

// Lays the buffers of a traversal of a graph with nv vertices and ne edges
// out from `base` (nullptr: sizes only); returns the bytes they take.
int64_t bfs_layout(int64_t nv, int64_t ne, char* base, BfsBufs* b);
// BFS TopDown from `source` into dist[nv] over laid-out buffers, reading the
// per-level totals back through `host` (pinned, at least 16 bytes).
// Allocates nothing, except the two timing events when `may_time` and
// stats->time_kernels ask for kernel timing.
hn_status bfs_run(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats, hipStream_t st,
                  const BfsBufs& b, bool may_time, uint64_t* host);

}  // namespace hn
