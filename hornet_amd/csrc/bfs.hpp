// bfs.hpp -- the single-source BFS level loop over caller-placed scratch
// (bfs.hip), shared by the entry points that own their scratch differently.
#pragma once
#include "graph.hpp"

namespace hn {

// Scratch of one traversal, carved from one device buffer: the handle's
// s_bfs, or the static array of the no-allocation entry point.
// The frontier needs no ping-pong: q is gathered into (qdeg, qoff) and
// scanned into P / tile_lo before collect overwrites q with level L+1.
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
    uint32_t* cnt;      // [0] next frontier size
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
