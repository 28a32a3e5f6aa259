// bfs.hpp -- the single-source BFS level loop over caller-placed scratch
// (bfs.hip), shared by the entry points that own their scratch differently.
#pragma once
#include "graph.hpp"

namespace hn {

// Level-loop control on the device: the loop runs without a host round trip
// per level, and the host polls it every few levels.  Launch L (from 1)
// expands the frontier of level L and collects level L's discoveries, whose
// distance is L; the host passes L to both kernels.
struct BfsCtl {
    uint64_t fc[2];    // frontier of level L: fc[L & 1] = size << 32 | groups
                       // (collect L adds into fc[(L + 1) & 1], which expand L clears)
    uint32_t levels;   // last non-empty level = number of non-empty frontiers
    uint32_t pad_;
    uint64_t visited;     // sum of the frontier sizes
    uint64_t scanned[2];  // sum of the frontiers' degrees (m_C) as of expansion L, in [L & 1]
    uint64_t src_deg;     // the source's degree
    uint64_t dsum;        // sum over the expanded frontiers of size x distance
};

// Scratch of one traversal, carved from one device buffer: the handle's
// s_bfs, or the static array of the no-allocation entry point.
// The frontier needs no ping-pong: collect writes each discovered vertex's
// degree, adjacency offset (in 128-byte units) and group prefix into
// (qdeg, qbase, P), and the first frontier index of every expansion tile into
// tile_lo; the next level's expansion reads them before that level's collect
// overwrites them.
struct BfsBufs {
    uint32_t* bitmap;   // visited as of the start of the level
    uint32_t* full;     // bit w: bitmap word w is all ones (every one of its 32 vertices visited)
    int32_t* qdeg;
    uint32_t* qbase;
    uint32_t* P;
    uint32_t* tile_lo;
    uint8_t* state;     // one byte per vertex, padded to whole bitmap words
    uint64_t* part;     // per collect workgroup: the degrees of its discoveries so far
    BfsCtl* ctl;
};

// Lays the buffers of a traversal of a graph with nv vertices and ne edges
// out from `base` (nullptr: sizes only); returns the bytes they take.
int64_t bfs_layout(int64_t nv, int64_t ne, char* base, BfsBufs* b);
// BFS TopDown from `source` into dist[nv] over laid-out buffers, reading the
// level counters back through `host` (pinned, at least sizeof(BfsCtl) bytes).
// With max_levels > 0 the traversal stops after that many levels (vertices
// farther than max_levels keep INT32_MAX; the 2-hop queries of the serving
// loop).  Allocates nothing, except the timing events when `may_time` and
// stats->time_kernels ask for kernel timing.
hn_status bfs_run(hn_graph* g, int32_t source, int32_t* dist, hn_bfs_stats* stats, hipStream_t st,
                  const BfsBufs& b, bool may_time, uint64_t* host, int32_t max_levels = 0);

}  // namespace hn
