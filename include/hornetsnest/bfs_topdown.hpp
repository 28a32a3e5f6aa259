// hornetsnest/bfs_topdown.hpp -- BfsTopDown written against the operator layer.
//
// Same API and level loop as HornetsNest's TopDown
// (hornetsnest/include/Static/BreadthFirstSearch/TopDown.cuh:52-70,
//  hornetsnest/src/Static/BreadthFirstSearch/TopDown.cu:68-128): distances
// start at INT_MAX, the source gets 0, level L writes L to unvisited
// neighbours of the level L-1 frontier through forAllEdges with
// BinarySearch load balancing and a TwoLevelQueue.  The operator is the
// reference's non-atomic check-then-set, so a vertex may be enqueued more
// than once per level; every writer stores the same level, so distances are
// exact.
//
// run() is the generic operator path; run_native() calls the library's fused
// level step (hn_bfs_topdown: atomic-free expand + collect, each vertex
// enqueued once), which gives identical distances.  validate() compares with
// a sequential host BFS of the exported CSR (the comparison the reference
// left commented out, TopDown.cu:113-128).
//
// BfsTopDown2 is the reference's deterministic variant
// (hornetsnest/include/Static/BreadthFirstSearch/TopDown2.cuh,
//  src/Static/BreadthFirstSearch/TopDown2.cu:70-120): the operator claims a
// vertex with atomicCAS(dist, INF, level), so each vertex is enqueued once.
#pragma once

#include <limits>
#include <vector>

#include "hornet/operators.hpp"

namespace hornets_nest {

using dist_t = int;
using HornetGraph = ::hornet::gpu::Hornet<vert_t>;
using HornetInit = ::hornet::HornetInit<vert_t>;

// distances equal to a sequential host BFS of the exported CSR
inline bool bfs_matches_host(HornetGraph& hornet, vert_t bfs_source, const dist_t* d_distances) {
    constexpr dist_t INF = std::numeric_limits<dist_t>::max();
    const int64_t nv = hornet.nV(), ne = hornet.nE64();
    int64_t* d_off = nullptr;
    int32_t* d_adj = nullptr;
    gpu::allocate(d_off, (size_t)nv + 1);
    gpu::allocate(d_adj, (size_t)(ne > 0 ? ne : 1));
    hornet::check(hn_graph_export_csr(hornet.handle(), d_off, d_adj, ne, hornet.stream()), "validate");
    std::vector<int64_t> off((size_t)nv + 1);
    std::vector<int32_t> adj((size_t)ne), got((size_t)nv);
    gpu::copyToHost(d_off, (size_t)nv + 1, off.data());
    if (ne > 0) gpu::copyToHost(d_adj, (size_t)ne, adj.data());
    gpu::copyToHost(d_distances, (size_t)nv, got.data());
    gpu::free(d_off, d_adj);
    std::vector<int32_t> dist((size_t)nv, INF), q;
    q.reserve((size_t)nv);
    dist[bfs_source] = 0;
    q.push_back(bfs_source);
    for (size_t h = 0; h < q.size(); h++) {
        const int32_t u = q[h];
        for (int64_t j = off[u]; j < off[u + 1]; j++)
            if (dist[adj[j]] == INF) {
                dist[adj[j]] = dist[u] + 1;
                q.push_back(adj[j]);
            }
    }
    return dist == got;
}

struct BfsTopDownOperator {
    dist_t* __restrict__ d_distances;
    dist_t current_level;
    TwoLevelQueue<vert_t> queue;

    OPERATOR(Vertex&, Edge& edge) {
        const vert_t dst = edge.dst_id();
        if (d_distances[dst] == std::numeric_limits<dist_t>::max()) {
            d_distances[dst] = current_level;
            queue.insert(dst);
        }
    }
};

class BfsTopDown : public StaticAlgorithm<HornetGraph> {
public:
    static constexpr dist_t INF = std::numeric_limits<dist_t>::max();

    explicit BfsTopDown(HornetGraph& g) : StaticAlgorithm(g), queue(g), load_balancing(g) {
        queue.set_stream(g.stream());
        gpu::allocate(d_distances, (size_t)g.nV());
        reset();
    }
    ~BfsTopDown() override { gpu::free(d_distances); }

    void reset() override {
        current_level = 1;
        queue.clear();
        dist_t* distances = d_distances;
        forAllnumV(hornet, [=] __device__(int i) { distances[i] = INF; });
    }

    void set_parameters(vert_t source) {
        bfs_source = source;
        queue.insert(bfs_source);                 // source in the frontier
        gpu::memsetZero(d_distances + bfs_source);  // dist[source] = 0
    }

    void run() override {
        while (queue.size() > 0) {
            forAllEdges(hornet, queue, BfsTopDownOperator{d_distances, current_level, queue}, load_balancing);
            current_level++;
            queue.swap();
        }
    }

    // the library's fused level step for the same traversal
    hn_bfs_stats run_native() {
        hn_bfs_stats st{};
        hornet::check(hn_bfs_topdown(hornet.handle(), bfs_source, d_distances, &st, hornet.stream()),
                      "BfsTopDown::run_native");
        return st;
    }

    void release() override {
        gpu::free(d_distances);
        d_distances = nullptr;
    }

    bool validate() override { return bfs_matches_host(hornet, bfs_source, d_distances); }

    const dist_t* distances() const noexcept { return d_distances; }
    dist_t levels() const noexcept { return current_level - 1; }

private:
    TwoLevelQueue<vert_t> queue;
    load_balancing::BinarySearch load_balancing;
    dist_t* d_distances = nullptr;
    vert_t bfs_source = 0;
    dist_t current_level = 0;
};

struct BfsOperatorAtomic {
    dist_t current_level;
    dist_t* __restrict__ d_distances;
    TwoLevelQueue<vert_t> queue;

    OPERATOR(Vertex&, Edge& edge) {
        const vert_t dst = edge.dst_id();
        if (atomicCAS(d_distances + dst, std::numeric_limits<dist_t>::max(), current_level) ==
            std::numeric_limits<dist_t>::max())
            queue.insert(dst);
    }
};

class BfsTopDown2 : public StaticAlgorithm<HornetGraph> {
public:
    static constexpr dist_t INF = std::numeric_limits<dist_t>::max();

    // every vertex enters the queue at most once, so nV entries suffice
    explicit BfsTopDown2(HornetGraph& g) : StaticAlgorithm(g), queue(g, 1.0f), load_balancing(g) {
        queue.set_stream(g.stream());
        gpu::allocate(d_distances, (size_t)g.nV());
        reset();
    }
    ~BfsTopDown2() override { gpu::free(d_distances); }

    void reset() override {
        current_level = 1;
        queue.clear();
        dist_t* distances = d_distances;
        forAllnumV(hornet, [=] __device__(int i) { distances[i] = INF; });
    }

    void set_parameters(vert_t source) {
        bfs_source = source;
        queue.insert(bfs_source);
        gpu::memsetZero(d_distances + bfs_source);
    }

    void run() override {
        while (queue.size() > 0) {
            forAllEdges(hornet, queue, BfsOperatorAtomic{current_level, d_distances, queue}, load_balancing);
            queue.swap();
            current_level++;
        }
    }

    void release() override {
        gpu::free(d_distances);
        d_distances = nullptr;
    }

    // the reference's validate() only prints the enqueue count
    bool validate() override { return bfs_matches_host(hornet, bfs_source, d_distances); }

    const dist_t* distances() const noexcept { return d_distances; }
    dist_t levels() const noexcept { return current_level - 1; }
    int enqueue_items() const noexcept { return queue.enqueue_items(); }

private:
    TwoLevelQueue<vert_t> queue;
    load_balancing::BinarySearch load_balancing;
    dist_t* d_distances = nullptr;
    vert_t bfs_source = 0;
    dist_t current_level = 0;
};

}  // namespace hornets_nest
