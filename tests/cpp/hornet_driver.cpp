// hornet_driver.cpp -- C++ user of the drop-in headers, shaped like the
// reference's drivers (hornetsnest/test/BFSTest.cu, TriangleTest2.cu,
// hornet/test/HornetInsertTest.cu).  Compiled with hipcc against
// include/hornet/*.hpp + include/hornetsnest/*.hpp and linked with
// libhornet_hip.so; tests/test_cpp_facade_gpu.py runs it and checks every
// printed result against the oracle.
//
// Input graph file: int64 nv, int64 ne, int64 offsets[nv+1], int32 edges[ne].
// Commands:
//   bfs    <graph> <source> <out.bin>   BfsTopDown::run (operator path) +
//                                       run_native; writes int32 dist[nv]
//   tc     <graph>                      TriangleCounting2::run on the stored graph
//   ops    <graph>                      forAll* / load-balancing checksums
//   update <graph> <batch> <out.bin>    insert (both dedups) then erase half;
//                                       batch file: int64 n, int32 src[n], dst[n];
//                                       writes the COO after each step
//   bfsfile <file> <out.bin>            BFSTest.cu:18-44 shape: graph::GraphStd
//                                       read of a .mtx/.gr/.bin file, BFS from
//                                       max_out_degree_id; writes int32 dist[nv]
//   tcfile  <file>                      TriangleTest2.cu:86-100 shape: read
//                                       UNDIRECTED, DIRECTED_BY_DEGREE | SORT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "Graph/GraphStd.hpp"
#include "hornet/hornet.hpp"
#include "hornet/operators.hpp"
#include "hornetsnest/bfs_topdown.hpp"
#include "hornetsnest/triangle2.hpp"

using namespace hornets_nest;

struct Csr {
    int64_t nv = 0, ne = 0;
    std::vector<int> off;
    std::vector<int> adj;
};

static Csr read_graph(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    Csr g;
    f.read(reinterpret_cast<char*>(&g.nv), 8);
    f.read(reinterpret_cast<char*>(&g.ne), 8);
    std::vector<int64_t> o((size_t)g.nv + 1);
    f.read(reinterpret_cast<char*>(o.data()), (std::streamsize)(8 * o.size()));
    g.off.assign(o.begin(), o.end());
    g.adj.resize((size_t)g.ne);
    f.read(reinterpret_cast<char*>(g.adj.data()), (std::streamsize)(4 * g.adj.size()));
    return g;
}

static void write_coo(std::ofstream& out, const HornetGraph& h) {
    auto coo = h.getCOO();
    std::vector<int> s, d;
    coo.toHost(s, d);
    int64_t n = (int64_t)s.size();
    out.write(reinterpret_cast<const char*>(&n), 8);
    out.write(reinterpret_cast<const char*>(s.data()), (std::streamsize)(4 * s.size()));
    out.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)(4 * d.size()));
}

struct CountEdges {
    unsigned long long* acc;   // [0] edges, [1] sum of dst, [2] sum of src
    OPERATOR(Vertex& v, Edge& e) {
        atomicAdd(acc, 1ull);
        atomicAdd(acc + 1, (unsigned long long)e.dst_id());
        atomicAdd(acc + 2, (unsigned long long)v.id());
    }
};
struct CountPairs {
    unsigned long long* acc;   // pairs with deg(src) > deg(dst)
    OPERATOR(Vertex& src, Vertex& dst) {
        if (src.degree() > dst.degree()) atomicAdd(acc, 1ull);
    }
};
struct SumDegrees {
    unsigned long long* acc;
    OPERATOR(Vertex& v) { atomicAdd(acc, (unsigned long long)v.degree()); }
};

static std::vector<unsigned long long> fetch(unsigned long long* d, int n) {
    std::vector<unsigned long long> h((size_t)n);
    gpu::copyToHost(d, (size_t)n, h.data());
    gpu::memsetZero(d, (size_t)n);
    return h;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: hornet_driver bfs|tc|ops|update <graph> ...\n");
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "bfsfile" && argc >= 4) {
        graph::GraphStd<vid_t, eoff_t> graph;
        graph.read(argv[2]);
        HornetInit hornet_init(graph.nV(), graph.nE(), graph.csr_out_offsets(), graph.csr_out_edges());
        HornetGraph hornet_graph(hornet_init);
        BfsTopDown bfs_top_down(hornet_graph);
        const vid_t root = graph.max_out_degree_id();
        bfs_top_down.set_parameters(root);
        bfs_top_down.run();
        std::vector<int> d((size_t)graph.nV());
        gpu::copyToHost(bfs_top_down.distances(), d.size(), d.data());
        std::printf("bfsfile nv=%d ne=%d root=%d levels=%d validate=%d\n", graph.nV(), graph.nE(), root,
                    bfs_top_down.levels(), bfs_top_down.validate() ? 1 : 0);
        std::ofstream out(argv[3], std::ios::binary);
        out.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)(4 * d.size()));
        return 0;
    }
    if (cmd == "tcfile") {
        graph::GraphStd<vid_t, eoff_t> graph(graph::structure_prop::UNDIRECTED);
        graph.read(argv[2], graph::parsing_prop::DIRECTED_BY_DEGREE | graph::parsing_prop::PRINT_INFO |
                                graph::parsing_prop::SORT);
        HornetInit hornet_init(graph.nV(), graph.nE(), graph.csr_out_offsets(), graph.csr_out_edges());
        HornetGraph hornet_graph(hornet_init);
        TriangleCounting2 tc(hornet_graph);
        tc.init();
        tc.run(1);
        std::printf("tcfile nv=%d ne=%d total=%llu\n", graph.nV(), graph.nE(), tc.total_count());
        return 0;
    }
    Csr g = read_graph(argv[2]);
    HornetInit init((int)g.nv, (int)g.ne, g.off.data(), g.adj.data());
    HornetGraph hornet_graph(init);
    std::printf("graph nv=%d ne=%d max_degree_id=%d max_degree=%d\n", hornet_graph.nV(), hornet_graph.nE(),
                hornet_graph.max_degree_id(), hornet_graph.max_degree());

    if (cmd == "bfs" && argc >= 5) {
        BfsTopDown bfs(hornet_graph);
        bfs.set_parameters(std::atoi(argv[3]));
        bfs.run();
        std::vector<int> d1((size_t)g.nv), d2((size_t)g.nv);
        gpu::copyToHost(bfs.distances(), (size_t)g.nv, d1.data());
        const bool ok = bfs.validate();
        hn_bfs_stats st = bfs.run_native();
        gpu::copyToHost(bfs.distances(), (size_t)g.nv, d2.data());
        std::printf("bfs levels=%d validate=%d native_equal=%d native_levels=%d\n", bfs.levels(), ok ? 1 : 0,
                    d1 == d2 ? 1 : 0, st.levels);
        std::ofstream out(argv[4], std::ios::binary);
        out.write(reinterpret_cast<const char*>(d1.data()), (std::streamsize)(4 * d1.size()));
        return 0;
    }
    if (cmd == "tc") {
        TriangleCounting2 tc(hornet_graph);
        tc.init();
        tc.run(1);
        std::printf("tc total=%llu sum=%llu\n", tc.total_count(), tc.countTriangles());
        return 0;
    }
    if (cmd == "ops") {
        unsigned long long* acc = nullptr;
        gpu::allocate(acc, 4);
        gpu::memsetZero(acc, 4);
        load_balancing::BinarySearch bs(hornet_graph);
        forAllEdges(hornet_graph, CountEdges{acc}, bs);
        auto a = fetch(acc, 3);
        std::printf("ops binary_search edges=%llu dst_sum=%llu src_sum=%llu\n", a[0], a[1], a[2]);
        load_balancing::VertexBased<4> vb4(hornet_graph);
        forAllEdges(hornet_graph, CountEdges{acc}, vb4);
        a = fetch(acc, 3);
        std::printf("ops vertex_based4 edges=%llu dst_sum=%llu src_sum=%llu\n", a[0], a[1], a[2]);
        load_balancing::VertexBased<64> vb64(hornet_graph);
        forAllEdges(hornet_graph, CountEdges{acc}, vb64);
        a = fetch(acc, 3);
        std::printf("ops vertex_based64 edges=%llu dst_sum=%llu src_sum=%llu\n", a[0], a[1], a[2]);
        forAllEdgeVertexPairs(hornet_graph, CountPairs{acc}, bs);
        a = fetch(acc, 1);
        std::printf("ops pairs_binary_search higher=%llu\n", a[0]);
        load_balancing::VertexBased1 vb1(hornet_graph);
        forAllEdgeVertexPairs(hornet_graph, CountPairs{acc}, vb1);
        a = fetch(acc, 1);
        std::printf("ops pairs_vertex_based1 higher=%llu\n", a[0]);
        forAllVertices(hornet_graph, SumDegrees{acc});
        a = fetch(acc, 1);
        std::printf("ops vertices degree_sum=%llu\n", a[0]);
        // a queue holding every even vertex
        TwoLevelQueue<vid_t> q(hornet_graph, 1.0f);
        std::vector<int> ev;
        for (int v = 0; v < (int)g.nv; v += 2) ev.push_back(v);
        q.insert(ev.data(), (int)ev.size());
        forAllEdges(hornet_graph, q, CountEdges{acc}, bs);
        a = fetch(acc, 3);
        std::printf("ops queue_even edges=%llu dst_sum=%llu src_sum=%llu\n", a[0], a[1], a[2]);
        unsigned long long* cnt = acc;
        forAllnumV(hornet_graph, [=] __device__(int i) { atomicAdd(cnt, (unsigned long long)i); });
        a = fetch(acc, 1);
        std::printf("ops numv id_sum=%llu\n", a[0]);
        gpu::free(acc);
        return 0;
    }
    if (cmd == "update" && argc >= 5) {
        std::ifstream f(argv[3], std::ios::binary);
        int64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        std::vector<int> s((size_t)n), d((size_t)n);
        f.read(reinterpret_cast<char*>(s.data()), (std::streamsize)(4 * n));
        f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)(4 * n));
        std::ofstream out(argv[4], std::ios::binary);
        hornet::BatchUpdatePtr<int, hornet::EMPTY, hornet::DeviceType::HOST> ptr((int)n, s.data(), d.data());
        hornet::gpu::BatchUpdate<int> batch(ptr);
        hornet_graph.insert(batch, true, true);
        std::printf("update inserted=%lld ne=%d\n", (long long)batch.applied(), hornet_graph.nE());
        write_coo(out, hornet_graph);
        hornet::BatchUpdatePtr<int, hornet::EMPTY, hornet::DeviceType::HOST> half((int)(n / 2), s.data(), d.data());
        hornet::gpu::BatchUpdate<int> ebatch(half);
        hornet_graph.erase(ebatch);
        std::printf("update erased=%lld ne=%d\n", (long long)ebatch.applied(), hornet_graph.nE());
        write_coo(out, hornet_graph);
        return 0;
    }
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
}
