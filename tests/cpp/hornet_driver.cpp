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
//   ops2   <graph> <source> <out.bin>   BfsTopDown2, forAll(queue), forAllVertexPairs,
//                                       forAllnumE, forAllEdges without LB object,
//                                       forAllAdjUnions (triangle credits per vertex
//                                       with WORK_FACTOR 0, 1, 1000 and from a pair
//                                       queue); writes int32 dist[nv] then 4 x
//                                       uint64 credits[nv]
//   snapshot <graph> <batch> <ckpt.bin> <out.bin>
//                                       freeze a HornetStatic copy, insert the batch
//                                       into the live graph, checkpoint the live
//                                       graph through GraphStd::writeBinary, resume it
//                                       as a HornetStatic; writes the COO of the
//                                       frozen copy, then of the resumed graph
//   meta   <graph> <batch> <out.bin>    HornetInsertTestWeighted.cu shape: a graph with
//                                       TypeList<int, double> edge meta-data, a weighted
//                                       insert, a forAllEdges operator reading
//                                       edge.field<N>(); writes the COO with both fields
//   bfsfile <file> <out.bin>            BFSTest.cu:18-44 shape: graph::GraphStd
//                                       read of a .mtx/.gr/.bin file, BFS from
//                                       max_out_degree_id; writes int32 dist[nv]
//   tcfile  <file>                      TriangleTest2.cu:86-100 shape: read
//                                       UNDIRECTED, DIRECTED_BY_DEGREE | SORT
//   api     <graph> <out.bin>           Hornet(), Hornet(nV) + insert, getCSR, reset,
//                                       print, insertEdgeData<N>, TriangleCounting2::validate;
//                                       writes the CSR after the insert and after the reset
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

// |N(u) ∩ N(v)| credited to u, honouring forAllAdjUnions' FLAG contract
struct AdjCount {
    unsigned long long* per_vertex;
    unsigned long long* calls;   // [0] FLAG 0 calls, [1] FLAG 1 calls
    OPERATOR(Vertex& u, Vertex&, vid_t* ub, vid_t* ue, vid_t* vb, vid_t* ve, int FLAG) {
        unsigned long long c = 0;
        if (!FLAG) {
            while (ub <= ue && vb <= ve) {
                if (*ub < *vb) ub++;
                else if (*vb < *ub) vb++;
                else c++, ub++, vb++;
            }
        } else {
            for (; ub <= ue; ub++) {
                const vid_t* lo = vb;
                const vid_t* hi = ve + 1;
                while (lo < hi) {
                    const vid_t* mid = lo + (hi - lo) / 2;
                    if (*mid < *ub) lo = mid + 1;
                    else hi = mid;
                }
                c += (lo <= ve && *lo == *ub);
            }
        }
        if (c) atomicAdd(per_vertex + u.id(), c);
        atomicAdd(calls + FLAG, 1ull);
    }
};
struct SumItems {
    unsigned long long* acc;
    OPERATOR(vid_t& x) { atomicAdd(acc, (unsigned long long)x); }
};
struct SumPairDegrees {
    unsigned long long* acc;
    OPERATOR(Vertex& a, Vertex& b) { atomicAdd(acc, (unsigned long long)(a.degree() * 3 + b.degree())); }
};

// sums of the two meta-data fields over all edges (the double field doubled)
struct SumFields {
    unsigned long long* acc;
    OPERATOR(Vertex&, Edge& e) {
        atomicAdd(acc, (unsigned long long)(long long)e.template field<0>());
        atomicAdd(acc + 1, (unsigned long long)(long long)(2.0 * e.template field<1, double>()));
    }
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
    if (cmd == "ops2" && argc >= 5) {
        std::ofstream out(argv[4], std::ios::binary);
        BfsTopDown2 bfs2(hornet_graph);
        bfs2.set_parameters(std::atoi(argv[3]));
        bfs2.run();
        std::vector<int> dist((size_t)g.nv);
        gpu::copyToHost(bfs2.distances(), (size_t)g.nv, dist.data());
        std::printf("ops2 bfs2 levels=%d validate=%d enqueued=%d\n", bfs2.levels(), bfs2.validate() ? 1 : 0,
                    bfs2.enqueue_items());
        out.write(reinterpret_cast<const char*>(dist.data()), (std::streamsize)(4 * dist.size()));
        unsigned long long* acc = nullptr;
        gpu::allocate(acc, 4);
        gpu::memsetZero(acc, 4);
        unsigned long long* cnt = acc;
        forAllnumE(hornet_graph, [=] __device__(int i) { atomicAdd(cnt, (unsigned long long)i); });
        auto a = fetch(acc, 1);
        std::printf("ops2 nume id_sum=%llu\n", a[0]);
        TwoLevelQueue<vid_t> q(hornet_graph, 1.0f);
        std::vector<int> third;
        for (int v = 0; v < (int)g.nv; v += 3) third.push_back(v);
        q.insert(third.data(), (int)third.size());
        forAll(q, SumItems{acc});
        a = fetch(acc, 1);
        std::printf("ops2 queue_items sum=%llu\n", a[0]);
        TwoLevelQueue<vid2_t> pq(hornet_graph, 1.0f);
        std::vector<vid2_t> pairs;
        for (int v = 0; v + 1 < (int)g.nv; v += 2) pairs.push_back(make_int2(v, v + 1));
        pq.insert(pairs.data(), (int)pairs.size());
        forAllVertexPairs(hornet_graph, pq, SumPairDegrees{acc});
        a = fetch(acc, 1);
        std::printf("ops2 vertex_pairs sum=%llu\n", a[0]);
        forAllEdges(hornet_graph, CountEdges{acc});
        a = fetch(acc, 3);
        std::printf("ops2 edges_default edges=%llu dst_sum=%llu src_sum=%llu\n", a[0], a[1], a[2]);
        unsigned long long* per = nullptr;
        gpu::allocate(per, (size_t)g.nv);
        std::vector<unsigned long long> hper((size_t)g.nv);
        auto adj_run = [&](const char* tag, int wf, bool from_queue) {
            gpu::memsetZero(per, (size_t)g.nv);
            if (from_queue) {
                auto coo = hornet_graph.getCOO();
                std::vector<int> cs, cd;
                coo.toHost(cs, cd);
                std::vector<vid2_t> ep(cs.size());
                for (size_t i = 0; i < cs.size(); i++) ep[i] = make_int2(cs[i], cd[i]);
                TwoLevelQueue<vid2_t> eq((size_t)std::max<size_t>(ep.size(), 1));
                eq.insert(ep.data(), (int)ep.size());
                forAllAdjUnions(hornet_graph, eq, AdjCount{per, acc}, wf);
            } else {
                forAllAdjUnions(hornet_graph, AdjCount{per, acc}, wf);
            }
            auto c = fetch(acc, 2);
            gpu::copyToHost(per, (size_t)g.nv, hper.data());
            unsigned long long tot = 0;
            for (auto x : hper) tot += x;
            std::printf("ops2 adj_unions_%s total=%llu merge_calls=%llu search_calls=%llu\n", tag, tot, c[0], c[1]);
            out.write(reinterpret_cast<const char*>(hper.data()), (std::streamsize)(8 * hper.size()));
        };
        adj_run("wf0", 0, false);
        adj_run("wf1", 1, false);
        adj_run("wf1000", 1000, false);
        adj_run("queue", 1, true);
        gpu::free(per);
        gpu::free(acc);
        return 0;
    }
    if (cmd == "meta" && argc >= 5) {
        using W = hornet::TypeList<int, double>;
        std::vector<int> w0((size_t)g.ne);
        std::vector<double> w1((size_t)g.ne);
        for (size_t j = 0; j < w0.size(); j++) w0[j] = (int)(j * 7 + 1), w1[j] = 0.5 * (double)j;
        hornet::HornetInit<int, hornet::EMPTY, W> winit((int)g.nv, (int)g.ne, g.off.data(), g.adj.data());
        winit.insertEdgeData(w0.data(), w1.data());
        hornet::gpu::Hornet<int, hornet::EMPTY, W> wh(winit);
        std::ifstream f(argv[3], std::ios::binary);
        int64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        std::vector<int> s((size_t)n), d((size_t)n), b0((size_t)n);
        std::vector<double> b1((size_t)n);
        f.read(reinterpret_cast<char*>(s.data()), (std::streamsize)(4 * n));
        f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)(4 * n));
        for (int64_t i = 0; i < n; i++) b0[(size_t)i] = (int)(100000 + i), b1[(size_t)i] = -1.0 * (double)i;
        hornet::BatchUpdatePtr<int, W, hornet::DeviceType::HOST> ptr((int)n, s.data(), d.data(), b0.data(), b1.data());
        hornet::gpu::BatchUpdate<int, W> batch(ptr);
        wh.insert(batch);
        unsigned long long* acc = nullptr;
        gpu::allocate(acc, 2);
        gpu::memsetZero(acc, 2);
        load_balancing::BinarySearch lb(wh);
        forAllEdges(wh, SumFields{acc}, lb);
        auto a = fetch(acc, 2);
        gpu::free(acc);
        std::printf("meta ne=%d inserted=%lld sum0=%lld sum1=%lld\n", wh.nE(), (long long)batch.applied(),
                    (long long)a[0], (long long)a[1]);
        auto coo = wh.getCOO();
        std::vector<int> cs, cd;
        std::vector<int> m0;
        std::vector<double> m1;
        coo.toHost(cs, cd);
        coo.edgeMetaToHost<0>(m0);
        coo.edgeMetaToHost<1>(m1);
        std::ofstream out(argv[4], std::ios::binary);
        int64_t m = (int64_t)cs.size();
        out.write(reinterpret_cast<const char*>(&m), 8);
        out.write(reinterpret_cast<const char*>(cs.data()), (std::streamsize)(4 * m));
        out.write(reinterpret_cast<const char*>(cd.data()), (std::streamsize)(4 * m));
        out.write(reinterpret_cast<const char*>(m0.data()), (std::streamsize)(4 * m));
        out.write(reinterpret_cast<const char*>(m1.data()), (std::streamsize)(8 * m));
        return 0;
    }
    if (cmd == "snapshot" && argc >= 6) {
        static_assert(hornet::IsHornetStatic<hornet::gpu::HornetStatic<int>>::value, "trait");
        static_assert(!hornet::IsHornetStatic<HornetGraph>::value, "trait");
        hornet::gpu::HornetStatic<int> frozen = hornet_graph.snapshot();
        std::ifstream f(argv[3], std::ios::binary);
        int64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        std::vector<int> s((size_t)n), d((size_t)n);
        f.read(reinterpret_cast<char*>(s.data()), (std::streamsize)(4 * n));
        f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)(4 * n));
        hornet::BatchUpdatePtr<int, hornet::EMPTY, hornet::DeviceType::HOST> ptr((int)n, s.data(), d.data());
        hornet::gpu::BatchUpdate<int> batch(ptr);
        hornet_graph.insert(batch, true, true);
        // checkpoint: CSR of the live graph -> GraphStd -> binary file
        const int64_t nv = hornet_graph.nV(), ne = hornet_graph.nE64();
        int64_t* d_off = nullptr;
        int32_t* d_adj = nullptr;
        gpu::allocate(d_off, (size_t)nv + 1);
        gpu::allocate(d_adj, (size_t)std::max<int64_t>(ne, 1));
        hornet::check(hn_graph_export_csr(hornet_graph.handle(), d_off, d_adj, ne, nullptr), "export_csr");
        std::vector<int64_t> off((size_t)nv + 1);
        std::vector<int> adj((size_t)ne);
        gpu::copyToHost(d_off, off.size(), off.data());
        if (ne > 0) gpu::copyToHost(d_adj, adj.size(), adj.data());
        gpu::free(d_off, d_adj);
        {
            graph::GraphStd<int, int64_t> ckpt(off.data(), (int)nv, adj.data(), ne);
            ckpt.writeBinary(argv[4]);
        }
        graph::GraphStd<int, int> back;
        back.read(argv[4]);
        HornetInit rinit(back.nV(), back.nE(), back.csr_out_offsets(), back.csr_out_edges());
        hornet::gpu::HornetStatic<int> resumed(rinit);
        std::printf("snapshot frozen_ne=%d live_ne=%d resumed_ne=%d inserted=%lld\n", frozen.nE(), hornet_graph.nE(),
                    resumed.nE(), (long long)batch.applied());
        std::ofstream out(argv[5], std::ios::binary);
        auto put = [&](const hornet::gpu::HornetStatic<int>& h) {
            auto coo = h.getCOO();
            std::vector<int> cs, cd;
            coo.toHost(cs, cd);
            int64_t m = (int64_t)cs.size();
            out.write(reinterpret_cast<const char*>(&m), 8);
            out.write(reinterpret_cast<const char*>(cs.data()), (std::streamsize)(4 * cs.size()));
            out.write(reinterpret_cast<const char*>(cd.data()), (std::streamsize)(4 * cd.size()));
        };
        put(frozen);
        put(resumed);
        return 0;
    }
    if (cmd == "api" && argc >= 4) {
        // Hornet(void); Hornet(nV) filled by one insert; getCSR; reset(HornetInit&);
        // print; HornetInit::insertEdgeData<N>; TriangleCounting2::validate
        hornet::gpu::Hornet<int> none;
        std::printf("api default nv=%d ne=%d\n", none.nV(), none.nE());
        HornetGraph grown((int)g.nv);
        std::vector<int> s, d;
        for (int64_t v = 0; v < g.nv; v++)
            for (int j = g.off[(size_t)v]; j < g.off[(size_t)v + 1]; j++) s.push_back((int)v), d.push_back(g.adj[(size_t)j]);
        hornet::BatchUpdatePtr<int, hornet::EMPTY, hornet::DeviceType::HOST> ptr((int)s.size(), s.data(), d.data());
        hornet::gpu::BatchUpdate<int> batch(ptr);
        grown.insert(batch);
        std::ofstream out(argv[3], std::ios::binary);
        auto put_csr = [&](const HornetGraph& h, const char* tag) {
            auto csr = h.getCSR();
            std::vector<int64_t> o64;
            std::vector<int> a;
            csr.toHost(o64, a);
            std::vector<int> o32(o64.size());
            (void)hipMemcpy(o32.data(), csr.offset(), sizeof(int) * o32.size(), hipMemcpyDeviceToHost);
            int same = 1;
            for (size_t i = 0; i < o64.size(); i++) same &= (int64_t)o32[i] == o64[i] ? 1 : 0;
            std::printf("api %s nv=%lld ne=%lld offsets32_equal=%d\n", tag, (long long)csr.nV(), (long long)csr.nE(),
                        same);
            int64_t n[2] = {(int64_t)o64.size() - 1, (int64_t)a.size()};
            out.write(reinterpret_cast<const char*>(n), 16);
            out.write(reinterpret_cast<const char*>(o64.data()), (std::streamsize)(8 * o64.size()));
            out.write(reinterpret_cast<const char*>(a.data()), (std::streamsize)(4 * a.size()));
        };
        std::printf("api grown inserted=%lld\n", (long long)batch.applied());
        put_csr(grown, "grown");
        // reset to the rows of the first half of the vertices
        const int64_t half = g.nv / 2;
        std::vector<int> hoff(g.off.begin(), g.off.end());
        for (int64_t v = half + 1; v <= g.nv; v++) hoff[(size_t)v] = g.off[(size_t)half];
        HornetInit hinit((int)g.nv, g.off[(size_t)half], hoff.data(), g.adj.data());
        grown.reset(hinit);
        put_csr(grown, "reset");
        TriangleCounting2 tc(hornet_graph);
        tc.init();
        tc.run_undirected();
        const int v_und = tc.validate() ? 1 : 0;
        const unsigned long long t_und = tc.countTriangles();
        tc.run();
        std::printf("api tc undirected=%llu validate_undirected=%d stored=%llu validate_stored=%d\n", t_und, v_und,
                    tc.total_count(), tc.validate() ? 1 : 0);
        // one meta-data field given, the other read as zero
        using W = hornet::TypeList<int, double>;
        std::vector<double> w1((size_t)g.ne);
        for (size_t j = 0; j < w1.size(); j++) w1[j] = 0.25 * (double)j;
        hornet::HornetInit<int, hornet::EMPTY, W> winit((int)g.nv, (int)g.ne, g.off.data(), g.adj.data());
        winit.insertEdgeData<1>(w1.data());
        hornet::gpu::Hornet<int, hornet::EMPTY, W> wh(winit);
        auto wcsr = wh.getCSR();
        std::vector<int> f0((size_t)g.ne);
        std::vector<double> f1((size_t)g.ne);
        if (g.ne > 0) {
            (void)hipMemcpy(f0.data(), wcsr.edgeMetaPtr<0>(), 4 * f0.size(), hipMemcpyDeviceToHost);
            (void)hipMemcpy(f1.data(), wcsr.edgeMetaPtr<1>(), 8 * f1.size(), hipMemcpyDeviceToHost);
        }
        long long z = 0;
        double sw = 0;
        for (size_t j = 0; j < f0.size(); j++) z += f0[j] != 0, sw += f1[j];
        std::printf("api meta nonzero0=%lld sum1_x4=%lld\n", z, (long long)(sw * 4.0));
        std::fflush(stdout);
        // print() (HornetInitialize.i.cuh:152-170) of the reset graph
        grown.print();
        std::fflush(stdout);
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
