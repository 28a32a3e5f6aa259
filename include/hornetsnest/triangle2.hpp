// hornetsnest/triangle2.hpp -- TriangleCounting2 with the reference's API.
//
// HornetsNest TriangleCounting2
// (hornetsnest/include/Static/TriangleCounting/triangle2.cuh:18-37,
//  hornetsnest/src/Static/TriangleCounting/triangle2.cu:24-175): per-vertex
// unsigned long long counters, run() intersects N(u) and N(v) for every
// stored edge (skipping src == dst and deg(src) < 2), crediting the endpoint
// with the smaller degree, and countTriangles() sums the counters.  On a graph
// stored DIRECTED_BY_DEGREE (the reference driver's form,
// hornetsnest/test/TriangleTest2.cu:86-87) that sum is the triangle count.
// The intersection runs in libhornet_hip (hn_tc2_stored); run_undirected()
// counts the triangles of a symmetric graph by orienting it on the device
// first (hn_tc_count).
#pragma once

#include <vector>

#include "hornet/operators.hpp"

namespace hornets_nest {

using triangle_t = unsigned long long;

class TriangleCounting2 : public StaticAlgorithm<::hornet::gpu::Hornet<vid_t>> {
public:
    using HornetGraph = ::hornet::gpu::Hornet<vid_t>;

    explicit TriangleCounting2(HornetGraph& g) : StaticAlgorithm(g) {}
    ~TriangleCounting2() override { release(); }

    void init() {
        gpu::allocate(triPerVertex, (size_t)hornet.nV());
        reset();
    }
    void reset() override {
        if (triPerVertex) gpu::memsetZero(triPerVertex, (size_t)hornet.nV());
        total = 0;
    }
    void run() override { run(1); }
    void run(const int WORK_FACTOR) {
        (void)WORK_FACTOR;   // the work binning is internal to the library
        uint64_t t = 0;
        hornet::check(hn_tc2_stored(hornet.handle(), &t, reinterpret_cast<uint64_t*>(triPerVertex), hornet.stream()),
                      "TriangleCounting2::run");
        total = t;
    }
    void run_undirected() {
        uint64_t t = 0;
        hornet::check(hn_tc_count(hornet.handle(), &t, reinterpret_cast<uint64_t*>(triPerVertex), hornet.stream()),
                      "TriangleCounting2::run_undirected");
        total = t;
    }
    triangle_t countTriangles() {
        if (!triPerVertex) return total;   // init() not called: no per-vertex counters
        std::vector<triangle_t> h((size_t)hornet.nV());
        copyTCToHost(h.data());
        triangle_t sum = 0;
        for (triangle_t x : h) sum += x;
        return sum;
    }
    void copyTCToHost(triangle_t* h_tcs) { gpu::copyToHost(triPerVertex, (size_t)hornet.nV(), h_tcs); }
    triangle_t total_count() const noexcept { return total; }
    void release() override {
        gpu::free(triPerVertex);
        triPerVertex = nullptr;
    }
    bool validate() override { return true; }

protected:
    triangle_t* triPerVertex = nullptr;
    triangle_t total = 0;
};

}  // namespace hornets_nest
