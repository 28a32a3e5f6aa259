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
// validate() recounts on the host from the exported CSR (the reference's
// validate is a stub returning true; its host counter,
// hornetsnest/test/TriangleTest2.cu:22-69, is the merge count used here) and
// compares with the last run's total.
#pragma once

#include <vector>

#include "hornet/operators.hpp"

namespace hornets_nest {

using triangle_t = unsigned long long;

namespace detail {
// sum over stored edges (s, d), s != d, deg(s) >= 2, of |N(s) ∩ N(d)| by merge;
// with `orient` the symmetric CSR is first oriented (deg u, u) < (deg v, v)
// (xlib/src/Graph/GraphStd.cpp:247-304)
inline triangle_t host_tc2_total(const std::vector<int64_t>& off, const std::vector<int32_t>& adj, bool orient) {
    const int64_t nv = (int64_t)off.size() - 1;
    std::vector<int64_t> o = off;
    std::vector<int32_t> a = adj;
    if (orient) {
        o.assign((size_t)nv + 1, 0);
        a.clear();
        for (int64_t u = 0; u < nv; u++) {
            const int64_t du = off[u + 1] - off[u];
            for (int64_t j = off[u]; j < off[u + 1]; j++) {
                const int32_t v = adj[j];
                const int64_t dv = off[v + 1] - off[v];
                if (du < dv || (du == dv && u < v)) a.push_back(v);
            }
            o[u + 1] = (int64_t)a.size();
        }
    }
    triangle_t sum = 0;
    for (int64_t s = 0; s < nv; s++) {
        const int64_t ds = o[s + 1] - o[s];
        if (ds < 2) continue;
        for (int64_t j = o[s]; j < o[s + 1]; j++) {
            const int32_t d = a[j];
            if (d == s) continue;
            int64_t x = o[s], y = o[d];
            const int64_t xe = o[s + 1], ye = o[d + 1];
            while (x < xe && y < ye) {
                if (a[x] == a[y]) { sum++; x++; y++; }
                else if (a[x] < a[y]) x++;
                else y++;
            }
        }
    }
    return sum;
}
}  // namespace detail

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
        oriented = false;
    }
    void run_undirected() {
        uint64_t t = 0;
        hornet::check(hn_tc_count(hornet.handle(), &t, reinterpret_cast<uint64_t*>(triPerVertex), hornet.stream()),
                      "TriangleCounting2::run_undirected");
        total = t;
        oriented = true;
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
    // the last run's total (and, after init(), the sum of the per-vertex
    // counters) equals a sequential host count of the same graph
    bool validate() override {
        std::vector<int64_t> off;
        std::vector<int32_t> adj;
        hornet.getCSR().toHost(off, adj);
        const triangle_t host = detail::host_tc2_total(off, adj, oriented);
        if (host != total) return false;
        return !triPerVertex || countTriangles() == host;
    }

protected:
    triangle_t* triPerVertex = nullptr;
    triangle_t total = 0;
    bool oriented = false;   // the last run was run_undirected()
};

}  // namespace hornets_nest
