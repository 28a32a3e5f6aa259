// Graph/GraphStd.hpp -- graph::GraphStd over the C ABI of libhornet_hip.so
// (hn_gstd_* in hornet_hip.h), with the reference's names and arguments:
//   graph::GraphStd<vid_t, eoff_t>  xlib/include/Graph/GraphStd.hpp
//   structure_prop / parsing_prop   xlib/include/Graph/GraphBase.hpp:46-48,92-123
//   read(filename, prop)            xlib/src/Graph/GraphBase.cpp:130-206
//   csr_*_offsets / csr_*_edges     xlib/include/Graph/GraphStd.i.hpp:239-256
//   max_out_degree_id               GraphStd.i.hpp:296-299
//   writeBinary / writeMarket       xlib/src/Graph/GraphStd.cpp:387-437
// Files: MatrixMarket (.mtx), DIMACS9 (.gr), binary (.bin).  vid_t is int;
// eoff_t is int or int64_t (the library keeps int64 offsets and this header
// narrows them for eoff_t = int, throwing if they do not fit).  Where the
// reference prints and exits on a bad file this throws graph::Error.
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "hornet_hip.h"

namespace graph {

class Error : public std::runtime_error {
public:
    Error(hn_status s, const std::string& where)
        : std::runtime_error(where + ": " + hn_status_string(s) + ": " + hn_last_error()), status(s) {}
    hn_status status;
};

class StructureProp {
public:
    constexpr explicit StructureProp(uint32_t v = 0) noexcept : _v(v) {}
    constexpr StructureProp operator|(StructureProp o) const noexcept { return StructureProp(_v | o._v); }
    constexpr uint32_t value() const noexcept { return _v; }

private:
    uint32_t _v;
};

class ParsingProp {
public:
    constexpr explicit ParsingProp(uint32_t v = 0) noexcept : _v(v) {}
    constexpr ParsingProp operator|(ParsingProp o) const noexcept { return ParsingProp(_v | o._v); }
    constexpr uint32_t value() const noexcept { return _v; }

private:
    uint32_t _v;
};

namespace structure_prop {
constexpr StructureProp NONE(0);
constexpr StructureProp DIRECTED(HN_GS_DIRECTED);
constexpr StructureProp UNDIRECTED(HN_GS_UNDIRECTED);
constexpr StructureProp ENABLE_INGOING(HN_GS_ENABLE_INGOING);
}  // namespace structure_prop

namespace parsing_prop {
constexpr ParsingProp NONE(0);
constexpr ParsingProp RANDOMIZE(HN_PARSE_RANDOMIZE);
constexpr ParsingProp SORT(HN_PARSE_SORT);
constexpr ParsingProp PRINT_INFO(HN_PARSE_PRINT_INFO);
constexpr ParsingProp RM_SINGLETON(HN_PARSE_RM_SINGLETON);
constexpr ParsingProp DIRECTED_BY_DEGREE(HN_PARSE_DIRECTED_BY_DEGREE);
}  // namespace parsing_prop

using namespace structure_prop;
using namespace parsing_prop;

template <typename vid_t = int, typename eoff_t = int>
class GraphStd {
    static_assert(std::is_same<vid_t, int>::value, "vertex ids are int");
    static_assert(std::is_same<eoff_t, int>::value || std::is_same<eoff_t, int64_t>::value,
                  "edge offsets are int or int64_t");

public:
    using degree_t = int;

    explicit GraphStd(StructureProp structure = structure_prop::NONE) noexcept : _structure(structure) {}
    GraphStd(StructureProp structure, const char* filename, ParsingProp prop = parsing_prop::NONE)
        : _structure(structure) {
        read(filename, prop);
    }
    // GraphStd(csr_offsets, nV, csr_edges, nE) -- xlib/src/Graph/GraphStd.cpp:50-59
    // (structure UNDIRECTED), e.g. to write a graph's CSR with writeBinary
    GraphStd(const eoff_t* csr_offsets, vid_t nV, const vid_t* csr_edges, eoff_t nE)
        : _structure(structure_prop::UNDIRECTED) {
        std::vector<int64_t> off(csr_offsets, csr_offsets + nV + 1);
        check(hn_gstd_from_csr(nV, (int64_t)nE, off.data(), csr_edges, &_g), "GraphStd(csr)");
        load();
    }
    GraphStd(const GraphStd&) = delete;
    GraphStd& operator=(const GraphStd&) = delete;
    ~GraphStd() { hn_gstd_destroy(_g); }

    // `seed` drives RANDOMIZE (the reference seeds from the clock)
    void read(const char* filename, ParsingProp prop = parsing_prop::NONE, uint64_t seed = 0) {
        hn_gstd_destroy(_g);
        _g = nullptr;
        check(hn_gstd_read(filename, _structure.value(), prop.value(), seed, &_g), "GraphStd::read");
        load();
    }

private:
    void load() {
        int32_t nv;
        int64_t ne;
        uint32_t st;
        check(hn_gstd_info(_g, &nv, &ne, &st), "GraphStd::read");
        if (std::is_same<eoff_t, int>::value && ne > (int64_t)INT32_MAX)
            throw std::overflow_error("GraphStd<int, int>: more edges than int offsets hold");
        _nV = nv;
        _nE = (eoff_t)ne;
        _st = st;
        const int64_t* o;
        const int32_t* e;
        check(hn_gstd_csr_out(_g, &o, &e), "GraphStd::read");
        _out_edges = e;
        _out_off.assign(o, o + nv + 1);
        _in_off.clear();
        _in_edges = nullptr;
        if (hn_gstd_csr_in(_g, &o, &e) == HN_OK) {
            _in_edges = e;
            _in_off.assign(o, o + nv + 1);
        }
    }

public:
    vid_t nV() const noexcept { return _nV; }
    eoff_t nE() const noexcept { return _nE; }
    bool is_directed() const noexcept { return _st & HN_GS_DIRECTED; }
    bool is_undirected() const noexcept { return _st & HN_GS_UNDIRECTED; }

    const eoff_t* csr_out_offsets() const noexcept { return _out_off.data(); }
    const vid_t* csr_out_edges() const noexcept { return _out_edges; }
    // nullptr unless DIRECTED | ENABLE_INGOING or UNDIRECTED
    const eoff_t* csr_in_offsets() const noexcept { return _in_off.empty() ? nullptr : _in_off.data(); }
    const vid_t* csr_in_edges() const noexcept { return _in_edges; }

    degree_t out_degree(vid_t v) const noexcept { return (degree_t)(_out_off[v + 1] - _out_off[v]); }
    vid_t max_out_degree_id() const {
        int32_t id;
        check(hn_gstd_max_out_degree_id(_g, &id), "GraphStd::max_out_degree_id");
        return id;
    }
    degree_t max_out_degree() const { return _nV ? out_degree(max_out_degree_id()) : 0; }

    void writeBinary(const std::string& filename, bool /*print*/ = true) const {
        check(hn_gstd_write_binary(_g, filename.c_str(), (int32_t)sizeof(eoff_t)), "GraphStd::writeBinary");
    }
    void writeMarket(const std::string& filename, bool /*print*/ = true) const {
        check(hn_gstd_write_market(_g, filename.c_str()), "GraphStd::writeMarket");
    }

private:
    static void check(hn_status s, const char* where) {
        if (s != HN_OK) throw Error(s, where);
    }

    StructureProp _structure;
    hn_gstd* _g = nullptr;
    vid_t _nV = 0;
    eoff_t _nE = 0;
    uint32_t _st = 0;
    std::vector<eoff_t> _out_off, _in_off;
    const vid_t* _out_edges = nullptr;
    const vid_t* _in_edges = nullptr;
};

}  // namespace graph
