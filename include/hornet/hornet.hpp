// hornet/hornet.hpp -- the dynamic graph object and its update batches.
//
// Header-only C++ over the C ABI of libhornet_hip.so (hornet_hip.h), with the
// names, arguments and defaults of the reference:
//   hornet::HornetInit<vid_t>        hornet/include/Core/HornetInitialize/HornetInit.cuh:54-97
//   hornet::gpu::Hornet<vid_t>       hornet/include/Core/Hornet.cuh:54-131
//   hornet::gpu::HornetStatic<vid_t> hornet/include/Core/Static/HornetStatic.cuh:16-66
//   hornet::BatchUpdatePtr / gpu::BatchUpdate
//                                    hornet/include/Core/BatchUpdate/BatchUpdate.cuh:77-272
//   hornet::COO                      hornet/include/Core/Static/Static.cuh:19-130
// vid_t = degree_t = int; vertex meta types are not supported (EMPTY); edge
// meta types are a TypeList of up to four 4- or 8-byte fields
// (hornet/include/Core/Conf/HornetConf.cuh), read on the device with
// edge.template field<N, T>() (T defaults to int).
// Errors: the reference prints and exits (xlib/src/Device/Util/SafeCudaAPI.cpp:71-94);
// here a failed call throws hornet::Error carrying the hn_status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "hornet_hip.h"
#include "hornet/device.hpp"

namespace hornet {

struct EMPTY {};
enum class DeviceType { HOST, DEVICE };

class Error : public std::runtime_error {
public:
    // detail defaults to the library's message for the failed call
    Error(hn_status s, const std::string& where, const std::string& detail = std::string())
        : std::runtime_error(where + ": " + hn_status_string(s) + ": " + (detail.empty() ? hn_last_error() : detail)),
          status(s) {}
    hn_status status;
};

inline void check(hn_status s, const char* where) {
    if (s != HN_OK) throw Error(s, where);
}

// Edge meta-data types (hornet/include/Core/Conf/HornetConf.cuh TypeList):
// up to HN_MAX_EDGE_META fields of 4 or 8 bytes, stored beside the edges
template <typename... Ts>
struct TypeList {};

namespace detail {
template <typename L>
struct MetaTraits;
template <>
struct MetaTraits<EMPTY> {
    static constexpr int n = 0;
    static void layout(hn_edge_meta& m) { m = hn_edge_meta{}; }
};
template <typename... Ts>
struct MetaTraits<TypeList<Ts...>> {
    static constexpr int n = (int)sizeof...(Ts);
    static_assert(n <= HN_MAX_EDGE_META, "at most HN_MAX_EDGE_META edge meta-data fields");
    static_assert(((sizeof(Ts) == 4 || sizeof(Ts) == 8) && ...), "edge meta-data fields are 4 or 8 bytes");
    static void layout(hn_edge_meta& m) {
        m = hn_edge_meta{};
        m.nfields = n;
        int i = 0;
        ((m.bytes[i++] = (int32_t)sizeof(Ts)), ...);
    }
};
template <unsigned N, typename L>
struct FieldAt;
template <unsigned N, typename T, typename... Ts>
struct FieldAt<N, TypeList<T, Ts...>> : FieldAt<N - 1, TypeList<Ts...>> {};
template <typename T, typename... Ts>
struct FieldAt<0, TypeList<T, Ts...>> {
    using type = T;
};
}  // namespace detail

// Host CSR handed to the Hornet constructor (pointers are borrowed), with the
// edge meta-data of HornetInit::insertEdgeData (HornetInit.cuh:75-97)
template <typename VID = vid_t, typename VMeta = EMPTY, typename EMeta = EMPTY, typename DEG = degree_t>
class HornetInit {
    static_assert(std::is_same<VID, int>::value && std::is_same<DEG, int>::value, "int ids and degrees");
    static_assert(std::is_same<VMeta, EMPTY>::value, "no vertex meta data");

public:
    static constexpr int num_edge_meta = detail::MetaTraits<EMeta>::n;

    HornetInit(VID nV, DEG nE, const DEG* csr_offsets, const VID* csr_edges) noexcept
        : _nV(nV), _nE(nE), _off(csr_offsets), _edges(csr_edges) {}
    template <typename... Ps>
    void insertEdgeData(const Ps*... edge_meta_data) noexcept {
        static_assert(sizeof...(Ps) == (size_t)num_edge_meta, "one array per edge meta-data field");
        int i = 0;
        ((_meta[i++] = static_cast<const void*>(edge_meta_data)), ...);
    }
    // one field at a time (HornetInit.cuh:78-79); fields not given read as zero
    template <unsigned N>
    void insertEdgeData(const typename detail::FieldAt<N, EMeta>::type* edge_meta_data) noexcept {
        static_assert(N < (unsigned)num_edge_meta, "no such edge meta-data field");
        _meta[N] = static_cast<const void*>(edge_meta_data);
    }
    VID nV() const noexcept { return _nV; }
    DEG nE() const noexcept { return _nE; }
    const DEG* csr_offsets() const noexcept { return _off; }
    const VID* csr_edges() const noexcept { return _edges; }
    const void* const* edge_meta() const noexcept { return _meta; }

private:
    VID _nV;
    DEG _nE;
    const DEG* _off;
    const VID* _edges;
    const void* _meta[HN_MAX_EDGE_META] = {};
};

// (n, src, dst, meta...) batch on the host or the device (BatchUpdate.cuh:95-99)
template <typename VID = vid_t, typename EMeta = EMPTY, DeviceType device = DeviceType::DEVICE,
          typename DEG = degree_t>
class BatchUpdatePtr {
public:
    template <typename... Ps>
    BatchUpdatePtr(DEG n, VID* src, VID* dst, Ps*... edge_data) noexcept : _n(n), _src(src), _dst(dst) {
        static_assert(sizeof...(Ps) == 0 || sizeof...(Ps) == (size_t)detail::MetaTraits<EMeta>::n,
                      "one array per edge meta-data field");
        int i = 0;
        ((_meta[i++] = static_cast<const void*>(edge_data)), ...);
    }
    DEG nE() const noexcept { return _n; }
    VID* src() const noexcept { return _src; }
    VID* dst() const noexcept { return _dst; }
    const void* const* edge_meta() const noexcept { return _meta[0] ? _meta : nullptr; }
    static constexpr bool on_device = device == DeviceType::DEVICE;

private:
    DEG _n;
    VID* _src;
    VID* _dst;
    const void* _meta[HN_MAX_EDGE_META] = {};
};

// Flat (src, dst, meta...) edge list; device storage owned by the object
template <DeviceType device = DeviceType::DEVICE, typename VID = vid_t, typename EMeta = EMPTY,
          typename DEG = degree_t>
class COO {
    static constexpr int NM = detail::MetaTraits<EMeta>::n;

public:
    COO() = default;
    explicit COO(int64_t n) : _n(n) {
        if (n > 0) {
            check(hipMalloc(&_src, sizeof(VID) * (size_t)n) == hipSuccess ? HN_OK : HN_E_OOM, "COO");
            check(hipMalloc(&_dst, sizeof(VID) * (size_t)n) == hipSuccess ? HN_OK : HN_E_OOM, "COO");
            hn_edge_meta m;
            detail::MetaTraits<EMeta>::layout(m);
            for (int f = 0; f < NM; f++)
                check(hipMalloc(&_meta[f], (size_t)m.bytes[f] * (size_t)n) == hipSuccess ? HN_OK : HN_E_OOM, "COO");
        }
    }
    COO(const COO&) = delete;
    COO& operator=(const COO&) = delete;
    COO(COO&& o) noexcept : _n(o._n), _src(o._src), _dst(o._dst) {
        for (int f = 0; f < HN_MAX_EDGE_META; f++) _meta[f] = o._meta[f], o._meta[f] = nullptr;
        o._src = o._dst = nullptr;
        o._n = 0;
    }
    ~COO() {
        if (_src) (void)hipFree(_src);
        if (_dst) (void)hipFree(_dst);
        for (int f = 0; f < HN_MAX_EDGE_META; f++)
            if (_meta[f]) (void)hipFree(_meta[f]);
    }
    // device array of edge meta-data field N
    template <unsigned N>
    typename detail::FieldAt<N, EMeta>::type* edgeMetaPtr() const noexcept {
        return static_cast<typename detail::FieldAt<N, EMeta>::type*>(_meta[N]);
    }
    void* const* edge_meta() const noexcept { return NM ? _meta : nullptr; }
    template <unsigned N>
    void edgeMetaToHost(std::vector<typename detail::FieldAt<N, EMeta>::type>& out) const {
        out.resize((size_t)_n);
        if (_n > 0)
            (void)hipMemcpy(out.data(), _meta[N], sizeof(out[0]) * (size_t)_n, hipMemcpyDeviceToHost);
    }
    int64_t size() const noexcept { return _n; }
    VID* srcPtr() const noexcept { return _src; }
    VID* dstPtr() const noexcept { return _dst; }
    // copy to host vectors
    void toHost(std::vector<VID>& s, std::vector<VID>& d) const {
        s.resize((size_t)_n);
        d.resize((size_t)_n);
        if (_n > 0) {
            (void)hipMemcpy(s.data(), _src, sizeof(VID) * (size_t)_n, hipMemcpyDeviceToHost);
            (void)hipMemcpy(d.data(), _dst, sizeof(VID) * (size_t)_n, hipMemcpyDeviceToHost);
        }
    }

private:
    int64_t _n = 0;
    VID* _src = nullptr;
    VID* _dst = nullptr;
    void* _meta[HN_MAX_EDGE_META] = {};
};

namespace detail {
template <typename DEG>
__global__ void narrow_offsets_kernel(int64_t n, const int64_t* __restrict__ in, DEG* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (DEG)in[i];
}
}  // namespace detail

// CSR (hornet/include/Core/Static/Static.cuh:92-129): offsets[nV+1] and the
// index (destinations and edge meta-data) in device memory owned by the
// object.  offset() holds degree_t offsets as in the reference (valid while
// nE fits a degree_t); offset64() holds the same offsets as int64.
template <DeviceType device = DeviceType::DEVICE, typename VID = vid_t, typename EMeta = EMPTY,
          typename DEG = degree_t>
class CSR {
    static_assert(device == DeviceType::DEVICE, "getCSR returns device storage");
    static constexpr int NM = detail::MetaTraits<EMeta>::n;

public:
    CSR() = default;
    CSR(int64_t nv, int64_t ne) : _nv(nv), _ne(ne), _index(ne) {
        check(hipMalloc(&_off64, sizeof(int64_t) * (size_t)(nv + 1)) == hipSuccess ? HN_OK : HN_E_OOM, "CSR");
        check(hipMalloc(&_off, sizeof(DEG) * (size_t)(nv + 1)) == hipSuccess ? HN_OK : HN_E_OOM, "CSR");
    }
    CSR(const CSR&) = delete;
    CSR& operator=(const CSR&) = delete;
    CSR(CSR&& o) noexcept : _nv(o._nv), _ne(o._ne), _off(o._off), _off64(o._off64), _index(std::move(o._index)) {
        o._off = nullptr;
        o._off64 = nullptr;
        o._nv = o._ne = 0;
    }
    ~CSR() {
        if (_off) (void)hipFree(_off);
        if (_off64) (void)hipFree(_off64);
    }
    DEG* offset() noexcept { return _off; }
    int64_t* offset64() noexcept { return _off64; }
    VID* index() noexcept { return _index.dstPtr(); }
    template <unsigned N>
    typename detail::FieldAt<N, EMeta>::type* edgeMetaPtr() const noexcept {
        return _index.template edgeMetaPtr<N>();
    }
    int64_t nV() const noexcept { return _nv; }
    int64_t nE() const noexcept { return _ne; }
    // copy to host vectors (int64 offsets)
    void toHost(std::vector<int64_t>& off, std::vector<VID>& adj) const {
        off.resize((size_t)_nv + 1);
        adj.resize((size_t)_ne);
        if (_off64) (void)hipMemcpy(off.data(), _off64, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost);
        if (_ne > 0) (void)hipMemcpy(adj.data(), _index.dstPtr(), sizeof(VID) * (size_t)_ne, hipMemcpyDeviceToHost);
    }
    // the index as a COO whose source array is unused (destinations + meta-data)
    COO<device, VID, EMeta, DEG>& index_storage() noexcept { return _index; }

private:
    int64_t _nv = 0;
    int64_t _ne = 0;
    DEG* _off = nullptr;
    int64_t* _off64 = nullptr;
    COO<device, VID, EMeta, DEG> _index;
};

namespace gpu {

// A batch as seen by insert/erase: the library sorts and deduplicates it in
// its own scratch, so the caller's arrays are only read
template <typename VID = vid_t, typename EMeta = EMPTY, typename DEG = degree_t>
class BatchUpdate {
public:
    template <DeviceType D>
    explicit BatchUpdate(const BatchUpdatePtr<VID, EMeta, D, DEG>& p) noexcept
        : _n(p.nE()), _src(p.src()), _dst(p.dst()), _meta(p.edge_meta()), _device(D == DeviceType::DEVICE) {}
    template <DeviceType D>
    explicit BatchUpdate(COO<D, VID, EMeta, DEG>& coo) noexcept
        : _n((DEG)coo.size()), _src(coo.srcPtr()), _dst(coo.dstPtr()),
          _meta(const_cast<const void* const*>(coo.edge_meta())), _device(D == DeviceType::DEVICE) {}

    DEG nE() const noexcept { return _n; }
    DEG size() const noexcept { return _n; }
    const VID* src() const noexcept { return _src; }
    const VID* dst() const noexcept { return _dst; }
    // edge meta-data arrays, one per field (nullptr: none given)
    const void* const* edge_meta() const noexcept { return _meta; }
    uint32_t flags() const noexcept { return _device ? HN_DEVICE_PTRS : 0u; }
    // entries that took effect in the last insert / erase
    int64_t applied() const noexcept { return _applied; }
    void set_applied(int64_t a) noexcept { _applied = a; }

private:
    DEG _n;
    const VID* _src;
    const VID* _dst;
    const void* const* _meta;
    bool _device;
    int64_t _applied = 0;
};

// HornetStatic (hornet/include/Core/Static/HornetStatic.cuh:16-66): an
// immutable graph with Hornet's queries and device view; Hornet adds updates.
template <typename VID = vid_t, typename VMeta = EMPTY, typename EMeta = EMPTY, typename DEG = degree_t>
class HornetStatic {
    static_assert(std::is_same<VID, int>::value && std::is_same<DEG, int>::value, "int ids and degrees");
    static_assert(std::is_same<VMeta, EMPTY>::value, "no vertex meta data");

public:
    using VertexType = VID;
    using DegreeType = DEG;
    using HornetDeviceT = HornetDevice;
    using HInitT = HornetInit<VID, VMeta, EMeta, DEG>;

    // HornetStatic(HInitT&): no room for growth is reserved
    explicit HornetStatic(HInitT& init, hipStream_t stream = nullptr) : HornetStatic(init, 1.0, stream) {}
    // adopt a handle created through the C ABI
    explicit HornetStatic(hn_graph* g, hipStream_t stream = nullptr) noexcept : _g(g), _stream(stream) {}
    HornetStatic(const HornetStatic&) = delete;
    HornetStatic& operator=(const HornetStatic&) = delete;
    HornetStatic(HornetStatic&& o) noexcept : _g(o._g), _stream(o._stream) { o._g = nullptr; }
    ~HornetStatic() { hn_graph_destroy(_g); }

    DEG nV() const {
        int32_t v = 0;
        check(hn_graph_nv(_g, &v), "Hornet::nV");
        return v;
    }
    DEG nE() const {
        int64_t e = 0;
        check(hn_graph_ne(_g, &e), "Hornet::nE");
        return (DEG)e;
    }
    int64_t nE64() const {
        int64_t e = 0;
        check(hn_graph_ne(_g, &e), "Hornet::nE");
        return e;
    }
    HornetDevice device() const {
        hn_dview v;
        check(hn_graph_device_view(_g, &v), "Hornet::device");
        return HornetDevice(v);
    }
    // HornetQuery.i.cuh:10-36 (lowest id on ties)
    VID max_degree_id() const {
        int32_t id = 0, d = 0;
        check(hn_graph_max_degree(_g, &id, &d, _stream), "Hornet::max_degree_id");
        return id;
    }
    DEG max_degree() const {
        int32_t id = 0, d = 0;
        check(hn_graph_max_degree(_g, &id, &d, _stream), "Hornet::max_degree");
        return d;
    }
    // getCOO (HornetQuery.i.cuh:111-151): adjacency is kept sorted, so the
    // result is always sorted by (src, dst)
    COO<DeviceType::DEVICE, VID, EMeta, DEG> getCOO(bool sortAdjacencyList = true) const {
        (void)sortAdjacencyList;
        int64_t ne = nE64();
        COO<DeviceType::DEVICE, VID, EMeta, DEG> coo(ne);
        int64_t n = 0;
        if (ne > 0)
            check(hn_graph_export_coo_meta(_g, coo.srcPtr(), coo.dstPtr(), coo.edge_meta(), ne, &n, _stream),
                  "Hornet::getCOO");
        return coo;
    }
    // getCSR (HornetQuery.i.cuh:69-107; the reference's flatten leaves the
    // index unsorted unless asked): rows are always sorted here
    CSR<DeviceType::DEVICE, VID, EMeta, DEG> getCSR(bool sortAdjacencyList = false) const {
        (void)sortAdjacencyList;
        const int64_t nv = nV(), ne = nE64();
        CSR<DeviceType::DEVICE, VID, EMeta, DEG> csr(nv, ne);
        check(hn_graph_export_csr_meta(_g, csr.offset64(), csr.index(), csr.index_storage().edge_meta(), ne,
                                       _stream),
              "Hornet::getCSR");
        const int64_t n = nv + 1;
        detail::narrow_offsets_kernel<DEG><<<(unsigned)((n + 255) / 256), 256, 0, _stream>>>(n, csr.offset64(),
                                                                                           csr.offset());
        check(hipStreamSynchronize(_stream) == hipSuccess ? HN_OK : HN_E_HIP, "Hornet::getCSR");
        return csr;
    }
    // print (HornetInitialize.i.cuh:152-170): "v : degree | neighbours" per vertex
    void print() const {
        std::vector<int64_t> off;
        std::vector<VID> adj;
        getCSR().toHost(off, adj);
        for (int64_t v = 0; v + 1 < (int64_t)off.size(); v++) {
            std::printf("%lld : %lld | ", (long long)v, (long long)(off[v + 1] - off[v]));
            for (int64_t j = off[v]; j < off[v + 1]; j++) std::printf("%d ", (int)adj[j]);
            std::printf("\n");
        }
    }
    hn_graph* handle() const noexcept { return _g; }
    hipStream_t stream() const noexcept { return _stream; }

protected:
    HornetStatic(HInitT& init, double arena_slack, hipStream_t stream) : _stream(stream) {
        hn_edge_meta layout;
        detail::MetaTraits<EMeta>::layout(layout);
        check(hn_graph_create_csr_meta(init.nV(), init.nE(), init.csr_offsets(), init.csr_edges(), &layout,
                                       init.edge_meta(), HN_OFFSETS_INT32, arena_slack, stream, &_g),
              "Hornet(HornetInit&)");
    }

    hn_graph* _g = nullptr;
    hipStream_t _stream = nullptr;
};

template <typename VID = vid_t, typename VMeta = EMPTY, typename EMeta = EMPTY, typename DEG = degree_t>
class Hornet : public HornetStatic<VID, VMeta, EMeta, DEG> {
    using Base = HornetStatic<VID, VMeta, EMeta, DEG>;
    using Base::_g;
    using Base::_stream;

public:
    using typename Base::HInitT;

    // Hornet(void) / Hornet(degree_t nV) (HornetInitialize.i.cuh:72-84): a
    // graph of nV vertices and no edges, which inserts then fill
    Hornet() : Hornet(DEG(0)) {}
    explicit Hornet(DEG nV, hipStream_t stream = nullptr) : Base(empty_graph(nV, stream), stream) {}
    // Hornet(HornetInit&) (HornetInitialize.i.cuh:86-146); `arena_slack` sizes
    // the edge arena as a multiple of the blocks the input needs
    explicit Hornet(HInitT& init, double arena_slack = 0.0, hipStream_t stream = nullptr)
        : Base(init, arena_slack, stream) {}
    // reset(HornetInit&) (HornetInitialize.i.cuh:188-204): the graph becomes
    // the one the init describes; on failure it is left unchanged
    void reset(HInitT& init, double arena_slack = 0.0) {
        Hornet fresh(init, arena_slack, _stream);
        std::swap(_g, fresh._g);
    }
    // adopt a handle created through the C ABI
    explicit Hornet(hn_graph* g, hipStream_t stream = nullptr) noexcept : Base(g, stream) {}

    // Hornet::insert (HornetInsert.i.cuh:12-27)
    void insert(BatchUpdate<VID, EMeta, DEG>& batch, bool removeBatchDuplicates = false,
                bool removeGraphDuplicates = false) {
        uint32_t f = batch.flags() | (removeBatchDuplicates ? HN_DEDUP_BATCH : 0u) |
                     (removeGraphDuplicates ? HN_DEDUP_GRAPH : 0u);
        int64_t n = 0;
        check(hn_graph_insert_meta(_g, batch.src(), batch.dst(), batch.edge_meta(), batch.nE(), f, &n, _stream),
              "Hornet::insert");
        batch.set_applied(n);
    }
    // Hornet::erase (HornetInsert.i.cuh:71-89)
    void erase(BatchUpdate<VID, EMeta, DEG>& batch, bool removeBatchDuplicates = false) {
        uint32_t f = batch.flags() | (removeBatchDuplicates ? HN_DEDUP_BATCH : 0u);
        int64_t n = 0;
        check(hn_graph_erase(_g, batch.src(), batch.dst(), batch.nE(), f, &n, _stream), "Hornet::erase");
        batch.set_applied(n);
    }
    // a frozen copy of the graph as it is now (hn_graph_snapshot; no reference
    // counterpart): later updates of this graph do not reach it
    Base snapshot(double arena_slack = 1.0) const {
        hn_graph* c = nullptr;
        check(hn_graph_snapshot(_g, arena_slack, _stream, &c), "Hornet::snapshot");
        return Base(c, _stream);
    }

private:
    static hn_graph* empty_graph(DEG nV, hipStream_t stream) {
        if (nV < 0) throw Error(HN_E_INVALID_ARG, "Hornet(nV)", "negative vertex count");
        std::vector<int64_t> off((size_t)nV + 1, 0);
        hn_edge_meta layout;
        detail::MetaTraits<EMeta>::layout(layout);
        const void* none[HN_MAX_EDGE_META] = {};
        hn_graph* g = nullptr;
        check(hn_graph_create_csr_meta(nV, 0, off.data(), nullptr, &layout, none, 0, 0.0, stream, &g),
              "Hornet(nV)");
        return g;
    }
};

}  // namespace gpu

// HornetStatic.cuh:76-80
template <typename>
struct IsHornetStatic : std::false_type {};
template <typename V, typename VM, typename EM, typename D>
struct IsHornetStatic<gpu::HornetStatic<V, VM, EM, D>> : std::true_type {};

}  // namespace hornet
