// hornet/device.hpp -- device views of a graph, passed by value into kernels.
//
// Counterparts of the reference's HornetDevice / Vertex / Edge
// (hornet/include/Core/HornetDevice/HornetDevice.cuh:52-90,
//  impl/Vertex.i.cuh:59-146, impl/Edge.i.cuh:43-116).  Where the reference
// reads three SoA fields to build an edge pointer, a vertex here is one 16-byte
// record load {offset, degree, log2 capacity}; the adjacency is
// arena[offset .. offset+degree), sorted ascending.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hornet_hip.h"

namespace hornet {

using vid_t = int;
using degree_t = int;

class Vertex;
class Edge;

// POD view (hn_dview) with the reference's accessors
class HornetDevice {
public:
    using VertexType = vid_t;
    using DegreeType = degree_t;

    HornetDevice() = default;
    __host__ __device__ explicit HornetDevice(const hn_dview& v) : _nv(v.nv), _ne(v.ne), _rec(v.rec), _arena(v.arena) {
        for (int f = 0; f < HN_MAX_EDGE_META; f++) _meta[f] = v.meta[f];
    }

    __host__ __device__ __forceinline__ vid_t nV() const { return _nv; }
    __host__ __device__ __forceinline__ int64_t nE() const { return _ne; }
    __device__ __forceinline__ Vertex vertex(vid_t id) const;
    __device__ __forceinline__ degree_t degree(vid_t id) const { return _rec[id].deg; }
    __host__ __device__ __forceinline__ const hn_vrec* records() const { return _rec; }
    __host__ __device__ __forceinline__ vid_t* arena() const { return _arena; }
    // edge meta-data field f, parallel to the arena
    __host__ __device__ __forceinline__ void* edge_meta(int f) const { return _meta[f]; }

private:
    vid_t _nv = 0;
    int64_t _ne = 0;
    const hn_vrec* _rec = nullptr;
    vid_t* _arena = nullptr;
    void* _meta[HN_MAX_EDGE_META] = {};
};

class Vertex {
public:
    __device__ __forceinline__ Vertex(const HornetDevice& h, vid_t id) : _h(h), _id(id), _r(h.records()[id]) {}
    __device__ __forceinline__ Vertex(const HornetDevice& h, vid_t id, const hn_vrec& r) : _h(h), _id(id), _r(r) {}

    __device__ __forceinline__ vid_t id() const { return _id; }
    __device__ __forceinline__ degree_t degree() const { return _r.deg; }
    // capacity of the block holding the adjacency (Vertex::limit, Vertex.i.cuh:95-101)
    __device__ __forceinline__ degree_t limit() const { return _r.cap_log2 < 0 ? 0 : (degree_t(1) << _r.cap_log2); }
    __device__ __forceinline__ vid_t* neighbor_ptr() const { return _h.arena() + _r.off; }
    __device__ __forceinline__ Edge edge(degree_t i) const;

private:
    HornetDevice _h;
    vid_t _id;
    hn_vrec _r;
};

class Edge {
public:
    __device__ __forceinline__ Edge(const HornetDevice& h, vid_t src, vid_t* ptr) : _h(h), _src(src), _ptr(ptr) {}

    __device__ __forceinline__ vid_t src_id() const { return _src; }
    __device__ __forceinline__ vid_t dst_id() const { return *_ptr; }
    __device__ __forceinline__ Vertex src() const { return Vertex(_h, _src); }
    __device__ __forceinline__ Vertex dst() const { return Vertex(_h, *_ptr); }
    // edge meta-data field N (Edge::field<N>(), hornet/include/Core/HornetDevice/Edge.cuh:105-110);
    // T is the field's type (int by default)
    template <unsigned N, typename T = int>
    __device__ __forceinline__ T& field() const {
        return reinterpret_cast<T*>(_h.edge_meta((int)N))[_ptr - _h.arena()];
    }

private:
    HornetDevice _h;
    vid_t _src;
    vid_t* _ptr;
};

__device__ __forceinline__ Vertex HornetDevice::vertex(vid_t id) const { return Vertex(*this, id); }
__device__ __forceinline__ Edge Vertex::edge(degree_t i) const { return Edge(_h, _id, _h.arena() + _r.off + i); }

}  // namespace hornet
