// hornet/operators.hpp -- the HornetsNest operator layer, header-only HIP.
//
// User functors cannot cross a C ABI, so the forAll* templates and their
// load-balancing kernels are instantiated in the caller's translation unit
// (compiled with hipcc for gfx950), exactly as in the reference, and work on
// the POD device view exported by libhornet_hip.  Reference counterparts:
//   forAll / forAllnumV / forAllVertices / forAllEdges / forAllEdgeVertexPairs
//                              primitives/Operator++.cuh:68-215, Operator++.i.cuh:524-645
//   OPERATOR                   primitives/Operator++.cuh:57-59
//   TwoLevelQueue<T>           primitives/Queue/TwoLevelQueue.cuh:62-204
//   load_balancing::BinarySearch  primitives/LoadBalancing/BinarySearch.cuh:52-94
//   load_balancing::VertexBased<VW> primitives/LoadBalancing/VertexBased.cuh:48-79
//   StaticAlgorithm<H>         primitives/HornetAlg.cuh:51-150
//   gpu::allocate/free/...     primitives/StandardAPI.hpp:43-155
// Differences: launches go to the graph's stream with no device-wide sync
// after each (the reference's CHECK_CUDA_ERROR syncs after every launch,
// xlib/include/Device/Util/SafeCudaAPI.cuh:71-79); the queue's device insert
// is wave64-aggregated (one atomicAdd per wave) and bounds-checked; errors
// throw hornet::Error instead of exiting.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <utility>

#include "hornet/hornet.hpp"
#include "hornet/detail/wave.hpp"

#define OPERATOR __device__ __forceinline__ void operator()

namespace hornets_nest {

using vid_t = int;
using vert_t = int;
using degree_t = int;
using eoff_t = int;
using Vertex = hornet::Vertex;
using Edge = hornet::Edge;

inline void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw hornet::Error(HN_E_HIP, where, hipGetErrorString(e));
}

// ------------------------------------------------------------------ memory
namespace gpu {
template <typename T>
void allocate(T*& ptr, size_t num_items) {
    ptr = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * (num_items ? num_items : 1)), "gpu::allocate");
}
template <typename T>
void free(T* ptr) {
    if (ptr) (void)hipFree(ptr);
}
template <typename T, typename... Ts>
void free(T* ptr, Ts*... rest) {
    free(ptr);
    free(rest...);
}
template <typename T>
void memsetZero(T* ptr, size_t num_items = 1) {
    hip_check(hipMemset(ptr, 0, sizeof(T) * num_items), "gpu::memsetZero");
}
template <typename T>
void memsetOne(T* ptr, size_t num_items = 1) {
    hip_check(hipMemset(ptr, 0xFF, sizeof(T) * num_items), "gpu::memsetOne");
}
template <typename T>
void copyToHost(const T* d_ptr, size_t num_items, T* h_ptr) {
    hip_check(hipMemcpy(h_ptr, d_ptr, sizeof(T) * num_items, hipMemcpyDeviceToHost), "gpu::copyToHost");
}
template <typename T>
void copyToDevice(const T* d_src, size_t num_items, T* d_dst) {
    hip_check(hipMemcpy(d_dst, d_src, sizeof(T) * num_items, hipMemcpyDeviceToDevice), "gpu::copyToDevice");
}
}  // namespace gpu

namespace host {
template <typename T>
void allocate(T*& ptr, size_t num_items) {
    ptr = new T[num_items ? num_items : 1];
}
template <typename T>
void free(T* ptr) {
    delete[] ptr;
}
template <typename T>
void copyToDevice(const T* h_ptr, size_t num_items, T* d_ptr) {
    hip_check(hipMemcpy(d_ptr, h_ptr, sizeof(T) * num_items, hipMemcpyHostToDevice), "host::copyToDevice");
}
template <typename T>
void copyToDevice(T value, T* d_ptr) {
    hip_check(hipMemcpy(d_ptr, &value, sizeof(T), hipMemcpyHostToDevice), "host::copyToDevice");
}
}  // namespace host

// ------------------------------------------------------------ TwoLevelQueue
// Ping-pong frontier: the operator appends to the output side on the device,
// swap() makes it the input (one 8-byte readback, as the reference's swap).
template <typename T>
class TwoLevelQueue {
public:
    TwoLevelQueue() = default;
    template <typename HornetClass>
    explicit TwoLevelQueue(const HornetClass& hornet, const float work_factor = 2.0f) {
        initialize((size_t)std::max<double>(1.0, (double)hornet.nV() * work_factor));
    }
    explicit TwoLevelQueue(size_t max_allocated_items) { initialize(max_allocated_items); }
    // kernel copies share the buffers and never free them
    __host__ __device__ TwoLevelQueue(const TwoLevelQueue<T>& o) noexcept
        : _in(o._in), _out(o._out), _cnt(o._cnt), _max(o._max), _in_size(o._in_size), _stream(o._stream),
          _kernel_copy(true) {}
    __host__ __device__ ~TwoLevelQueue() noexcept {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (!_kernel_copy) {
            if (_in) (void)hipFree(_in);
            if (_out) (void)hipFree(_out);
            if (_cnt) (void)hipFree(_cnt);
        }
#endif
    }
    TwoLevelQueue& operator=(const TwoLevelQueue&) = delete;

    void initialize(size_t max_allocated_items) {
        _max = (int)max_allocated_items;
        gpu::allocate(_in, max_allocated_items);
        gpu::allocate(_out, max_allocated_items);
        gpu::allocate(_cnt, 2);
        gpu::memsetZero(_cnt, 2);
    }

    // device: append from any lane (wave-aggregated, bounds-checked)
    __device__ __forceinline__ void insert(const T& item) const {
        const uint64_t m = __ballot(1);
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (__lane_id() == leader) base = atomicAdd(_cnt, (int)__popcll(m));
        base = __shfl(base, leader, hn::WAVE);
        const int pos = base + (int)__popcll(m & hn::lanemask_lt());
        if (pos < _max) _out[pos] = item;
        else atomicOr(_cnt + 1, 1);
    }
    // host: append to the input side
    void insert(const T& item) {
        if (_in_size >= _max) throw hornet::Error(HN_E_CAPACITY, "TwoLevelQueue::insert", "queue full");
        hip_check(hipMemcpy(_in + _in_size, &item, sizeof(T), hipMemcpyHostToDevice), "TwoLevelQueue::insert");
        _in_size++;
    }
    void insert(const T* items_array, int num_items) {
        if (_in_size + num_items > _max) throw hornet::Error(HN_E_CAPACITY, "TwoLevelQueue::insert", "queue full");
        hip_check(hipMemcpy(_in + _in_size, items_array, sizeof(T) * (size_t)num_items, hipMemcpyHostToDevice),
                  "TwoLevelQueue::insert");
        _in_size += num_items;
    }
    void swap() {
        int h[2] = {0, 0};
        hip_check(hipMemcpyAsync(h, _cnt, sizeof(h), hipMemcpyDeviceToHost, _stream), "TwoLevelQueue::swap");
        hip_check(hipStreamSynchronize(_stream), "TwoLevelQueue::swap");
        if (h[1]) throw hornet::Error(HN_E_CAPACITY, "TwoLevelQueue::swap", "device inserts overflowed the queue");
        std::swap(_in, _out);
        _in_size = h[0];
        hip_check(hipMemsetAsync(_cnt, 0, sizeof(int) * 2, _stream), "TwoLevelQueue::swap");
        _enqueued += h[0];
    }
    void clear() {
        _in_size = 0;
        hip_check(hipMemsetAsync(_cnt, 0, sizeof(int) * 2, _stream), "TwoLevelQueue::clear");
    }
    int size() const noexcept { return _in_size; }
    int output_size() const {
        int c = 0;
        hip_check(hipMemcpy(&c, _cnt, sizeof(int), hipMemcpyDeviceToHost), "TwoLevelQueue::output_size");
        return c;
    }
    const T* device_input_ptr() const noexcept { return _in; }
    const T* device_output_ptr() const noexcept { return _out; }
    int enqueue_items() const noexcept { return _enqueued; }
    void set_stream(hipStream_t s) noexcept { _stream = s; }
    hipStream_t stream() const noexcept { return _stream; }

private:
    T* _in = nullptr;
    T* _out = nullptr;
    int* _cnt = nullptr;   // [0] output size, [1] overflow flag
    int _max = 0;
    int _in_size = 0;
    hipStream_t _stream = nullptr;
    bool _kernel_copy = false;
    int _enqueued = 0;
};

// -------------------------------------------------------------- kernels
namespace detail {

constexpr int LB_IPT = 8;
constexpr int LB_TILE = hn::BLOCK * LB_IPT;

struct QueueDegIn {
    hornet::HornetDevice h;
    const vid_t* q;   // null -> identity (all vertices)
    __device__ __forceinline__ int64_t operator()(int64_t i) const {
        return h.degree(q ? q[i] : (vid_t)i);
    }
};

// op(Vertex&, Edge&) for every edge of the segment list (queue or all vertices)
template <typename Operator>
__global__ __launch_bounds__(hn::BLOCK) void binary_search_edges_kernel(hornet::HornetDevice h, const vid_t* q,
                                                                        int64_t F, const int64_t* P, Operator op) {
    __shared__ hn::LbShared<LB_IPT, int64_t> sh;
    const int64_t W = P[F];
    for (int64_t w0 = (int64_t)blockIdx.x * LB_TILE; w0 < W; w0 += (int64_t)gridDim.x * LB_TILE) {
        hn::lb_tile<LB_IPT>(
            P, F, w0, W, sh,
            [&](int64_t s) {
                const vid_t v = q ? q[s] : (vid_t)s;
                return h.records()[v].off - P[s];
            },
            [&](int64_t s, int64_t w, int64_t base) {
                const vid_t v = q ? q[s] : (vid_t)s;
                Vertex vx(h, v);
                Edge e(h, v, h.arena() + base + w);
                op(vx, e);
            });
    }
}

// VW lanes per vertex stride over its edges (VertexBasedKernel.cuh:45-129)
template <int VW, bool PAIRS, typename Operator>
__global__ __launch_bounds__(hn::BLOCK) void vertex_based_kernel(hornet::HornetDevice h, const vid_t* q, int64_t F,
                                                                 Operator op) {
    const int64_t gid = ((int64_t)blockIdx.x * hn::BLOCK + threadIdx.x) / VW;
    const int sub = threadIdx.x % VW;
    const int64_t stride = (int64_t)gridDim.x * hn::BLOCK / VW;
    for (int64_t i = gid; i < F; i += stride) {
        const vid_t v = q ? q[i] : (vid_t)i;
        Vertex vx(h, v);
        for (degree_t k = sub; k < vx.degree(); k += VW) {
            Edge e = vx.edge(k);
            if constexpr (PAIRS) {
                Vertex dst(h, e.dst_id());
                op(vx, dst);
            } else {
                op(vx, e);
            }
        }
    }
}

template <typename Operator>
__global__ void for_all_kernel(int64_t n, Operator op) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        op((int)i);
}

template <typename Operator>
__global__ void for_all_vertices_kernel(hornet::HornetDevice h, const vid_t* q, int64_t n, Operator op) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Vertex v(h, q ? q[i] : (vid_t)i);
        op(v);
    }
}

// adapter: op(src Vertex&, dst Vertex&) on top of an edge operator
template <typename Operator>
struct PairsAdapter {
    Operator op;
    hornet::HornetDevice h;
    __device__ __forceinline__ void operator()(Vertex& v, Edge& e) {
        Vertex d(h, e.dst_id());
        op(v, d);
    }
};

inline unsigned grid_for(int64_t n) {
    int64_t g = hn::ceil_div(n > 0 ? n : 1, hn::BLOCK);
    return (unsigned)std::min<int64_t>(g, 256 * 16);
}

}  // namespace detail

// --------------------------------------------------------- load balancing
namespace load_balancing {

// Edge-granular LB: scan of the segment degrees, then the wave64 LDS-staged
// tile mapping (BinarySearch.i.cuh:44-107 with xlib::binarySearchLB)
class BinarySearch {
public:
    template <typename HornetClass>
    explicit BinarySearch(HornetClass& hornet, const float work_factor = 2.0f) {
        reserve((size_t)std::max<double>((double)hornet.nV() * work_factor, (double)hornet.nV()) + 1);
    }
    BinarySearch(const BinarySearch&) = delete;
    BinarySearch& operator=(const BinarySearch&) = delete;
    ~BinarySearch() { release(); }

    template <typename HornetClass, typename Operator>
    void apply(HornetClass& hornet, const vid_t* d_input, int num_vertices, const Operator& op) {
        run(hornet, d_input, num_vertices, op);
    }
    template <typename HornetClass, typename Operator>
    void apply(HornetClass& hornet, const Operator& op) {
        run(hornet, nullptr, hornet.nV(), op);
    }

private:
    template <typename HornetClass, typename Operator>
    void run(HornetClass& hornet, const vid_t* q, int64_t F, const Operator& op) {
        if (F <= 0) return;
        reserve((size_t)F + 1);
        hipStream_t st = hornet.stream();
        hornet::HornetDevice h = hornet.device();
        hip_check(hn::exclusive_scan_dev<int64_t>(detail::QueueDegIn{h, q}, hn::ArrayOut<int64_t>{_work}, _work + F,
                                                  F, _tmp, st),
                  "BinarySearch::apply (scan)");
        detail::binary_search_edges_kernel<Operator><<<256 * 8, hn::BLOCK, 0, st>>>(h, q, F, _work, op);
        hip_check(hipGetLastError(), "BinarySearch::apply");
    }
    void reserve(size_t items) {
        if (items <= _cap) return;
        release();
        gpu::allocate(_work, items + 1);
        hip_check(hipMalloc(&_tmp, hn::scan_temp_bytes((int64_t)items)), "BinarySearch");
        _cap = items;
    }
    void release() {
        gpu::free(_work);
        if (_tmp) (void)hipFree(_tmp);
        _work = nullptr;
        _tmp = nullptr;
        _cap = 0;
    }
    int64_t* _work = nullptr;
    void* _tmp = nullptr;
    size_t _cap = 0;
};

// VW lanes per vertex (VertexBased.i.cuh:43-89); VW divides 64
template <int VW_SIZE>
class VertexBased {
    static_assert(VW_SIZE >= 1 && VW_SIZE <= 64 && (64 % VW_SIZE) == 0, "VW must divide the wave size");

public:
    template <typename HornetClass>
    explicit VertexBased(HornetClass&) noexcept {}

    template <typename HornetClass, typename Operator>
    void apply(HornetClass& hornet, const vid_t* d_input, int num_vertices, const Operator& op) const {
        launch<false>(hornet, d_input, num_vertices, op);
    }
    template <typename HornetClass, typename Operator>
    void apply(HornetClass& hornet, const Operator& op) const {
        launch<false>(hornet, nullptr, hornet.nV(), op);
    }
    template <typename HornetClass, typename Operator>
    void applyVertexPairs(HornetClass& hornet, const Operator& op) const {
        launch<true>(hornet, nullptr, hornet.nV(), op);
    }

private:
    template <bool PAIRS, typename HornetClass, typename Operator>
    void launch(HornetClass& hornet, const vid_t* q, int64_t F, const Operator& op) const {
        if (F <= 0) return;
        const unsigned grid = detail::grid_for(F * VW_SIZE);
        detail::vertex_based_kernel<VW_SIZE, PAIRS, Operator><<<grid, hn::BLOCK, 0, hornet.stream()>>>(
            hornet.device(), q, F, op);
        hip_check(hipGetLastError(), "VertexBased::apply");
    }
};
using VertexBased1 = VertexBased<1>;

}  // namespace load_balancing

// ---------------------------------------------------------------- forAll*
template <typename Operator>
void forAll(size_t size, const Operator& op, hipStream_t stream = nullptr) {
    if (size == 0) return;
    detail::for_all_kernel<Operator><<<detail::grid_for((int64_t)size), hn::BLOCK, 0, stream>>>((int64_t)size, op);
    hip_check(hipGetLastError(), "forAll");
}

template <typename HornetClass, typename Operator>
void forAllnumV(HornetClass& hornet, const Operator& op) {
    const int64_t n = hornet.nV();
    if (n == 0) return;
    detail::for_all_kernel<Operator><<<detail::grid_for(n), hn::BLOCK, 0, hornet.stream()>>>(n, op);
    hip_check(hipGetLastError(), "forAllnumV");
}

template <typename HornetClass, typename Operator>
void forAllVertices(HornetClass& hornet, const Operator& op) {
    const int64_t n = hornet.nV();
    if (n == 0) return;
    detail::for_all_vertices_kernel<Operator><<<detail::grid_for(n), hn::BLOCK, 0, hornet.stream()>>>(
        hornet.device(), nullptr, n, op);
    hip_check(hipGetLastError(), "forAllVertices");
}

template <typename HornetClass, typename Operator>
void forAllVertices(HornetClass& hornet, const vid_t* vertex_array, int size, const Operator& op) {
    if (size <= 0) return;
    detail::for_all_vertices_kernel<Operator><<<detail::grid_for(size), hn::BLOCK, 0, hornet.stream()>>>(
        hornet.device(), vertex_array, size, op);
    hip_check(hipGetLastError(), "forAllVertices");
}

template <typename HornetClass, typename Operator>
void forAllVertices(HornetClass& hornet, const TwoLevelQueue<vid_t>& queue, const Operator& op) {
    forAllVertices(hornet, queue.device_input_ptr(), queue.size(), op);
}

template <typename HornetClass, typename Operator, typename LoadBalancing>
void forAllEdges(HornetClass& hornet, const Operator& op, LoadBalancing& load_balancing) {
    load_balancing.apply(hornet, op);
}

template <typename HornetClass, typename Operator, typename LoadBalancing>
void forAllEdges(HornetClass& hornet, const vid_t* vertex_array, int size, const Operator& op,
                 LoadBalancing& load_balancing) {
    load_balancing.apply(hornet, vertex_array, size, op);
}

template <typename HornetClass, typename Operator, typename LoadBalancing>
void forAllEdges(HornetClass& hornet, const TwoLevelQueue<vid_t>& queue, const Operator& op,
                 LoadBalancing& load_balancing) {
    load_balancing.apply(hornet, queue.device_input_ptr(), queue.size(), op);
}

// op(Vertex& src, Vertex& dst) for every stored edge (Operator++.i.cuh:580-600)
template <typename HornetClass, typename Operator, typename LoadBalancing>
void forAllEdgeVertexPairs(HornetClass& hornet, const Operator& op, LoadBalancing& load_balancing) {
    load_balancing.apply(hornet, detail::PairsAdapter<Operator>{op, hornet.device()});
}

template <typename HornetClass, typename Operator, int VW>
void forAllEdgeVertexPairs(HornetClass& hornet, const Operator& op, load_balancing::VertexBased<VW>& lb) {
    lb.applyVertexPairs(hornet, op);
}

// --------------------------------------------------------- StaticAlgorithm
template <typename HornetClass>
class StaticAlgorithm {
public:
    explicit StaticAlgorithm(HornetClass& hornet) noexcept : hornet(hornet) {}
    virtual ~StaticAlgorithm() noexcept = default;
    virtual void run() = 0;
    virtual void reset() = 0;
    virtual void release() = 0;
    virtual bool validate() = 0;

protected:
    HornetClass& hornet;
};

}  // namespace hornets_nest
