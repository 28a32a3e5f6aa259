// hornet/operators.hpp -- the HornetsNest operator layer, header-only HIP.
//
// User functors cannot cross a C ABI, so the forAll* templates and their
// load-balancing kernels are instantiated in the caller's translation unit
// (compiled with hipcc for gfx950), exactly as in the reference, and work on
// the POD device view exported by libhornet_hip.  Reference counterparts:
//   forAll / forAllnumV / forAllnumE / forAllVertices / forAllEdges /
//   forAllEdgeVertexPairs / forAllVertexPairs
//                              primitives/Operator++.cuh:68-215, Operator++.i.cuh:524-645
//   forAllAdjUnions            Operator++.cuh:171-180, Operator++.i.cuh:298-476
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
using vid2_t = int2;   // xlib::Make2Str<vid_t>::type

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

template <typename T, typename Operator>
__global__ void for_all_items_kernel(const T* __restrict__ items, int64_t n, Operator op) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T value = items[i];
        op(value);
    }
}

template <typename T, typename Operator>
__global__ void for_all_vertex_pairs_kernel(hornet::HornetDevice h, const T* __restrict__ pairs, int64_t n,
                                            Operator op) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T p = pairs[i];
        Vertex a(h, p.x), b(h, p.y);
        op(a, b);
    }
}

// ---- forAllAdjUnions
// One (src, dst) pair becomes an intersection of u = the endpoint of smaller
// degree (dst on ties) with v = the other, split over a power-of-two number of
// lanes (at most a wave) from its work: FLAG 1 (binary search) when
// WORK_FACTOR * (|u| + |v| + log|u|) >= |u| * log|v|, as the reference's bin
// rule (Operator++.i.cuh:312-320), else FLAG 0 (merge).  Pairs with src == dst
// or deg(src) < 2 are skipped (Operator++.i.cuh:212-214).
constexpr int ADJ_UNION_WORK = 64;   // merge steps / probes per lane

struct AdjUnionCase {
    vid_t u, v;
    degree_t ul, vl;
    int flag;
    int threads;   // 0: skipped
};

struct CooPairsAt {
    const vid_t* s;
    const vid_t* d;
    __device__ __forceinline__ int2 operator()(int64_t i) const { return make_int2(s[i], d[i]); }
};
template <typename T>
struct ArrayPairsAt {
    const T* p;
    __device__ __forceinline__ int2 operator()(int64_t i) const { return make_int2(p[i].x, p[i].y); }
};

template <typename Pairs>
__device__ __forceinline__ AdjUnionCase adj_union_case(const hornet::HornetDevice& h, const Pairs& pairs, int64_t i,
                                                       int work_factor) {
    const int2 p = pairs(i);
    const degree_t sl = h.degree(p.x), dl = h.degree(p.y);
    AdjUnionCase c{0, 0, 0, 0, 0, 0};
    if (p.x == p.y || sl < 2) return c;
    const bool ss = sl < dl;
    c.u = ss ? p.x : p.y;
    c.v = ss ? p.y : p.x;
    c.ul = ss ? sl : dl;
    c.vl = ss ? dl : sl;
    const int log_u = min(32 - __clz(c.ul), 31), log_v = min(32 - __clz(c.vl), 31);
    const int64_t bin_work = (int64_t)c.ul * log_v;
    c.flag = (int64_t)work_factor * ((int64_t)c.ul + c.vl + log_u) >= bin_work;
    const int64_t work = c.flag ? bin_work : (int64_t)c.ul + c.vl;
    int64_t t = (work + ADJ_UNION_WORK - 1) / ADJ_UNION_WORK;
    t = min(t, (int64_t)c.ul);
    int lanes = 1;
    while (lanes < t && lanes < hn::WAVE) lanes <<= 1;
    c.threads = t > 0 ? lanes : 0;
    return c;
}

template <typename Pairs>
struct AdjUnionThreadsIn {
    hornet::HornetDevice h;
    Pairs pairs;
    int work_factor;
    __device__ __forceinline__ int64_t operator()(int64_t i) const {
        return adj_union_case(h, pairs, i, work_factor).threads;
    }
};

// Merge-path split of sorted A (a) and B (b) after d merged elements, A first
// on ties: i from A, j = d - i from B.  A match A[i-1] == B[j] is moved into
// the earlier part (j + 1), so no equal pair is cut.
__device__ __forceinline__ void merge_split(const vid_t* A, int a, const vid_t* B, int b, int d, int& i, int& j) {
    int lo = d - b > 0 ? d - b : 0, hi = d < a ? d : a;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    i = lo;
    j = d - lo;
    if (i > 0 && j < b && A[i - 1] == B[j]) j++;
}

template <typename Pairs, typename Operator>
__global__ __launch_bounds__(hn::BLOCK) void adj_union_kernel(hornet::HornetDevice h, Pairs pairs, int64_t n,
                                                              const int64_t* P, int work_factor, Operator op) {
    __shared__ hn::LbShared<LB_IPT, int64_t> sh;
    const int64_t W = P[n];
    for (int64_t w0 = (int64_t)blockIdx.x * LB_TILE; w0 < W; w0 += (int64_t)gridDim.x * LB_TILE) {
        hn::lb_tile<LB_IPT>(
            P, n, w0, W, sh, [&](int64_t s) { return P[s]; },
            [&](int64_t s, int64_t w, int64_t base) {
                const AdjUnionCase c = adj_union_case(h, pairs, s, work_factor);
                const int k = (int)(w - base);
                Vertex uv(h, c.u), vv(h, c.v);
                vid_t* un = uv.neighbor_ptr();
                vid_t* vn = vv.neighbor_ptr();
                if (c.flag) {
                    const int b = (int)((int64_t)c.ul * k / c.threads);
                    const int e = (int)((int64_t)c.ul * (k + 1) / c.threads);
                    if (b < e) op(uv, vv, un + b, un + e - 1, vn, vn + c.vl - 1, 1);
                } else {
                    const int64_t tot = (int64_t)c.ul + c.vl;
                    int i0, j0, i1, j1;
                    merge_split(un, c.ul, vn, c.vl, (int)(tot * k / c.threads), i0, j0);
                    merge_split(un, c.ul, vn, c.vl, (int)(tot * (k + 1) / c.threads), i1, j1);
                    if (i0 < i1 && j0 < j1) op(uv, vv, un + i0, un + i1 - 1, vn + j0, vn + j1 - 1, 0);
                }
            });
    }
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

// op(item) for every item of the queue's input side (Operator++.i.cuh:534-543)
template <typename T, typename Operator>
void forAll(const TwoLevelQueue<T>& queue, const Operator& op) {
    const int64_t n = queue.size();
    if (n == 0) return;
    detail::for_all_items_kernel<T, Operator><<<detail::grid_for(n), hn::BLOCK, 0, queue.stream()>>>(
        queue.device_input_ptr(), n, op);
    hip_check(hipGetLastError(), "forAll");
}

// op(Vertex&, Vertex&) for every (x, y) pair of the queue (Operator++.i.cuh:545-556)
template <typename HornetClass, typename T, typename Operator>
void forAllVertexPairs(HornetClass& hornet, const TwoLevelQueue<T>& queue, const Operator& op) {
    const int64_t n = queue.size();
    if (n == 0) return;
    detail::for_all_vertex_pairs_kernel<T, Operator><<<detail::grid_for(n), hn::BLOCK, 0, hornet.stream()>>>(
        hornet.device(), queue.device_input_ptr(), n, op);
    hip_check(hipGetLastError(), "forAllVertexPairs");
}

// op(i) for i in [0, nE) (Operator++.i.cuh:570-577)
template <typename HornetClass, typename Operator>
void forAllnumE(HornetClass& hornet, const Operator& op) {
    const int64_t n = hornet.nE64();
    if (n == 0) return;
    detail::for_all_kernel<Operator><<<detail::grid_for(n), hn::BLOCK, 0, hornet.stream()>>>(n, op);
    hip_check(hipGetLastError(), "forAllnumE");
}

// forAllEdges without a load-balancing object (declared at Operator++.cuh:129-130):
// BinarySearch
template <typename HornetClass, typename Operator>
void forAllEdges(HornetClass& hornet, const Operator& op) {
    load_balancing::BinarySearch lb(hornet, 1.0f);
    lb.apply(hornet, op);
    hip_check(hipStreamSynchronize(hornet.stream()), "forAllEdges");
}

namespace detail {
template <typename HornetClass, typename Pairs, typename Operator>
void adj_unions_run(HornetClass& hornet, Pairs pairs, int64_t n, const Operator& op, int work_factor) {
    if (n <= 0) return;
    hipStream_t st = hornet.stream();
    hornet::HornetDevice h = hornet.device();
    int64_t* P = nullptr;
    void* tmp = nullptr;
    gpu::allocate(P, (size_t)n + 1);
    hip_check(hipMalloc(&tmp, hn::scan_temp_bytes(n)), "forAllAdjUnions");
    hip_check(hn::exclusive_scan_dev<int64_t>(AdjUnionThreadsIn<Pairs>{h, pairs, work_factor},
                                              hn::ArrayOut<int64_t>{P}, P + n, n, tmp, st),
              "forAllAdjUnions (scan)");
    adj_union_kernel<Pairs, Operator><<<256 * 8, hn::BLOCK, 0, st>>>(h, pairs, n, P, work_factor, op);
    hip_check(hipGetLastError(), "forAllAdjUnions");
    hip_check(hipStreamSynchronize(st), "forAllAdjUnions");
    gpu::free(P);
    (void)hipFree(tmp);
}
}  // namespace detail

// op(Vertex& u, Vertex& v, vid_t* ui_begin, vid_t* ui_end, vid_t* vi_begin,
// vid_t* vi_end, int FLAG) over every stored edge (Operator++.i.cuh:369-375)
// or the queue's vertex pairs (:377-476; an empty queue means every edge).
// Ends are inclusive.  FLAG 0: merge the two sub-ranges; the lanes' sub-ranges
// partition the merge path so each common element is seen by exactly one
// lane.  FLAG 1: binary-search each element of the u sub-range in all of v.
template <typename HornetClass, typename Operator>
void forAllAdjUnions(HornetClass& hornet, const Operator& op, const int WORK_FACTOR = 1) {
    auto coo = hornet.getCOO();
    detail::adj_unions_run(hornet, detail::CooPairsAt{coo.srcPtr(), coo.dstPtr()}, coo.size(), op, WORK_FACTOR);
}

template <typename HornetClass, typename Operator>
void forAllAdjUnions(HornetClass& hornet, TwoLevelQueue<vid2_t> vertex_pairs, const Operator& op,
                     const int WORK_FACTOR = 1) {
    if (vertex_pairs.size() == 0) {
        forAllAdjUnions(hornet, op, WORK_FACTOR);
        return;
    }
    detail::adj_unions_run(hornet, detail::ArrayPairsAt<vid2_t>{vertex_pairs.device_input_ptr()},
                           vertex_pairs.size(), op, WORK_FACTOR);
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
