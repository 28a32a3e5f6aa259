// prims.hpp -- status-returning wrappers of the shared wave64 scan
// (include/hornet/detail/wave.hpp) for the library.
#pragma once
#include "common.hpp"

namespace hn {

inline hn_status scan_status(hipError_t e) {
    if (e == hipSuccess) return HN_OK;
    set_error(std::string("scan launch: ") + hipGetErrorString(e));
    return HN_E_HIP;
}

// Exclusive scan of in(0..n) delivered to out(i, prefix, value); *total = sum.
template <typename Out, typename InFn, typename OutFn>
hn_status exclusive_scan_fn(InFn in, OutFn out, Out* total, int64_t n, void* tmp, hipStream_t st) {
    return scan_status(exclusive_scan_dev<Out>(in, out, total, n, tmp, st));
}

// out[0..n) = exclusive prefix of in(0..n), out[n] = total.
template <typename Out, typename InFn>
hn_status exclusive_scan(InFn in, Out* out, int64_t n, void* tmp, hipStream_t st) {
    return scan_status(exclusive_scan_dev<Out>(in, ArrayOut<Out>{out}, out + n, n, tmp, st));
}

}  // namespace hn
