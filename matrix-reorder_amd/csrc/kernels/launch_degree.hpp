// launch_degree.hpp -- host launchers of the degree-sorted and DBG relabels
// (degree_order.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace mr {
namespace k {

// workspace of degree_order for n vertices
size_t degree_order_ws_bytes(int64_t n);
// perm[v] = rank of v in (key desc, id asc) with key = cnt[v] (grouped = 0)
// or ceil(log2(cnt[v] + 1)) (grouped = 1)
hipError_t degree_order(int64_t n, const int32_t* cnt, int grouped, int32_t* perm, void* ws, size_t ws_bytes,
                        hipStream_t s);

}  // namespace k
}  // namespace mr
