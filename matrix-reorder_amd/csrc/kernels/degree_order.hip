// degree_order.hip -- the degree-sorted and degree-grouped (DBG) relabels.
// A vertex's key is its column count in M (how often it appears as a column
// index), or the group ceil(log2(count + 1)) of that count; the new order is
// (key descending, id ascending).  The pairs (inverted key, id) start in id
// order, so one stable LSD radix sort (radix.hip) by the inverted key alone
// yields that order exactly.
#include <cstring>
#include <hip/hip_runtime.h>

#include "launch_degree.hpp"
#include "launch_radix.hpp"

namespace mr {
namespace k {

namespace {

// the largest key: column counts are below 2^31, groups at most 32
constexpr uint32_t kMaxCount = 0x7fffffffu;
constexpr uint32_t kMaxGroup = 63u;

__global__ void degree_keys_kernel(int64_t n, const int32_t* __restrict__ cnt, int grouped,
                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ ids)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = cnt[v];
        // group = ceilf(log2(d + 1.0)), evaluated as the reference does
        // (src/class.cu:1683-1691): the double log2 is rounded to float first
        const uint32_t key = grouped ? (uint32_t)ceilf((float)log2((double)d + 1.0)) : (uint32_t)d;
        keys[v] = (grouped ? kMaxGroup : kMaxCount) - key;   // ascending = key descending
        ids[v] = (uint32_t)v;
    }
}

__global__ void rank_kernel(int64_t n, const uint32_t* __restrict__ sorted_ids, int32_t* __restrict__ perm)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        perm[sorted_ids[i]] = (int32_t)i;
}

unsigned grid_of(int64_t n)
{
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

size_t degree_order_ws_bytes(int64_t n)
{
    return radix_sort_ws_bytes(n) + 4 * sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) + 256;
}

hipError_t degree_order(int64_t n, const int32_t* cnt, int grouped, int32_t* perm, void* ws, size_t ws_bytes,
                        hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const size_t arrays = 4 * sizeof(uint32_t) * (size_t)n;
    if (ws_bytes < arrays + radix_sort_ws_bytes(n)) return hipErrorInvalidValue;
    uint32_t* keys = static_cast<uint32_t*>(ws);
    uint32_t* ids = keys + n;
    uint32_t* keys_out = ids + n;
    uint32_t* ids_out = keys_out + n;
    void* tmp = static_cast<char*>(ws) + (arrays + 255) / 256 * 256;
    const size_t tmp_bytes = ws_bytes - (arrays + 255) / 256 * 256;
    degree_keys_kernel<<<grid_of(n), 256, 0, s>>>(n, cnt, grouped, keys, ids);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(keys, ids, keys_out, ids_out, n, grouped ? 6 : 31, tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    rank_kernel<<<grid_of(n), 256, 0, s>>>(n, ids_out, perm);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
