// degree_order.hip -- the degree-sorted and degree-grouped (DBG) relabels.
// A vertex's key is its column count in M (how often it appears as a column
// index), or the group ceil(log2(count + 1)) of that count; the new order is
// (key descending, id ascending).  One 64-bit composite key per vertex,
// (~key << 32) | id, makes every key distinct, so one ascending radix sort of
// the composite keys yields that order exactly.
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "launch_degree.hpp"

namespace mr {
namespace k {

namespace {

__global__ void degree_keys_kernel(int64_t n, const int32_t* __restrict__ cnt, int grouped,
                                   unsigned long long* __restrict__ keys)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = cnt[v];
        // group = ceilf(log2(d + 1.0)), evaluated as the reference does
        // (src/class.cu:1683-1691): the double log2 is rounded to float first
        const uint32_t key = grouped ? (uint32_t)ceilf((float)log2((double)d + 1.0)) : (uint32_t)d;
        keys[v] = ((unsigned long long)(0xffffffffu - key) << 32) | (unsigned long long)(uint32_t)v;
    }
}

__global__ void rank_kernel(int64_t n, const unsigned long long* __restrict__ sorted, int32_t* __restrict__ perm)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        perm[(uint32_t)(sorted[i] & 0xffffffffull)] = (int32_t)i;
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
    size_t tmp = 0;
    rocprim::radix_sort_keys(nullptr, tmp, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                             (size_t)n, 0, 64, (hipStream_t)0);
    return tmp + 2 * sizeof(unsigned long long) * (size_t)n + 256;
}

hipError_t degree_order(int64_t n, const int32_t* cnt, int grouped, int32_t* perm, void* ws, size_t ws_bytes,
                        hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    unsigned long long* keys = static_cast<unsigned long long*>(ws);
    unsigned long long* sorted = keys + n;
    void* tmp = sorted + n;
    size_t tmp_bytes = ws_bytes - 2 * sizeof(unsigned long long) * (size_t)n;
    degree_keys_kernel<<<grid_of(n), 256, 0, s>>>(n, cnt, grouped, keys);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // bits 32..63 hold the inverted key, bits 0..31 the id; ids < 2^31
    e = rocprim::radix_sort_keys(tmp, tmp_bytes, keys, sorted, (size_t)n, 0, 64, s);
    if (e != hipSuccess) return e;
    rank_kernel<<<grid_of(n), 256, 0, s>>>(n, sorted, perm);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
