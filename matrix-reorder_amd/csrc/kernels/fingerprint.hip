// fingerprint.hip -- position-keyed 64-bit fingerprint of an int32 array (the
// key of the on-disk plan cache).  One grid-stride pass; each wave adds its
// partial sum to the result with one 64-bit integer atomic, so the result is
// exact and independent of the schedule.
#include <hip/hip_runtime.h>

#include "launch_cache.hpp"

namespace mr {
namespace k {

namespace {

__global__ __launch_bounds__(256) void fingerprint_kernel(const int32_t* __restrict__ a, int64_t n, uint64_t salt,
                                                          unsigned long long* out)
{
    uint64_t acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += fingerprint_term((uint64_t)i, (uint32_t)a[i], salt);
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor((unsigned long long)acc, off);
    if ((threadIdx.x & 63) == 0 && acc != 0) atomicAdd(out, (unsigned long long)acc);
}

}  // namespace

hipError_t fingerprint_sum(const int32_t* a, int64_t n, uint64_t salt, uint64_t* out, int cu_count,
                           hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    const int64_t cap = (int64_t)(cu_count > 0 ? cu_count : 256) * 8;
    if (g > cap) g = cap;
    fingerprint_kernel<<<(unsigned)g, 256, 0, s>>>(a, n, salt, reinterpret_cast<unsigned long long*>(out));
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
