// serve.hip -- device helpers of the serving loop: right-hand sides drawn from
// a per-request seed, written straight into a batch's X arena, and an
// order-independent 64-bit hash of each request's result.
//
// Both kernels walk the concatenation of all requests of a batch (begin[r] =
// first element of request r) with a binary search over the begin array held
// in LDS, like the batched MERBIT kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch_serve.hpp"

namespace mr {
namespace k {
namespace {

constexpr int kThreads = 256;
constexpr int kGrid = 2048;

__device__ __forceinline__ int find_req(const int64_t* s_begin, int n, int64_t e)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_begin[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kThreads) void rhs_kernel(const ServeReq* __restrict__ reqs,
                                                       const int64_t* __restrict__ begin, int n, int64_t total)
{
    __shared__ int64_t s_begin[kMaxServeReqs + 1];
    for (int i = threadIdx.x; i <= n; i += kThreads) s_begin[i] = begin[i];
    __syncthreads();
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kThreads) {
        const int r = find_req(s_begin, n, e);
        const int64_t j = e - s_begin[r];   // request-local index i * k + c
        const int k = reqs[r].k;
        const int64_t i = j / k, c = j - i * k;
        float* x = static_cast<float*>(reqs[r].x);
        x[i * reqs[r].ldx + c] = serve_rhs_value(reqs[r].seed, (uint64_t)j);
    }
}

__global__ __launch_bounds__(kThreads) void hash_kernel(const ServeReq* __restrict__ reqs,
                                                        const int64_t* __restrict__ begin, int n, int64_t total,
                                                        unsigned long long* __restrict__ out)
{
    __shared__ int64_t s_begin[kMaxServeReqs + 1];
    for (int i = threadIdx.x; i <= n; i += kThreads) s_begin[i] = begin[i];
    __syncthreads();
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total; e += (int64_t)gridDim.x * kThreads) {
        const int r = find_req(s_begin, n, e);
        const int64_t j = e - s_begin[r];
        const int k = reqs[r].k;
        const int64_t i = j / k, c = j - i * k;
        const float* y = static_cast<const float*>(reqs[r].y);
        const uint32_t bits = __float_as_uint(y[i * reqs[r].ldy + c]);
        // wrapping sum of per-element mixes: independent of the summation order
        atomicAdd(&out[r], (unsigned long long)serve_hash_term((uint64_t)j, bits));
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

hipError_t serve_rhs(const ServeReq* reqs, const int64_t* begin, int n, int64_t total, hipStream_t s)
{
    if (n <= 0 || total <= 0) return hipSuccess;
    if (n > kMaxServeReqs) return hipErrorInvalidValue;
    const int64_t g = (total + kThreads - 1) / kThreads;
    rhs_kernel<<<(unsigned)(g < kGrid ? g : kGrid), kThreads, 0, s>>>(reqs, begin, n, total);
    return hipGetLastError();
}

hipError_t serve_hash(const ServeReq* reqs, const int64_t* begin, int n, int64_t total, uint64_t* out,
                      hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (n > kMaxServeReqs) return hipErrorInvalidValue;
    auto* o = reinterpret_cast<unsigned long long*>(out);
    zero_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(o, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || total <= 0) return e;
    const int64_t g = (total + kThreads - 1) / kThreads;
    hash_kernel<<<(unsigned)(g < kGrid ? g : kGrid), kThreads, 0, s>>>(reqs, begin, n, total, o);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
