// scan.hip -- device exclusive prefix sum of int32 counts into int32 offsets
// (three-phase: per-tile sums, scan of the tile sums, tile scan plus offset).
#include "dtypes.hpp"
#include "launch_perm.hpp"

namespace mr {
namespace k {

namespace {

constexpr int kThreads = 256;
constexpr int kPer = 4;
constexpr int kTile = kThreads * kPer;   // 1024 elements per tile

__device__ __forceinline__ int64_t block_incl_scan(int64_t v, int64_t* s_w)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    if (lane == 63) s_w[wid] = v;
    __syncthreads();
    int64_t off = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) off += (w < wid) ? s_w[w] : 0;
    __syncthreads();
    return v + off;
}

__global__ __launch_bounds__(kThreads) void tile_sums_kernel(const int32_t* in, int64_t n, int64_t* sums)
{
    __shared__ int64_t s_w[kThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kTile + threadIdx.x * kPer;
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
        if (base + i < n) v += in[base + i];
    v = block_incl_scan(v, s_w);
    if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = v;
}

__global__ __launch_bounds__(kThreads) void tile_sums64_kernel(const int64_t* in, int64_t n, int64_t* sums)
{
    __shared__ int64_t s_w[kThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kTile + threadIdx.x * kPer;
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
        if (base + i < n) v += in[base + i];
    v = block_incl_scan(v, s_w);
    if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = v;
}

// exclusive scan of int64 tile sums in place (n <= kTile * kTile handled by recursion on host)
__global__ __launch_bounds__(kThreads) void scan_sums_kernel(int64_t* sums, int64_t n, const int64_t* carry_in)
{
    __shared__ int64_t s_w[kThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kTile + threadIdx.x * kPer;
    int64_t loc[kPer];
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        loc[i] = (base + i < n) ? sums[base + i] : 0;
        v += loc[i];
    }
    const int64_t incl = block_incl_scan(v, s_w);
    int64_t run = incl - v + (carry_in ? carry_in[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        if (base + i < n) sums[base + i] = run;
        run += loc[i];
    }
}

__global__ __launch_bounds__(kThreads) void tile_scan_kernel(const int32_t* in, int64_t n, const int64_t* offs,
                                                             int32_t* out, int64_t* total)
{
    __shared__ int64_t s_w[kThreads / 64];
    const int64_t base = blockIdx.x * (int64_t)kTile + threadIdx.x * kPer;
    int64_t loc[kPer];
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        loc[i] = (base + i < n) ? in[base + i] : 0;
        v += loc[i];
    }
    const int64_t incl = block_incl_scan(v, s_w);
    int64_t run = incl - v + offs[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        if (base + i < n) out[base + i] = (int32_t)run;
        run += loc[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) *total = run;
}

}  // namespace

size_t scan_workspace_bytes(int64_t n)
{
    size_t b = 0;
    int64_t m = (n + kTile - 1) / kTile;
    while (true) {
        b += sizeof(int64_t) * (size_t)(m + 1);
        if (m <= 1) break;
        m = (m + kTile - 1) / kTile;
    }
    return b + 256;
}

// exclusive scan of n int32 values into out[0..n); *total (device, optional) = sum
hipError_t exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, int64_t* total, void* ws,
                              hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t* sums = static_cast<int64_t*>(ws);
    const int64_t nt = (n + kTile - 1) / kTile;
    tile_sums_kernel<<<(unsigned)nt, kThreads, 0, s>>>(in, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // scan the tile sums: one block while they fit a tile, else two levels
    if (nt <= kTile) {
        scan_sums_kernel<<<1, kThreads, 0, s>>>(sums, nt, nullptr);
    } else {
        int64_t* sums2 = sums + nt + 1;
        const int64_t nt2 = (nt + kTile - 1) / kTile;
        if (nt2 > kTile) return hipErrorInvalidValue;   // > 2^30 elements
        tile_sums64_kernel<<<(unsigned)nt2, kThreads, 0, s>>>(sums, nt, sums2);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        scan_sums_kernel<<<1, kThreads, 0, s>>>(sums2, nt2, nullptr);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        scan_sums_kernel<<<(unsigned)nt2, kThreads, 0, s>>>(sums, nt, sums2);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    tile_scan_kernel<<<(unsigned)nt, kThreads, 0, s>>>(in, n, sums, out, total);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
