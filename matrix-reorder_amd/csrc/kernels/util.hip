// util.hip -- small gfx950 kernels: fills, value-format conversion, vector
// relabelling and CSR validation.
#include "dtypes.hpp"
#include "launch.hpp"

namespace mr {
namespace k {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int per_block = kBlock)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > 1048576) g = 1048576;   // grid-stride loops cover the rest
    return (unsigned)g;
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// double -> bfloat16 with a single round-to-nearest-even step: the float
// rounding of d can land exactly on a bf16 midpoint; the sign of d - (double)f
// then decides the direction.
__device__ __forceinline__ __hip_bfloat16 d2bf16(double d)
{
    const float f = (float)d;
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) {          // inf / nan pass through
        u = (u & 0x007fffffu) ? (u | 0x00400000u) : u;
        return __ushort_as_bfloat16((unsigned short)(u >> 16));
    }
    const uint32_t lower = u & 0xffffu;
    if (lower == 0x8000u && (double)f != d) {
        const bool above = fabs(d) > fabs((double)f);
        u = above ? ((u + 0x8000u) & 0xffff0000u) : (u & 0xffff0000u);
    } else {
        u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    }
    return __ushort_as_bfloat16((unsigned short)(u >> 16));
}

template <typename S> __device__ __forceinline__ double load_d(const S* p, int64_t i);
template <> __device__ __forceinline__ double load_d(const double* p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double load_d(const float* p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ double load_d(const __half* p, int64_t i) { return __half2float(p[i]); }
template <> __device__ __forceinline__ double load_d(const __hip_bfloat16* p, int64_t i) { return __bfloat162float(p[i]); }

template <typename D> __device__ __forceinline__ void store_d(D* p, int64_t i, double v);
template <> __device__ __forceinline__ void store_d(double* p, int64_t i, double v) { p[i] = v; }
template <> __device__ __forceinline__ void store_d(float* p, int64_t i, double v) { p[i] = (float)v; }
template <> __device__ __forceinline__ void store_d(__half* p, int64_t i, double v) { p[i] = __half(static_cast<_Float16>(v)); }
template <> __device__ __forceinline__ void store_d(__hip_bfloat16* p, int64_t i, double v) { p[i] = d2bf16(v); }

template <typename S, typename D>
__global__ void convert_kernel(const S* src, D* dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        store_d(dst, i, load_d(src, i));
}

template <typename D>
__global__ void ones_kernel(D* dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        store_d(dst, i, 1.0);
}

template <typename E>
__global__ void permute_vec_kernel(const int32_t* perm, int64_t n, int dir, const E* in, E* out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (dir == 0) out[perm[i]] = in[i];
        else out[i] = in[perm[i]];
    }
}

// streaming copy of 16-byte words: the achievable-bandwidth reference
__global__ __launch_bounds__(256) void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void validate_rows_kernel(int64_t n_rows, const int32_t* rp, int64_t nnz, int32_t* flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (rp[i + 1] < rp[i]) atomicOr(flag, 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (rp[0] != 0 || rp[n_rows] != nnz) atomicOr(flag, 1);
    }
}

__global__ void validate_cols_kernel(int64_t n_cols, const int32_t* ci, int64_t nnz, int32_t* flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = ci[i];
        if (c < 0 || c >= n_cols) atomicOr(flag, 2);
    }
}

}  // namespace

hipError_t fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    fill_u32_kernel<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t convert_values(const void* src, int src_dt, void* dst, int dst_dt, int64_t n,
                          hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    return dispatch_dt(src_dt, [&](auto sv) {
        using S = decltype(sv);
        return dispatch_dt(dst_dt, [&](auto dv) {
            using D = decltype(dv);
            convert_kernel<S, D><<<grid_for(n), kBlock, 0, s>>>(static_cast<const S*>(src),
                                                                static_cast<D*>(dst), n);
            return hipGetLastError();
        });
    });
}

hipError_t fill_ones(void* dst, int dt, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    return dispatch_dt(dt, [&](auto dv) {
        using D = decltype(dv);
        ones_kernel<D><<<grid_for(n), kBlock, 0, s>>>(static_cast<D*>(dst), n);
        return hipGetLastError();
    });
}

hipError_t permute_vec(const int32_t* perm, int64_t n, int dir, const void* in, void* out,
                       int elem, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (elem == 8)
        permute_vec_kernel<uint64_t><<<grid_for(n), kBlock, 0, s>>>(
            perm, n, dir, static_cast<const uint64_t*>(in), static_cast<uint64_t*>(out));
    else
        permute_vec_kernel<uint32_t><<<grid_for(n), kBlock, 0, s>>>(
            perm, n, dir, static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out));
    return hipGetLastError();
}

hipError_t copy_bytes(void* dst, const void* src, int64_t bytes, int cu_count, hipStream_t s)
{
    const int64_t n16 = bytes / 16;
    if (n16 <= 0) return hipSuccess;
    int64_t g = (int64_t)cu_count * 8;
    if (g > (n16 + 255) / 256) g = (n16 + 255) / 256;
    if (g < 1) g = 1;
    copy16_kernel<<<(unsigned)g, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    return hipGetLastError();
}

hipError_t validate_csr(int64_t n_rows, int64_t n_cols, const int32_t* rp, const int32_t* ci,
                        int64_t nnz, int32_t* flag, hipStream_t s)
{
    validate_rows_kernel<<<grid_for(n_rows > 0 ? n_rows : 1), kBlock, 0, s>>>(n_rows, rp, nnz, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nnz <= 0) return e;
    validate_cols_kernel<<<grid_for(nnz), kBlock, 0, s>>>(n_cols, ci, nnz, flag);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
