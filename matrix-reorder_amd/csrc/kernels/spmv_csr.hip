// spmv_csr.hip -- CSR scalar and vector SpMV and a CSR SpMM for gfx950.
//
// These replace the reference's cuSPARSE CSR_ALG1 calls (src/class.cu:182-207,
// 1359-1406).  Scalar: one lane per row.  Vector: a group of G lanes of a
// wave64 per row, G = the power of two at or above the mean row length clamped
// to [2, 64], reduced with xor shuffles inside the group (fixed order, so the
// result is deterministic).  SpMM: one group per row, lanes over the k
// right-hand-side columns (row-major X, Y), so each nonzero reads one
// contiguous X row segment.
#include "dtypes.hpp"
#include "launch.hpp"

namespace mr {
namespace k {

namespace {

constexpr int32_t kLongRow = 4096;   // rows above this go to a workgroup of their own

template <typename VT>
__global__ __launch_bounds__(256) void csr_scalar_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const VT* __restrict__ val,
                                                         const typename Acc<VT>::type* __restrict__ x,
                                                         typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    AT s = 0;
    const int32_t e = rp[r + 1];
    for (int32_t k = rp[r]; k < e; ++k) s += (AT)to_f(val[k]) * x[ci[k]];
    y[r] = s;
}

template <typename VT, int G>
__global__ __launch_bounds__(256) void csr_vector_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const VT* __restrict__ val,
                                                         const typename Acc<VT>::type* __restrict__ x,
                                                         typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    const int lane = threadIdx.x & (G - 1);
    const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
    if (r >= n_rows) return;   // whole groups leave together
    const int32_t e = rp[r + 1];
    if (e - rp[r] > kLongRow) return;   // written by csr_long_row_kernel
    AT s = 0;
    for (int32_t k = rp[r] + lane; k < e; k += G) s += (AT)to_f(val[k]) * x[ci[k]];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lane == 0) y[r] = s;
}

template <typename VT, int G>
__global__ __launch_bounds__(256) void csr_spmm_kernel(int64_t n_rows, int kk, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ ci,
                                                       const VT* __restrict__ val,
                                                       const typename Acc<VT>::type* __restrict__ X,
                                                       int64_t ldx, typename Acc<VT>::type* __restrict__ Y,
                                                       int64_t ldy)
{
    using AT = typename Acc<VT>::type;
    const int lane = threadIdx.x & (G - 1);
    const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
    if (r >= n_rows) return;
    const int32_t b = rp[r], e = rp[r + 1];
    for (int j0 = 0; j0 < kk; j0 += G) {
        const int j = j0 + lane;
        AT s = 0;
        if (j < kk)
            for (int32_t q = b; q < e; ++q) s += (AT)to_f(val[q]) * X[(int64_t)ci[q] * ldx + j];
        if (j < kk) Y[r * ldy + j] = s;
    }
}

// one workgroup per row longer than kLongRow: 256 lanes stride the row, then a
// fixed-order reduction (wave xor tree, waves summed in order)
template <typename VT>
__global__ __launch_bounds__(256) void csr_long_row_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci, const VT* __restrict__ val,
                                                           const typename Acc<VT>::type* __restrict__ x,
                                                           typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    __shared__ AT s_w[4];
    const int32_t r = rows[blockIdx.x];
    const int32_t b = rp[r], e = rp[r + 1];
    AT s = 0;
    for (int32_t k = b + (int32_t)threadIdx.x; k < e; k += 256) s += (AT)to_f(val[k]) * x[ci[k]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) y[r] = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
}

__global__ void long_rows_kernel(int64_t n_rows, const int32_t* __restrict__ rp, int32_t* __restrict__ list,
                                 int32_t* __restrict__ count)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * blockDim.x)
        if (rp[r + 1] - rp[r] > kLongRow) list[atomicAdd(count, 1)] = (int32_t)r;
}

template <typename VT, int G>
hipError_t launch_vector(int64_t n_rows, const int32_t* rp, const int32_t* ci, const void* val,
                         const void* x, void* y, hipStream_t s)
{
    using AT = typename Acc<VT>::type;
    const int64_t threads = n_rows * G;
    const unsigned grid = (unsigned)((threads + 255) / 256);
    csr_vector_kernel<VT, G><<<grid, 256, 0, s>>>(n_rows, rp, ci, static_cast<const VT*>(val),
                                                  static_cast<const AT*>(x), static_cast<AT*>(y));
    return hipGetLastError();
}

}  // namespace

hipError_t csr_long_rows(int64_t n_rows, const int32_t* rp, int32_t* list, int32_t* count, hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    int64_t g = (n_rows + 255) / 256;
    if (g > 4096) g = 4096;
    long_rows_kernel<<<(unsigned)g, 256, 0, s>>>(n_rows, rp, list, count);
    return hipGetLastError();
}

hipError_t csr_spmv(int dt, int vector, int64_t n_rows, const int32_t* rp, const int32_t* ci,
                    const void* val, const void* x, void* y, double avg_row, const int32_t* long_rows,
                    int32_t n_long, hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        using AT = typename Acc<VT>::type;
        if (!vector) {
            const unsigned grid = (unsigned)((n_rows + 255) / 256);
            csr_scalar_kernel<VT><<<grid, 256, 0, s>>>(n_rows, rp, ci, static_cast<const VT*>(val),
                                                       static_cast<const AT*>(x), static_cast<AT*>(y));
            return hipGetLastError();
        }
        hipError_t e;
        if (avg_row <= 2) e = launch_vector<VT, 2>(n_rows, rp, ci, val, x, y, s);
        else if (avg_row <= 4) e = launch_vector<VT, 4>(n_rows, rp, ci, val, x, y, s);
        else if (avg_row <= 8) e = launch_vector<VT, 8>(n_rows, rp, ci, val, x, y, s);
        else if (avg_row <= 16) e = launch_vector<VT, 16>(n_rows, rp, ci, val, x, y, s);
        else if (avg_row <= 32) e = launch_vector<VT, 32>(n_rows, rp, ci, val, x, y, s);
        else e = launch_vector<VT, 64>(n_rows, rp, ci, val, x, y, s);
        if (e != hipSuccess || n_long <= 0) return e;
        csr_long_row_kernel<VT><<<(unsigned)n_long, 256, 0, s>>>(long_rows, rp, ci, static_cast<const VT*>(val),
                                                                 static_cast<const AT*>(x), static_cast<AT*>(y));
        return hipGetLastError();
    });
}

hipError_t csr_spmm(int dt, int64_t n_rows, int kk, const int32_t* rp, const int32_t* ci,
                    const void* val, const void* X, int64_t ldx, void* Y, int64_t ldy,
                    hipStream_t s)
{
    if (n_rows <= 0 || kk <= 0) return hipSuccess;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        using AT = typename Acc<VT>::type;
        auto go = [&](auto gtag) -> hipError_t {
            constexpr int G = decltype(gtag)::value;
            const int64_t threads = n_rows * G;
            const unsigned grid = (unsigned)((threads + 255) / 256);
            csr_spmm_kernel<VT, G><<<grid, 256, 0, s>>>(n_rows, kk, rp, ci, static_cast<const VT*>(val),
                                                        static_cast<const AT*>(X), ldx,
                                                        static_cast<AT*>(Y), ldy);
            return hipGetLastError();
        };
        if (kk <= 4) return go(std::integral_constant<int, 4>{});
        if (kk <= 8) return go(std::integral_constant<int, 8>{});
        if (kk <= 16) return go(std::integral_constant<int, 16>{});
        if (kk <= 32) return go(std::integral_constant<int, 32>{});
        return go(std::integral_constant<int, 64>{});
    });
}

}  // namespace k
}  // namespace mr
