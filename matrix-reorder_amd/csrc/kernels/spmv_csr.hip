// spmv_csr.hip -- CSR SpMV (scalar and vector) and a CSR SpMM for gfx950.
//
// These replace the reference's cuSPARSE CSR_ALG1 calls (src/class.cu:182-207,
// 1359-1406).  The rows are binned once per matrix by length class -- <= 1,
// <= 2, 4, 8, 16, 32, 64, <= 4096, longer -- into one row list (a count pass,
// a scan of nine counts on the host, a fill pass), and one launch then gives
// every class its own shape, block ranges laid out class after class:
//   vector   class c <= 6: a group of 2^c lanes of a wave64 per row, so each
//            lane holds at most one entry of a row in its class;
//   scalar   classes <= 6: one lane per row (rows of at most 64 entries);
//   both     65..4096 entries: one wave per row; longer: one workgroup per row.
// Row sums are reduced in a fixed order (xor shuffles inside the group; the
// waves of a workgroup summed in order), so results are deterministic.  The
// launch runs the block ranges from the last class down, so the rows over 4096
// entries (17% of the nonzeros of R-MAT s22 after pGOrder, up to 98 k each)
// start first.  SpMM:
// one group per row, lanes over the k right-hand-side columns (row-major X,
// Y), so each nonzero reads one contiguous X row segment.
#include "dtypes.hpp"
#include "launch.hpp"

namespace mr {
namespace k {

namespace {

__device__ __forceinline__ int row_class(int32_t len)
{
    if (len <= 1) return 0;
    if (len > 4096) return 8;
    if (len > 64) return 7;
    return 32 - __clz(len - 1);   // 2 -> 1, 3..4 -> 2, ..., 33..64 -> 6
}

__global__ __launch_bounds__(256) void bin_count_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                        int32_t* __restrict__ counts)
{
    __shared__ int32_t c[kCsrClasses];
    if (threadIdx.x < kCsrClasses) c[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&c[row_class(rp[r + 1] - rp[r])], 1);
    __syncthreads();
    if (threadIdx.x < kCsrClasses && c[threadIdx.x]) atomicAdd(&counts[threadIdx.x], c[threadIdx.x]);
}

// rows of each class to list[offs[c] + ...] (order inside a class does not
// matter: every row is summed on its own)
__global__ __launch_bounds__(256) void bin_fill_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ offs, int32_t* __restrict__ cursor,
                                                       int32_t* __restrict__ list)
{
    __shared__ int32_t c[kCsrClasses], base[kCsrClasses];
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (threadIdx.x < kCsrClasses) c[threadIdx.x] = 0;
    __syncthreads();
    int cls = -1, at = 0;
    if (r < n_rows) {
        cls = row_class(rp[r + 1] - rp[r]);
        at = atomicAdd(&c[cls], 1);
    }
    __syncthreads();
    if (threadIdx.x < kCsrClasses)
        base[threadIdx.x] = c[threadIdx.x] ? offs[threadIdx.x] + atomicAdd(&cursor[threadIdx.x], c[threadIdx.x]) : 0;
    __syncthreads();
    if (cls >= 0) list[base[cls] + at] = (int32_t)r;
}

// rows of one class, G lanes per row (G <= 64), i = the row's index in the class
template <typename VT, int G>
__device__ __forceinline__ void rows_group(int64_t i, int32_t cnt, const int32_t* __restrict__ rows,
                                           const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                           const VT* __restrict__ val, const typename Acc<VT>::type* __restrict__ x,
                                           typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    if (i >= cnt) return;   // whole groups leave together
    const int lane = threadIdx.x & (G - 1);
    const int32_t r = rows[i];
    const int32_t b = rp[r], e = rp[r + 1];
    AT s = 0;
    if constexpr (G == 1) {
        // one lane per row of at most 64 entries: four loads in flight
        int32_t k = b;
        for (; k + 4 <= e; k += 4) {
            const int32_t c0 = ci[k], c1 = ci[k + 1], c2 = ci[k + 2], c3 = ci[k + 3];
            const AT v0 = (AT)to_f(val[k]), v1 = (AT)to_f(val[k + 1]), v2 = (AT)to_f(val[k + 2]),
                     v3 = (AT)to_f(val[k + 3]);
            const AT x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
            s += v0 * x0;
            s += v1 * x1;
            s += v2 * x2;
            s += v3 * x3;
        }
        for (; k < e; ++k) s += (AT)to_f(val[k]) * x[ci[k]];
    } else {
        for (int32_t k = b + lane; k < e; k += G) s += (AT)to_f(val[k]) * x[ci[k]];
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    }
    if (lane == 0) y[r] = s;
}

// one wave per row of 65..4096 entries: two strides in flight per lane
template <typename VT>
__device__ __forceinline__ void rows_wave(int64_t i, int32_t cnt, const int32_t* __restrict__ rows,
                                          const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                          const VT* __restrict__ val, const typename Acc<VT>::type* __restrict__ x,
                                          typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    if (i >= cnt) return;   // uniform over the wave
    const int lane = threadIdx.x & 63;
    const int32_t r = rows[i];
    const int32_t b = rp[r], e = rp[r + 1];
    AT s0 = 0, s1 = 0;
    int32_t k = b + lane;
    for (; k + 64 < e; k += 128) {
        const int32_t c0 = ci[k], c1 = ci[k + 64];
        const AT v0 = (AT)to_f(val[k]), v1 = (AT)to_f(val[k + 64]);
        s0 += v0 * x[c0];
        s1 += v1 * x[c1];
    }
    if (k < e) s0 += (AT)to_f(val[k]) * x[ci[k]];
    AT s = s0 + s1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[r] = s;
}

// one workgroup per row longer than 4096: 256 lanes stride the row, then a
// fixed-order reduction (wave xor tree, waves summed in order)
template <typename VT>
__device__ __forceinline__ void rows_block(int64_t i, const int32_t* __restrict__ rows, const int32_t* __restrict__ rp,
                                           const int32_t* __restrict__ ci, const VT* __restrict__ val,
                                           const typename Acc<VT>::type* __restrict__ x,
                                           typename Acc<VT>::type* __restrict__ y)
{
    using AT = typename Acc<VT>::type;
    __shared__ AT s_w[4];
    const int32_t r = rows[i];
    const int32_t b = rp[r], e = rp[r + 1];
    AT s0 = 0, s1 = 0;
    int32_t k = b + (int32_t)threadIdx.x;
    for (; k + 256 < e; k += 512) {
        const int32_t c0 = ci[k], c1 = ci[k + 256];
        const AT v0 = (AT)to_f(val[k]), v1 = (AT)to_f(val[k + 256]);
        s0 += v0 * x[c0];
        s1 += v1 * x[c1];
    }
    if (k < e) s0 += (AT)to_f(val[k]) * x[ci[k]];
    AT s = s0 + s1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) y[r] = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
}

// block ranges of the classes: class c owns blocks [blk[c], blk[c + 1]) and
// rows list[off[c], off[c] + cnt[c])
struct CsrLaunch {
    int32_t blk[kCsrClasses + 1];
    int32_t off[kCsrClasses];
    int32_t cnt[kCsrClasses];
};

template <typename VT, bool kScalar>
__global__ __launch_bounds__(256) void csr_binned_kernel(CsrLaunch L, const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci, const VT* __restrict__ val,
                                                         const typename Acc<VT>::type* __restrict__ x,
                                                         typename Acc<VT>::type* __restrict__ y)
{
    // blocks are dispatched in index order: the longest rows (the class
    // ranges' tail) go first, so their long walks overlap the short rows
    // instead of finishing alone after them
    const int32_t bid = (int32_t)(gridDim.x - 1 - blockIdx.x);
    int c = 0;
#pragma unroll
    for (int q = 1; q < kCsrClasses; ++q) c = bid >= L.blk[q] ? q : c;
    const int64_t local = bid - L.blk[c];
    const int32_t* rows = list + L.off[c];
    const int32_t cnt = L.cnt[c];
    const int tid = threadIdx.x;
    if (c == 8) {
        rows_block<VT>(local, rows, rp, ci, val, x, y);
        return;
    }
    if (c == 7) {
        rows_wave<VT>(local * 4 + (tid >> 6), cnt, rows, rp, ci, val, x, y);
        return;
    }
    if constexpr (kScalar) {
        rows_group<VT, 1>(local * 256 + tid, cnt, rows, rp, ci, val, x, y);
    } else {
        switch (c) {
        case 0: rows_group<VT, 1>(local * 256 + tid, cnt, rows, rp, ci, val, x, y); break;
        case 1: rows_group<VT, 2>(local * 128 + (tid >> 1), cnt, rows, rp, ci, val, x, y); break;
        case 2: rows_group<VT, 4>(local * 64 + (tid >> 2), cnt, rows, rp, ci, val, x, y); break;
        case 3: rows_group<VT, 8>(local * 32 + (tid >> 3), cnt, rows, rp, ci, val, x, y); break;
        case 4: rows_group<VT, 16>(local * 16 + (tid >> 4), cnt, rows, rp, ci, val, x, y); break;
        case 5: rows_group<VT, 32>(local * 8 + (tid >> 5), cnt, rows, rp, ci, val, x, y); break;
        default: rows_group<VT, 64>(local * 4 + (tid >> 6), cnt, rows, rp, ci, val, x, y); break;
        }
    }
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

}  // namespace

hipError_t csr_bin_count(int64_t n_rows, const int32_t* rp, int32_t* counts, hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    int64_t g = (n_rows + 255) / 256;
    if (g > 4096) g = 4096;
    bin_count_kernel<<<(unsigned)g, 256, 0, s>>>(n_rows, rp, counts);
    return hipGetLastError();
}

hipError_t csr_bin_fill(int64_t n_rows, const int32_t* rp, const int32_t* offs, int32_t* cursor, int32_t* list,
                        hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    bin_fill_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, s>>>(n_rows, rp, offs, cursor, list);
    return hipGetLastError();
}

hipError_t csr_spmv(int dt, int vector, const int32_t* list, const int32_t* counts, const int32_t* rp,
                    const int32_t* ci, const void* val, const void* x, void* y, hipStream_t s)
{
    CsrLaunch L{};
    int64_t blk = 0, off = 0;
    for (int c = 0; c < kCsrClasses; ++c) {
        const int64_t cnt = counts[c];
        L.blk[c] = (int32_t)blk;
        L.off[c] = (int32_t)off;
        L.cnt[c] = (int32_t)cnt;
        const int64_t rows_per_block = c == 8 ? 1 : c == 7 ? 4 : (vector ? 256 >> c : 256);
        blk += (cnt + rows_per_block - 1) / rows_per_block;
        off += cnt;
    }
    L.blk[kCsrClasses] = (int32_t)blk;
    if (blk <= 0) return hipSuccess;
    if (blk > INT32_MAX) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        using AT = typename Acc<VT>::type;
        if (vector)
            csr_binned_kernel<VT, false><<<(unsigned)blk, 256, 0, s>>>(L, list, rp, ci, static_cast<const VT*>(val),
                                                                       static_cast<const AT*>(x), static_cast<AT*>(y));
        else
            csr_binned_kernel<VT, true><<<(unsigned)blk, 256, 0, s>>>(L, list, rp, ci, static_cast<const VT*>(val),
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
