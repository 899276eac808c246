// merbit_mm.hip -- MERBIT merge-path SpMM (k right-hand sides) for gfx950.
//
// The merge path and its row-end bitmap are those of the SpMV plan (see
// merbit.hip; reference: include/merbit.cuh).  With k right-hand sides
// (row-major X[n_cols][ldx], Y[n_rows][ldy]) a group of G = k lanes (k a power
// of two, 2..64) owns one segment of S = 16*k consecutive merge-path items, one
// lane per right-hand-side column, so a block of 256 lanes covers 256/k
// segments = 4096 items -- the same block grid as the SpMV plan at sigma 16.
// Inside a segment the row ends split its nonzeros into runs; each run is a
// plain loop over contiguous nonzeros: the G lanes load the same (col, val)
// (one broadcast request) and read one contiguous X row slice (coalesced).
// Rows that end inside a segment after its first row end are stored straight
// to Y (coalesced over the k columns); the partial before the first row end
// and the tail after the last are combined over the block's segments through
// LDS in segment order, and the block's open row goes to a k-wide carry that
// the fix-up kernel adds in block order.  No floating-point atomics: Y is
// bit-reproducible.
#include "dtypes.hpp"
#include "launch.hpp"
#include "xcd_order.hpp"

namespace mr {
namespace k {

namespace {

constexpr int kLanes = 256;
constexpr int kItems = 4096;   // merge-path items per block

template <typename VT, int G>
__device__ __forceinline__ void merbit_mm_block(
    int64_t b, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const uint32_t* __restrict__ flags, const int32_t* __restrict__ blk_x, const int32_t* __restrict__ blk_y,
    const typename Acc<VT>::type* __restrict__ X, int64_t ldx, typename Acc<VT>::type* __restrict__ Y,
    int64_t ldy, int kk, int64_t nnz_total, typename Acc<VT>::type* __restrict__ carry,
    int32_t* __restrict__ carry_row)
{
    using AT = typename Acc<VT>::type;
    constexpr int NS = kLanes / G;         // segments per block
    constexpr int S = kItems / NS;         // items per segment (= 16 * G)
    constexpr int W = S / 32;              // flag words per segment
    __shared__ int s_cnt[NS];
    __shared__ int s_off[NS + 1];
    __shared__ AT s_tail[NS][G];
    __shared__ int s_has[NS];

    const int tid = threadIdx.x, j = tid & (G - 1), seg = tid / G;
    const int32_t x0 = blk_x[b];
    const int32_t y0 = blk_y[b], y1 = blk_y[b + 1];
    const bool col_ok = j < kk;

    // row ends per segment, then segment start coordinates
    const uint32_t* fw = flags + (b * kItems + (int64_t)seg * S) / 32;
    int part = 0;
    for (int w = j; w < W; w += G) part += __popc(fw[w]);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (j == 0) s_cnt[seg] = part;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int q = 0; q < NS; ++q) { s_off[q] = run; run += s_cnt[q]; }
        s_off[NS] = run;
    }
    __syncthreads();
    const int dy = s_off[seg];
    int32_t gx = x0 + (seg * S - dy);      // first nonzero of the segment
    int32_t row = y0 + dy;                 // row of the first item

    // walk the segment: runs of nonzeros separated by row ends
    AT acc = 0, head = 0;
    bool has_end = false;
    int32_t head_row = 0;
    for (int w = 0; w < W; ++w) {
        uint32_t f = fw[w];
        int pos = 0;                         // items of this word consumed
        while (pos < 32) {
            const uint32_t rest = f >> pos;
            const int run = rest ? __ffs(rest) - 1 : 32 - pos;   // nonzeros before the next row end
            // sum the run: contiguous nonzeros gx .. gx+run-1; past the last
            // real item the bitmap holds zeros that name no nonzero
            const int real = (int)min<int64_t>(run, max<int64_t>(0, nnz_total - gx));
            int q = 0;
            for (; q + 4 <= real; q += 4) {
                const int32_t c0 = ci[gx + q], c1 = ci[gx + q + 1], c2 = ci[gx + q + 2], c3 = ci[gx + q + 3];
                const AT v0 = (AT)to_f(val[gx + q]), v1 = (AT)to_f(val[gx + q + 1]);
                const AT v2 = (AT)to_f(val[gx + q + 2]), v3 = (AT)to_f(val[gx + q + 3]);
                if (col_ok) {
                    const AT a0 = X[(int64_t)c0 * ldx + j], a1 = X[(int64_t)c1 * ldx + j];
                    const AT a2 = X[(int64_t)c2 * ldx + j], a3 = X[(int64_t)c3 * ldx + j];
                    acc += v0 * a0;
                    acc += v1 * a1;
                    acc += v2 * a2;
                    acc += v3 * a3;
                }
            }
            for (; q < real; ++q) {
                const int32_t c = ci[gx + q];
                const AT v = (AT)to_f(val[gx + q]);
                if (col_ok) acc += v * X[(int64_t)c * ldx + j];
            }
            gx += run;
            pos += run;
            if (pos >= 32) break;
            // the item at pos ends row `row`
            if (!has_end) {
                head = acc;
                head_row = row;
                has_end = true;
            } else if (col_ok) {
                Y[(int64_t)row * ldy + j] = acc;
            }
            acc = 0;
            ++row;
            ++pos;
        }
    }

    // combine partials over the block's segments in segment order
    s_tail[seg][j] = acc;
    if (j == 0) s_has[seg] = has_end ? 1 : 0;
    __syncthreads();
    AT carry_in = 0;
    for (int q = seg - 1; q >= 0; --q) {
        carry_in += s_tail[q][j];
        if (s_has[q]) break;
    }
    if (has_end && col_ok) Y[(int64_t)head_row * ldy + j] = carry_in + head;
    if (seg == NS - 1) {
        // the block's open row: this segment's tail plus, when it has no row
        // end, the chain of tails before it
        AT out = acc;
        if (!has_end) out += carry_in;
        carry[b * G + j] = out;
        if (j == 0) carry_row[b] = y1;
    }
}

template <typename VT, int G>
__global__ __launch_bounds__(kLanes) void merbit_mm_kernel(
    int64_t n_blocks, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const uint32_t* __restrict__ flags, const int32_t* __restrict__ blk_x, const int32_t* __restrict__ blk_y,
    const typename Acc<VT>::type* __restrict__ X, int64_t ldx, typename Acc<VT>::type* __restrict__ Y,
    int64_t ldy, int kk, int64_t nnz_total, typename Acc<VT>::type* __restrict__ carry,
    int32_t* __restrict__ carry_row)
{
    merbit_mm_block<VT, G>(blockIdx.x, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz_total, carry, carry_row);
}

// ---- batched SpMM: many matrices (same k), one persistent launch ------------
template <typename VT>
struct MmJobT {   // same layout as BatchJob in launch.hpp
    const int32_t* ci;
    const VT* val;
    const uint32_t* flags;
    const int32_t* blk_x;
    const int32_t* blk_y;
    const typename Acc<VT>::type* x;
    typename Acc<VT>::type* y;
    typename Acc<VT>::type* carry;
    int32_t* carry_row;
    int64_t n_rows, nnz, ldx, ldy;
    int32_t kk, pad;
};

__device__ __forceinline__ int find_job_mm(const int64_t* s_begin, int n_jobs, int64_t g)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_begin[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename VT, int G>
__global__ __launch_bounds__(kLanes) void merbit_mm_batch_kernel(const MmJobT<VT>* __restrict__ jobs,
                                                                 const int64_t* __restrict__ begin, int n_jobs,
                                                                 int64_t total, int chunk)
{
    __shared__ int64_t s_begin[kMaxBatchJobs + 1];
    for (int i = threadIdx.x; i <= n_jobs; i += kLanes) s_begin[i] = begin[i];
    __syncthreads();
    // one tile per block (chunk > 0): tiles dealt to the XCDs in chunks of
    // consecutive tiles; else a grid-stride loop
    const int64_t g0 = chunk > 0 ? xcd_chunk_tile(blockIdx.x, chunk) : (int64_t)blockIdx.x;
    const int64_t step = chunk > 0 ? total : (int64_t)gridDim.x;
    for (int64_t g = g0; g < total; g += step) {
        const int j = find_job_mm(s_begin, n_jobs, g);
        const MmJobT<VT> J = jobs[j];
        merbit_mm_block<VT, G>(g - s_begin[j], J.ci, J.val, J.flags, J.blk_x, J.blk_y, J.x, J.ldx, J.y, J.ldy, J.kk,
                               J.nnz, J.carry, J.carry_row);
        __syncthreads();
    }
}

template <typename VT, int G>
__global__ void merbit_mm_batch_fixup_kernel(const MmJobT<VT>* __restrict__ jobs, const int64_t* __restrict__ begin,
                                             int n_jobs, int64_t total)
{
    using AT = typename Acc<VT>::type;
    __shared__ int64_t s_begin[kMaxBatchJobs + 1];
    for (int i = threadIdx.x; i <= n_jobs; i += blockDim.x) s_begin[i] = begin[i];
    __syncthreads();
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t g = t / G;
    const int jc = (int)(t % G);
    if (g >= total) return;
    const int j = find_job_mm(s_begin, n_jobs, g);
    if (jc >= jobs[j].kk) return;
    const int64_t b = g - s_begin[j], nb = s_begin[j + 1] - s_begin[j];
    const AT* carry = jobs[j].carry;
    const int32_t* crow = jobs[j].carry_row;
    const int32_t r = crow[b];
    if (r >= jobs[j].n_rows) return;
    if (b > 0 && crow[b - 1] == r) return;
    AT s = carry[b * G + jc];
    for (int64_t q = b + 1; q < nb && crow[q] == r; ++q) s += carry[q * G + jc];
    AT* Y = jobs[j].y;
    const int64_t ldy = jobs[j].ldy;
    Y[(int64_t)r * ldy + jc] = Y[(int64_t)r * ldy + jc] + s;
}

template <typename AT, int G>
__global__ void merbit_mm_fixup_kernel(int64_t n_blocks, int64_t n_rows, int kk, const AT* __restrict__ carry,
                                       const int32_t* __restrict__ carry_row, AT* __restrict__ Y, int64_t ldy)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t b = t / G;
    const int j = (int)(t % G);
    if (b >= n_blocks || j >= kk) return;
    const int32_t r = carry_row[b];
    if (r >= n_rows) return;
    if (b > 0 && carry_row[b - 1] == r) return;
    AT s = carry[b * G + j];
    for (int64_t q = b + 1; q < n_blocks && carry_row[q] == r; ++q) s += carry[q * G + j];
    Y[(int64_t)r * ldy + j] = Y[(int64_t)r * ldy + j] + s;
}

template <typename VT, int G>
hipError_t launch_mm(int64_t n_rows, int64_t n_blocks, const int32_t* ci, const void* val, const uint32_t* flags,
                     const int32_t* blk_x, const int32_t* blk_y, const void* X, int64_t ldx, void* Y, int64_t ldy,
                     int kk, int64_t nnz_total, void* carry, int32_t* carry_row, hipStream_t s)
{
    using AT = typename Acc<VT>::type;
    merbit_mm_kernel<VT, G><<<(unsigned)n_blocks, kLanes, 0, s>>>(
        n_blocks, ci, static_cast<const VT*>(val), flags, blk_x, blk_y, static_cast<const AT*>(X), ldx,
        static_cast<AT*>(Y), ldy, kk, nnz_total, static_cast<AT*>(carry), carry_row);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t threads = n_blocks * G;
    merbit_mm_fixup_kernel<AT, G><<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(
        n_blocks, n_rows, kk, static_cast<const AT*>(carry), carry_row, static_cast<AT*>(Y), ldy);
    return hipGetLastError();
}

}  // namespace

hipError_t merbit_batch_spmm(int dt, int kk, const void* jobs, const int64_t* begin, int n_jobs, int64_t total,
                             int grid, hipStream_t s)
{
    if (total <= 0 || n_jobs <= 0) return hipSuccess;
    if (kk > 64 || n_jobs > kMaxBatchJobs) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        auto go = [&](auto gtag) -> hipError_t {
            constexpr int G = decltype(gtag)::value;
            const MmJobT<VT>* J = static_cast<const MmJobT<VT>*>(jobs);
            const int chunk = grid >= total ? kBatchChunk : 0;
            const int64_t gr = chunk > 0 ? xcd_chunk_grid(total, chunk) : grid;
            merbit_mm_batch_kernel<VT, G><<<(unsigned)gr, kLanes, 0, s>>>(J, begin, n_jobs, total, chunk);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            const int64_t threads = total * G;
            merbit_mm_batch_fixup_kernel<VT, G><<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(J, begin, n_jobs,
                                                                                                  total);
            return hipGetLastError();
        };
        int g = 2;
        while (g < kk && g < 64) g <<= 1;
        switch (g) {
        case 2: return go(std::integral_constant<int, 2>{});
        case 4: return go(std::integral_constant<int, 4>{});
        case 8: return go(std::integral_constant<int, 8>{});
        case 16: return go(std::integral_constant<int, 16>{});
        case 32: return go(std::integral_constant<int, 32>{});
        default: return go(std::integral_constant<int, 64>{});
        }
    });
}

int merbit_mm_group(int kk)
{
    int g = 2;
    while (g < kk && g < 64) g <<= 1;
    return g;
}

hipError_t merbit_spmm(int dt, int64_t n_rows, int64_t nnz, int64_t n_blocks, const int32_t* ci, const void* val,
                       const uint32_t* flags, const int32_t* blk_x, const int32_t* blk_y, const void* X,
                       int64_t ldx, void* Y, int64_t ldy, int kk, void* carry, int32_t* carry_row, hipStream_t s)
{
    if (n_blocks <= 0 || n_rows <= 0 || kk <= 0) return hipSuccess;
    if (kk > 64) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        switch (merbit_mm_group(kk)) {
        case 2: return launch_mm<VT, 2>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        case 4: return launch_mm<VT, 4>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        case 8: return launch_mm<VT, 8>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        case 16: return launch_mm<VT, 16>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        case 32: return launch_mm<VT, 32>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        default: return launch_mm<VT, 64>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, X, ldx, Y, ldy, kk, nnz, carry, carry_row, s);
        }
    });
}

}  // namespace k
}  // namespace mr
