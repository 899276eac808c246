// merbit.hip -- MERBIT merge-path SpMV for gfx950 (wave64).
//
// Reference: TILE / Tile_Kernel / merbit_kernel (include/merbit.cuh:7-155,
// 264-493).  The merge path of a CSR with n rows and nnz nonzeros has
// n + nnz items: item p "ends a row" when p == row_ptr[r+1] + r for some row r,
// otherwise it consumes the next nonzero.  The reference cuts it into 32-lane
// tiles of SIGMA items per lane and stores a 32-bit descriptor per lane
// (delta_x, delta_y, row-end flags).  This build keeps the merge-path idea and
// reshapes the format for CDNA4:
//
//  * a block is 256 lanes = 4 wave64s, each lane owns S consecutive items
//    (S = 8, 16 or 32), T = 256*S items per block;
//  * metadata = the row-end bitmap over the merge path (one bit per item, the
//    concatenation of the reference's descriptor flag fields) plus the merge
//    coordinates (blk_x = nonzeros consumed, blk_y = rows ended) at every block
//    start.  A lane's own coordinates come from a block-wide exclusive scan of
//    the popcounts of its flag bits, so no per-lane descriptor is stored
//    (1/8 byte per item at S=16 instead of 4 bytes per 14 items);
//  * phase 1: the block's nonzeros [blk_x[b], blk_x[b+1]) are streamed with
//    coalesced dword loads, x is gathered, products go to LDS (the reference's
//    load_smem, merbit.cuh:291-303); all S loads of a lane are issued before
//    the gathers so each lane keeps S misses in flight;
//  * phase 2: each lane walks its S flag bits over the LDS products
//    (normal_track, merbit.cuh:340-384); completed rows go to an LDS row
//    buffer, the partial before the lane's first row end and the tail after
//    its last are combined by a segmented scan over the block (wave64 shuffles
//    + one LDS exchange between the 4 waves) instead of LDS float atomics;
//  * phase 3: the block writes its finished rows with coalesced stores; the
//    partial of the row still open at the block end goes to carry[b] /
//    carry_row[b] and a second, tiny kernel adds the carries of every row that
//    crosses block boundaries, in block order.  No floating-point atomics are
//    used anywhere, so y is bit-reproducible run to run (the reference's
//    cross-block atomicAdd, merbit.cuh:386-406, is order-dependent);
//  * blocks whose items all belong to one row (no row end) reduce straight
//    from registers (the reference's fast track, merbit.cuh:264-288);
//  * the hardware dispatches workgroups round-robin over the 8 XCDs; block ids
//    are remapped so each XCD walks a contiguous range of the merge path and
//    its private L2 sees the x window of that row range.
#include "dtypes.hpp"
#include "launch.hpp"
#include "ldst.hpp"
#include "xcd_order.hpp"

namespace mr {
namespace k {

namespace {

constexpr int kLanes = 256;
constexpr int kWaves = kLanes / 64;
constexpr int kMaxJobs = kMaxBatchJobs;

// blk_x/blk_y at diagonal d (merge_search, include/merbit.cuh:24-53)
__device__ __forceinline__ void merge_search(const int32_t* rp, int64_t n_rows, int64_t nnz,
                                             int64_t diag, int64_t& cx, int64_t& cy)
{
    int64_t y_min = diag - nnz > 0 ? diag - nnz : 0;
    int64_t y_max = diag < n_rows ? diag : n_rows;
    while (y_min < y_max) {
        const int64_t mid = (y_min + y_max) >> 1;
        if (rp[mid + 1] <= diag - mid - 1) y_min = mid + 1;
        else y_max = mid;
    }
    cx = diag - y_min;
    cy = y_min < n_rows ? y_min : n_rows;
}

__global__ void flags_kernel(int64_t n_rows, const int32_t* __restrict__ rp, uint32_t* __restrict__ flags)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = (int64_t)rp[r + 1] + r;
        atomicOr(&flags[p >> 5], 1u << (p & 31));
    }
}

__global__ void blocks_kernel(int64_t n_rows, int64_t nnz, int64_t items, int64_t T, int64_t n_blocks,
                              const int32_t* __restrict__ rp, int32_t* __restrict__ blk_x,
                              int32_t* __restrict__ blk_y)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b > n_blocks) return;
    int64_t d = b * T;
    if (d > items) d = items;
    int64_t cx, cy;
    merge_search(rp, n_rows, nnz, d, cx, cy);
    blk_x[b] = (int32_t)cx;
    blk_y[b] = (int32_t)cy;
}

// logical block of hardware block i: XCD (i % 8) takes a contiguous range
__device__ __forceinline__ int64_t remap_block(int64_t i, int64_t n)
{
    const int64_t full = (n / 8) * 8;
    if (i >= full) return i;
    return (i % 8) * (n / 8) + (i / 8);
}

// one MERBIT block b of one matrix (shared by the single-matrix kernel and the
// persistent batched kernel, which calls it for many (matrix, block) pairs)
template <typename VT, int S, bool NT>
__device__ __forceinline__ void merbit_block(
    int64_t b, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const uint32_t* __restrict__ flags, const int32_t* __restrict__ blk_x,
    const int32_t* __restrict__ blk_y, const typename Acc<VT>::type* __restrict__ x,
    typename Acc<VT>::type* __restrict__ y, typename Acc<VT>::type* __restrict__ carry,
    int32_t* __restrict__ carry_row)
{
    using AT = typename Acc<VT>::type;
    constexpr int T = kLanes * S;
    __shared__ AT s[T];
    __shared__ AT s_wv[kWaves];
    __shared__ int s_wf[kWaves];
    __shared__ int s_wc[kWaves];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t x0 = blk_x[b], x1 = blk_x[b + 1];
    const int32_t y0 = blk_y[b], y1 = blk_y[b + 1];
    const int nnzb = x1 - x0, rowsb = y1 - y0;
    const int32_t* cib = ci + x0;
    const VT* vb = val + x0;

    // ---- phase 1: the block's nonzeros, all S loads of a lane issued first ---
    int32_t c[S];
    VT v[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int e = u * kLanes + tid;
        c[u] = e < nnzb ? ld_stream<NT>(cib + e) : 0;
        v[u] = e < nnzb ? ld_stream<NT>(vb + e) : VT(0);
    }
    if (rowsb == 0) {
        // every item of the block consumes a nonzero of row y0 (fast track)
        AT acc = 0;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int e = u * kLanes + tid;
            if (e < nnzb) acc += (AT)to_f(v[u]) * x[c[u]];
        }
        acc = wave_sum(acc);
        if (lane == 0) s_wv[wid] = acc;
        __syncthreads();
        if (tid == 0) {
            AT t = s_wv[0];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) t += s_wv[w];
            carry[b] = t;
            carry_row[b] = y0;
        }
        return;
    }
    const int pad_end = T - rowsb;   // products [0,nnzb), zeros up to pad_end
    {
        AT xv[S];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int e = u * kLanes + tid;
            xv[u] = e < nnzb ? x[c[u]] : AT(0);
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int e = u * kLanes + tid;
            if (e < pad_end) s[e] = (AT)to_f(v[u]) * xv[u];
        }
    }

    // ---- lane coordinates: exclusive scan of row-end popcounts -------------
    const int64_t item0 = b * (int64_t)T + (int64_t)tid * S;
    uint32_t f = flags[item0 >> 5] >> (item0 & 31);
    if (S < 32) f &= (1u << S) - 1u;
    const int cnt = __popc(f);
    int inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) s_wc[wid] = inc;
    __syncthreads();
    int woff = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) woff += (w < wid) ? s_wc[w] : 0;
    const int dy = woff + inc - cnt;
    int dx = tid * S - dy;

    // ---- phase 2: walk the lane's items -------------------------------------
    AT* sy = s + pad_end;
    AT acc = 0, head = 0;
    bool has_end = false;
    int head_row = 0, row = dy;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        if ((f >> j) & 1u) {
            if (!has_end) { head = acc; head_row = row; has_end = true; }
            else sy[row] = acc;
            acc = 0;
            ++row;
        } else {
            acc += s[dx];
            ++dx;
        }
    }

    // ---- segmented scan of (tail, has_end) over the 256 lanes ----------------
    AT sv = acc;
    int fl = has_end ? 1 : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const AT vu = __shfl_up(sv, d, 64);
        const int fu = __shfl_up(fl, d, 64);
        if (lane >= d) {
            if (!fl) sv += vu;
            fl |= fu;
        }
    }
    if (lane == 63) { s_wv[wid] = sv; s_wf[wid] = fl; }
    __syncthreads();
    AT pv = 0;
#pragma unroll
    for (int w = 0; w < kWaves - 1; ++w)
        if (w < wid) pv = s_wf[w] ? s_wv[w] : pv + s_wv[w];
    const AT incl = fl ? sv : pv + sv;
    const AT up = __shfl_up(incl, 1, 64);
    const AT carry_in = lane == 0 ? pv : up;
    if (has_end) sy[head_row] = carry_in + head;
    __syncthreads();

    // ---- phase 3: finished rows out, open row to the carry buffer -------------
    for (int i = tid; i < rowsb; i += kLanes) st_stream<NT>(y + y0 + i, sy[i]);
    if (tid == kLanes - 1) {
        carry[b] = incl;
        carry_row[b] = y1;
    }
}

template <typename VT, int S, bool NT, bool REMAP>
__global__ __launch_bounds__(kLanes) void merbit_spmv_kernel(
    int64_t n_blocks, const int32_t* __restrict__ ci, const VT* __restrict__ val,
    const uint32_t* __restrict__ flags, const int32_t* __restrict__ blk_x,
    const int32_t* __restrict__ blk_y, const typename Acc<VT>::type* __restrict__ x,
    typename Acc<VT>::type* __restrict__ y, typename Acc<VT>::type* __restrict__ carry,
    int32_t* __restrict__ carry_row)
{
    const int64_t b = REMAP ? remap_block(blockIdx.x, n_blocks) : (int64_t)blockIdx.x;
    merbit_block<VT, S, NT>(b, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row);
}

// ---- batched: many matrices, one persistent launch ------------------------
// job j owns global blocks [begin[j], begin[j+1]); a block finds its job by
// binary search over the job table cached in LDS
template <typename VT>
struct BatchJobT {   // same layout as BatchJob in launch.hpp
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

__device__ __forceinline__ int find_job(const int64_t* s_begin, int n_jobs, int64_t g)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_begin[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename VT, int S, bool NT>
__global__ __launch_bounds__(kLanes) void merbit_batch_kernel(const BatchJobT<VT>* __restrict__ jobs,
                                                              const int64_t* __restrict__ begin, int n_jobs,
                                                              int64_t total, int chunk)
{
    __shared__ int64_t s_begin[kMaxJobs + 1];
    for (int i = threadIdx.x; i <= n_jobs; i += kLanes) s_begin[i] = begin[i];
    __syncthreads();
    // one tile per block (chunk > 0): tiles dealt to the XCDs in chunks of
    // consecutive tiles; else a grid-stride loop
    const int64_t g0 = chunk > 0 ? xcd_chunk_tile(blockIdx.x, chunk) : (int64_t)blockIdx.x;
    const int64_t step = chunk > 0 ? total : (int64_t)gridDim.x;
    for (int64_t g = g0; g < total; g += step) {
        const int j = find_job(s_begin, n_jobs, g);
        const BatchJobT<VT> J = jobs[j];
        merbit_block<VT, S, NT>(g - s_begin[j], J.ci, J.val, J.flags, J.blk_x, J.blk_y, J.x, J.y, J.carry,
                                J.carry_row);
        __syncthreads();   // LDS is reused by the next block
    }
}

template <typename VT>
__global__ void merbit_batch_fixup_kernel(const BatchJobT<VT>* __restrict__ jobs, const int64_t* __restrict__ begin,
                                          int n_jobs, int64_t total)
{
    using AT = typename Acc<VT>::type;
    __shared__ int64_t s_begin[kMaxJobs + 1];
    for (int i = threadIdx.x; i <= n_jobs; i += blockDim.x) s_begin[i] = begin[i];
    __syncthreads();
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int j = find_job(s_begin, n_jobs, g);
    const int64_t b = g - s_begin[j], nb = s_begin[j + 1] - s_begin[j];
    const AT* carry = jobs[j].carry;
    const int32_t* crow = jobs[j].carry_row;
    const int32_t r = crow[b];
    if (r >= jobs[j].n_rows) return;
    if (b > 0 && crow[b - 1] == r) return;
    AT s = carry[b];
    for (int64_t q = b + 1; q < nb && crow[q] == r; ++q) s += carry[q];
    jobs[j].y[r] = jobs[j].y[r] + s;
}

template <typename AT>
__global__ void merbit_fixup_kernel(int64_t n_blocks, int64_t n_rows, const AT* __restrict__ carry,
                                    const int32_t* __restrict__ carry_row, AT* __restrict__ y)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= n_blocks) return;
    const int32_t r = carry_row[b];
    if (r >= n_rows) return;
    if (b > 0 && carry_row[b - 1] == r) return;   // not the first block of the run
    AT s = carry[b];
    for (int64_t j = b + 1; j < n_blocks && carry_row[j] == r; ++j) s += carry[j];
    y[r] = y[r] + s;
}

template <typename VT, int S, bool NT, bool REMAP>
hipError_t launch_spmv(int64_t n_rows, int64_t n_blocks, const int32_t* ci, const void* val,
                       const uint32_t* flags, const int32_t* blk_x, const int32_t* blk_y,
                       const void* x, void* y, void* carry, int32_t* carry_row, hipStream_t s)
{
    using AT = typename Acc<VT>::type;
    merbit_spmv_kernel<VT, S, NT, REMAP><<<(unsigned)n_blocks, kLanes, 0, s>>>(
        n_blocks, ci, static_cast<const VT*>(val), flags, blk_x, blk_y, static_cast<const AT*>(x),
        static_cast<AT*>(y), static_cast<AT*>(carry), carry_row);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    merbit_fixup_kernel<AT><<<(unsigned)((n_blocks + 255) / 256), 256, 0, s>>>(
        n_blocks, n_rows, static_cast<const AT*>(carry), carry_row, static_cast<AT*>(y));
    return hipGetLastError();
}

template <typename VT, int S, bool NT>
hipError_t launch_batch(const void* jobs, const int64_t* begin, int n_jobs, int64_t total, int grid, hipStream_t s)
{
    const BatchJobT<VT>* J = static_cast<const BatchJobT<VT>*>(jobs);
    const int chunk = grid >= total ? kBatchChunk : 0;
    const int64_t g = chunk > 0 ? xcd_chunk_grid(total, chunk) : grid;
    merbit_batch_kernel<VT, S, NT><<<(unsigned)g, kLanes, 0, s>>>(J, begin, n_jobs, total, chunk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    merbit_batch_fixup_kernel<VT><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(J, begin, n_jobs, total);
    return hipGetLastError();
}

template <typename VT, int S>
hipError_t launch_policy(int policy, int64_t n_rows, int64_t n_blocks, const int32_t* ci, const void* val,
                         const uint32_t* flags, const int32_t* blk_x, const int32_t* blk_y, const void* x,
                         void* y, void* carry, int32_t* carry_row, hipStream_t s)
{
    switch (policy & 3) {
    case 0: return launch_spmv<VT, S, false, false>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
    case 1: return launch_spmv<VT, S, true, false>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
    case 2: return launch_spmv<VT, S, false, true>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
    default: return launch_spmv<VT, S, true, true>(n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
    }
}

}  // namespace

hipError_t merbit_build(int64_t n_rows, int64_t nnz, const int32_t* rp, int sigma, int64_t n_blocks,
                        uint32_t* flags, int64_t flag_words, int32_t* blk_x, int32_t* blk_y,
                        hipStream_t s)
{
    hipError_t e = fill_u32(flags, 0u, flag_words, s);
    if (e != hipSuccess) return e;
    if (n_rows > 0) {
        int64_t g = (n_rows + 255) / 256;
        if (g > 65536) g = 65536;
        flags_kernel<<<(unsigned)g, 256, 0, s>>>(n_rows, rp, flags);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const int64_t items = n_rows + nnz;
    const int64_t T = (int64_t)kLanes * sigma;
    blocks_kernel<<<(unsigned)((n_blocks + 1 + 255) / 256), 256, 0, s>>>(n_rows, nnz, items, T, n_blocks,
                                                                          rp, blk_x, blk_y);
    return hipGetLastError();
}

hipError_t merbit_batch_spmv(int dt, int sigma, int nt, const void* jobs, const int64_t* begin, int n_jobs,
                             int64_t total, int grid, hipStream_t s)
{
    if (total <= 0 || n_jobs <= 0) return hipSuccess;
    if (n_jobs > kMaxJobs) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        auto go = [&](auto stag) -> hipError_t {
            constexpr int S = decltype(stag)::value;
            if constexpr (S == 32 && sizeof(typename Acc<VT>::type) == 8) return hipErrorInvalidValue;
            else return nt ? launch_batch<VT, S, true>(jobs, begin, n_jobs, total, grid, s)
                           : launch_batch<VT, S, false>(jobs, begin, n_jobs, total, grid, s);
        };
        switch (sigma) {
        case 8: return go(std::integral_constant<int, 8>{});
        case 16: return go(std::integral_constant<int, 16>{});
        case 32: return go(std::integral_constant<int, 32>{});
        default: return hipErrorInvalidValue;
        }
    });
}

hipError_t merbit_coords(int64_t n_rows, int64_t nnz, const int32_t* rp, int64_t items_per_block, int64_t n_blocks,
                         int32_t* blk_x, int32_t* blk_y, hipStream_t s)
{
    blocks_kernel<<<(unsigned)((n_blocks + 1 + 255) / 256), 256, 0, s>>>(n_rows, nnz, n_rows + nnz, items_per_block,
                                                                          n_blocks, rp, blk_x, blk_y);
    return hipGetLastError();
}

hipError_t merbit_spmv(int dt, int sigma, int policy, int64_t n_rows, int64_t n_blocks, const int32_t* ci,
                       const void* val, const uint32_t* flags, const int32_t* blk_x,
                       const int32_t* blk_y, const void* x, void* y, void* carry,
                       int32_t* carry_row, hipStream_t s)
{
    if (n_blocks <= 0 || n_rows <= 0) return hipSuccess;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        switch (sigma) {
        case 8: return launch_policy<VT, 8>(policy, n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
        case 16: return launch_policy<VT, 16>(policy, n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
        case 32:
            // 256 x 32 doubles would exceed the 64 KB LDS of one workgroup
            if constexpr (sizeof(typename Acc<VT>::type) == 4)
                return launch_policy<VT, 32>(policy, n_rows, n_blocks, ci, val, flags, blk_x, blk_y, x, y, carry, carry_row, s);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
        }
    });
}

}  // namespace k
}  // namespace mr
