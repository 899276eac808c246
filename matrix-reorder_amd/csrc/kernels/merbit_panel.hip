// merbit_panel.hip -- MERBIT SpMV over column panels held in LDS (F32, F16 and
// BF16 storage, FP32 accumulation).
//
// The merge-path SpMV (include/merbit.cuh:7-155, 408-493 in the reference)
// gathers x[col] for every nonzero.  After pGOrder on a power-law matrix most
// of those gathers are random 4-byte reads spread over all of x, and on
// MI355X they are bound by outstanding L1 -> L2 requests, not by HBM bytes.
// This variant removes every global gather:
//
//  * the columns are cut into panels of kPanelCols = 32768 ids.  A workgroup
//    loads one panel of x (128 KB, coalesced) into LDS and multiplies only the
//    nonzeros whose columns fall in it, so every x read is an LDS read;
//  * a panel's nonzeros are stored panel-major: row by row, each row's run of
//    nonzeros in the panel (a "segment") in column order.  An item is a 16-bit
//    panel-local column with bit 15 marking the last item of a segment, and
//    the value: 2 + |val| bytes, against 4 + |val| for CSR;
//  * the stream is cut into wave tiles of 64 lanes x S items (lane l owns S
//    consecutive items).  Inside a tile the 16-byte chunks are stored chunk
//    row by chunk row (lane l's chunk c at c * 64 + l), so every vector load
//    of a wave reads 1 KB of consecutive bytes; every panel starts on a tile
//    and a segment still open at a tile end is closed there, so no sum ever
//    crosses a tile: there are no carries and no fix-up pass;
//  * each closed segment writes one partial sum; tseg[t] = index of tile t's
//    first partial.  A second kernel adds each row's partials in panel order,
//    so y is bit-reproducible.  Its input is a stream too: every row's
//    partial indices followed by a row-end marker, packed into wave tiles of
//    64 lanes x 8 items that no row crosses (padding fills a tile's tail), so
//    a lane issues eight independent partial loads and no sum leaves its
//    tile.  Rows with more partials than a tile holds are added by a third,
//    one-workgroup-per-row kernel.
//
// The plan (streams, partial lists and the workgroups' job table) is built on
// the device by the kernels at the end of this file; the host only lays the
// panels out on tiles and splits them into jobs (panel_plan.cpp).
#include "dtypes.hpp"
#include "launch_panel.hpp"
#include "ldst.hpp"

namespace mr {
namespace k {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool NT> __device__ __forceinline__ u32x4 ld_vec(const u32x4* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename VT> __device__ __forceinline__ float value_at(const uint32_t* w, int j);
template <> __device__ __forceinline__ float value_at<float>(const uint32_t* w, int j) { return __uint_as_float(w[j]); }
template <> __device__ __forceinline__ float value_at<__hip_bfloat16>(const uint32_t* w, int j)
{
    const uint32_t h = (j & 1) ? (w[j >> 1] >> 16) : (w[j >> 1] & 0xffffu);
    return __uint_as_float(h << 16);
}
template <> __device__ __forceinline__ float value_at<__half>(const uint32_t* w, int j)
{
    const uint32_t h = (j & 1) ? (w[j >> 1] >> 16) : (w[j >> 1] & 0xffffu);
    return __half2float(__ushort_as_half((unsigned short)h));
}

// ---- SpMV ------------------------------------------------------------------
// the 16-byte vectors of lane `lane` in tile t, and the tile's first partial
template <typename VT, int S, bool NT>
__device__ __forceinline__ void load_tile(const uint16_t* __restrict__ pcol, const VT* __restrict__ pval,
                                          const int32_t* __restrict__ tseg, int32_t t, int lane, uint32_t* cw,
                                          uint32_t* vw, int32_t* pb, int32_t* pe)
{
    constexpr int T = 64 * S;
    constexpr int CW = S / 2;
    constexpr int VW = S * (int)sizeof(VT) / 4;
    // chunk q of every lane is one contiguous 1 KB row of the tile: each load
    // instruction reads 64 consecutive 16-byte words
    const int64_t base = (int64_t)t * T;
    const u32x4* cp = reinterpret_cast<const u32x4*>(pcol + base) + lane;
    const u32x4* vp = reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(pval) + base * (int64_t)sizeof(VT)) +
                      lane;
#pragma unroll
    for (int q = 0; q < CW / 4; ++q) {
        const u32x4 a = ld_vec<NT>(cp + 64 * q);
        cw[4 * q] = a.x; cw[4 * q + 1] = a.y; cw[4 * q + 2] = a.z; cw[4 * q + 3] = a.w;
    }
#pragma unroll
    for (int q = 0; q < VW / 4; ++q) {
        const u32x4 b = ld_vec<NT>(vp + 64 * q);
        vw[4 * q] = b.x; vw[4 * q + 1] = b.y; vw[4 * q + 2] = b.z; vw[4 * q + 3] = b.w;
    }
    *pb = tseg[t];
    *pe = tseg[t + 1];
}

// one tile: products from the LDS panel, a wave scan ranks the closed
// segments, a segmented scan joins each lane's open head to the lanes before
// it.  The tile's partials (the contiguous range part[pbase, pend), numbered
// by the plan) are ranked into the wave's LDS staging area and written out
// with full-wave stores: one store instruction per 64 partials instead of one
// per item slot.
template <typename VT, int S, bool NT>
__device__ __forceinline__ void panel_tile(const uint32_t* cw, const uint32_t* vw, int32_t pbase, int32_t pend,
                                           int lane, const float* xs, float* stg, float* __restrict__ part)
{
    float p[S];
    uint32_t ends = 0;
#pragma unroll
    for (int jj = 0; jj < S; ++jj) {
        const uint32_t c16 = (jj & 1) ? (cw[jj >> 1] >> 16) : (cw[jj >> 1] & 0xffffu);
        p[jj] = xs[c16 & 0x7fffu];
        ends |= (c16 >> 15) << jj;
    }
#pragma unroll
    for (int jj = 0; jj < S; ++jj) p[jj] *= value_at<VT>(vw, jj);

    // rank of this lane's first partial in the tile
    const int cnt = __popc(ends);
    const int r0 = wave_incl_sum(cnt) - cnt;
    // Without branches: every item's running sum is written at the rank of
    // the next segment end at or after it, and a later write replaces it --
    // the end's own sum in this lane, or, for the slot of a lane's first end,
    // the write after the segmented scan below (a lane's items after its last
    // end land there too).  An item that is not an end leaves at most 64 S - 1
    // partials in the tile, so every rank written stays inside this wave's
    // staging area.
    int r = r0;
    float acc = 0.0f, head = 0.0f;
#pragma unroll
    for (int jj = 0; jj < S; ++jj) {
        acc += p[jj];
        const bool e = (ends >> jj) & 1u;
        stg[r] = acc;
        head = (e && r == r0) ? acc : head;
        r += e ? 1 : 0;
        acc = e ? 0.0f : acc;
    }
    const bool has_end = ends != 0;
    // segmented scan of (tail, has_end): the open sum each lane's first
    // closed segment continues from the lanes before it
    float sv = acc;
    int fl = has_end ? 1 : 0;
    wave_seg_scan(sv, fl);
    const float up = wave_shr1(sv);
    if (has_end) stg[r0] = up + head;
    // LDS operations of one wave complete in order: the reads below see the
    // writes above.  Plain stores: the partial sums stay cached for the
    // kernel that adds them.
    for (int q = pbase + lane; q < pend; q += 64) part[q] = stg[q - pbase];
}

// jobs[3j .. 3j+2] = (panel, first tile, end tile); job list i runs jobs
// [wg[i], wg[i+1]) in order.  part[tseg[t] + i] = i-th segment closed in tile t.
template <typename VT, int S, int W, bool NT>
__global__ __launch_bounds__(64 * W) void panel_spmv_kernel(const uint16_t* __restrict__ pcol,
                                                            const VT* __restrict__ pval,
                                                            const int32_t* __restrict__ tseg,
                                                            const int32_t* __restrict__ jobs,
                                                            const int32_t* __restrict__ wg,
                                                            const float* __restrict__ x, int64_t n_cols,
                                                            float* __restrict__ part)
{
    constexpr int CW = S / 2;                       // column words (two 16-bit items each)
    constexpr int VW = S * (int)sizeof(VT) / 4;     // value words
    static_assert(CW % 4 == 0 && VW % 4 == 0, "lanes load whole 16-byte vectors");
    extern __shared__ float xs[];
    // the wave index is uniform across the wave; saying so keeps the tile
    // loop's control flow and the tile indices scalar
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool x16 = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    float* stg = xs + kPanelCols + wid * (64 * S);   // this wave's partial staging

    // the next job's x panel, loaded into registers while this job's tiles
    // run (full panels of a 16-byte aligned x): a job switch then writes LDS
    // from registers instead of waiting for memory
    constexpr int XN = kPanelCols / 4 / (64 * W);   // 16-byte vectors per thread
    u32x4 xn[XN];
    bool have_xn = false;
    // job lists in order are spread so that each XCD runs a contiguous eighth
    // of them (blocks go round-robin over the 8 XCDs; gridDim.x is a multiple
    // of 8): neighbouring lists share a panel, whose second load then hits
    // that XCD's L2
    const int li = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
    const int j1 = wg[li + 1];
    for (int j = wg[li]; j < j1; ++j) {
        const int32_t pn = jobs[3 * j], t0 = jobs[3 * j + 1], t1 = jobs[3 * j + 2];
        const int64_t c0 = (int64_t)pn * kPanelCols;
        const int nc = (int)((n_cols - c0) < kPanelCols ? (n_cols - c0) : kPanelCols);
        int32_t t = t0 + wid;
        // three register sets in turn: tile t is multiplied while tiles t + W
        // and t + 2W are in flight.  Loads and stores share one in-order
        // counter, so waiting for a tile also waits for the partial stores
        // issued before its load; two tiles ahead, those stores have had a
        // whole tile to complete.  Every load is unconditional (clamped to the
        // job's last tile): a load under a branch makes the compiler wait for
        // all outstanding loads.
        const int32_t tl = t1 - 1;
        uint32_t ca[CW], va[VW], cb[CW], vb[VW], cc[CW], vc[VW];
        int32_t pa = 0, pb = 0, pc = 0, ea = 0, eb = 0, ec = 0;
        if (t < t1) {
            load_tile<VT, S, NT>(pcol, pval, tseg, t, lane, ca, va, &pa, &ea);
            load_tile<VT, S, NT>(pcol, pval, tseg, t + W < t1 ? t + W : tl, lane, cb, vb, &pb, &eb);
        }
        __syncthreads();   // the previous job's tiles are done with xs
        if (have_xn) {
#pragma unroll
            for (int q = 0; q < XN; ++q) reinterpret_cast<u32x4*>(xs)[threadIdx.x + q * 64 * W] = xn[q];
        } else if (x16) {
            const float4* xv = reinterpret_cast<const float4*>(x + c0);
            for (int i = threadIdx.x; 4 * i < nc; i += 64 * W) {
                if (4 * i + 3 < nc) {
                    reinterpret_cast<float4*>(xs)[i] = xv[i];
                } else {
                    for (int e = 4 * i; e < nc; ++e) xs[e] = x[c0 + e];
                }
            }
        } else {
            for (int i = threadIdx.x; i < nc; i += 64 * W) xs[i] = x[c0 + i];
        }
        __syncthreads();
        // issued after this job's first tile loads and before the rest: the
        // first tile's wait covers these loads too, the later ones do not
        have_xn = false;
        if (j + 1 < j1) {
            const int64_t c1 = (int64_t)jobs[3 * (j + 1)] * kPanelCols;
            if (x16 && c1 + kPanelCols <= n_cols) {
                const u32x4* xv = reinterpret_cast<const u32x4*>(x + c1);
#pragma unroll
                for (int q = 0; q < XN; ++q) xn[q] = xv[threadIdx.x + q * 64 * W];
                have_xn = true;
            }
        }

        // this wave's tiles t + k W, k < n; the main loop runs whole rounds of
        // three with no exit inside, so each register set stays in place
        const int n = t < t1 ? (t1 - t + W - 1) / W : 0;
        auto tile_at = [&](int k) { return t + k * W < t1 ? t + k * W : tl; };
        int k = 0;
        for (; k + 3 <= n; k += 3) {
            load_tile<VT, S, NT>(pcol, pval, tseg, tile_at(k + 2), lane, cc, vc, &pc, &ec);
            panel_tile<VT, S, NT>(ca, va, pa, ea, lane, xs, stg, part);
            load_tile<VT, S, NT>(pcol, pval, tseg, tile_at(k + 3), lane, ca, va, &pa, &ea);
            panel_tile<VT, S, NT>(cb, vb, pb, eb, lane, xs, stg, part);
            load_tile<VT, S, NT>(pcol, pval, tseg, tile_at(k + 4), lane, cb, vb, &pb, &eb);
            panel_tile<VT, S, NT>(cc, vc, pc, ec, lane, xs, stg, part);
        }
        if (k < n) panel_tile<VT, S, NT>(ca, va, pa, ea, lane, xs, stg, part);
        if (k + 1 < n) panel_tile<VT, S, NT>(cb, vb, pb, eb, lane, xs, stg, part);
    }
}

// y[r] = sum of row r's partials.  Tile t of the row-sum stream holds rows
// trow[t] .. trow[t + 1] - 1: each row's partial indices (>= 0), the last one
// carrying the count c of rows that end there (the row and the c - 1 empty
// rows after it), or a marker -c for a row with no partials in the tile;
// kRowPad fills the tile's tail.  Lane l owns the 8 consecutive items
// 8l .. 8l + 7, stored as two 16-byte chunks at chunk rows 0 and 1
// (c * 64 + l), so each load instruction of a wave reads 1 KB of consecutive
// bytes.
//
// A wave takes two consecutive tiles: both item loads, then all sixteen
// partial gathers, then the two sums (one tile per wave left the gathers
// latency-bound; three or four tiles per wave ran slower at R-MAT s22, and so
// did persistent waves, whose next-tile loads and y stores share the in-order
// memory counter the gathers wait on).  The blocks are dispatched round-robin
// over the 8 XCDs.  The tile groups are dealt to the XCDs in chunks of
// kRowSumChunk consecutive blocks: the rows of a chunk are consecutive, so the
// partial lines they share are fetched into one L2, and every XCD gets chunks
// from all over the matrix.  (One contiguous eighth per XCD left the XCD with
// the hub rows, whose partials spread over more lines, finishing last.)
constexpr int kRowSumChunk = 16;
__device__ __forceinline__ void rowsum_load(const int32_t* __restrict__ items, int64_t t, int lane, u32x4* a, u32x4* b)
{
    const u32x4* ip = reinterpret_cast<const u32x4*>(items + t * kRowSumTile) + lane;
    *a = __builtin_nontemporal_load(ip);
    *b = __builtin_nontemporal_load(ip + 64);
}

// the tile's eight partials per lane, each loaded at an in-bounds index
// (markers read part[0]; the sum masks them), so the loads go out together
__device__ __forceinline__ void rowsum_gather(const u32x4 a, const u32x4 b, uint32_t imask,
                                              const float* __restrict__ part, float* g)
{
    const int32_t it[8] = {(int32_t)a.x, (int32_t)a.y, (int32_t)a.z, (int32_t)a.w,
                           (int32_t)b.x, (int32_t)b.y, (int32_t)b.z, (int32_t)b.w};
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const uint32_t neg = (uint32_t)(it[jj] >> 31);   // all ones for a marker
        g[jj] = part[(uint32_t)it[jj] & imask & ~neg];
    }
}

__device__ __forceinline__ void rowsum_reduce(const u32x4 a, const u32x4 b, const float* g, int ib, int32_t r0,
                                              int32_t r1, int lane, float* stg, float* __restrict__ y)
{
    const int32_t it[8] = {(int32_t)a.x, (int32_t)a.y, (int32_t)a.z, (int32_t)a.w,
                           (int32_t)b.x, (int32_t)b.y, (int32_t)b.z, (int32_t)b.w};
    float p[8];
    int c[8];       // rows that end after each item
    uint32_t ends = 0;
    int cnt = 0;    // rows this lane's items close
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const uint32_t neg = (uint32_t)(it[jj] >> 31);
        p[jj] = __uint_as_float(__float_as_uint(g[jj]) & ~neg);
        c[jj] = neg ? (it[jj] == kRowPad ? 0 : -it[jj]) : (int)((uint32_t)it[jj] >> ib);
        ends |= (uint32_t)(c[jj] != 0) << jj;
        cnt += c[jj];
    }
    // empty rows are never written below: the staging starts at zero
#pragma unroll
    for (int q = 0; q < kRowSumRows / 64; ++q) stg[lane + 64 * q] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    // row offset of this lane's first end marker in the tile
    const int rk0 = wave_incl_sum(cnt) - cnt;
    int rk = rk0;
    float acc = 0.0f, head = 0.0f;
    bool has_end = false;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        acc += p[jj];
        if ((ends >> jj) & 1u) {
            if (!has_end) {
                head = acc;
                has_end = true;
            } else {
                stg[rk] = acc;
            }
            rk += c[jj];
            acc = 0.0f;
        }
    }
    // segmented scan of (tail, has_end): the sum each lane's first row
    // continues from the lanes before it.  No row crosses the tile, so the
    // last lane's tail is zero.
    float sv = acc;
    int fl = has_end ? 1 : 0;
    wave_seg_scan(sv, fl);
    const float up = wave_shr1(sv);
    if (has_end) stg[rk0] = up + head;
    __builtin_amdgcn_wave_barrier();
    // the tile's rows are consecutive: full-wave stores of y[r0, r1)
    for (int q = r0 + lane; q < r1; q += 64) y[q] = stg[q - r0];
    // the next tile's staging writes come after every lane's reads above
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void panel_rowsum_kernel(const int32_t* __restrict__ items,
                                                           const int32_t* __restrict__ trow, int64_t n_tiles,
                                                           int ib, const float* __restrict__ part,
                                                           float* __restrict__ y)
{
    __shared__ float stg_all[4][kRowSumRows];
    const uint32_t imask = ib >= 31 ? 0x7fffffffu : (1u << ib) - 1u;
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // block b is the i-th block of XCD x = b % 8; gridDim.x is a multiple of 8.
    // Its group: chunk i / kRowSumChunk of XCD x, dealt round-robin, or past
    // the last whole round of chunks, a contiguous share of the rest.
    const int per_xcd = gridDim.x / 8, xcd = blockIdx.x % 8, i = blockIdx.x / 8;
    const int full = per_xcd / kRowSumChunk * kRowSumChunk;
    const int64_t g = i < full ? ((int64_t)(i / kRowSumChunk) * 8 + xcd) * kRowSumChunk + i % kRowSumChunk
                               : (int64_t)8 * full + (int64_t)xcd * (per_xcd - full) + (i - full);
    const int64_t t = (g * 4 + wid) * 2;
    if (t >= n_tiles) return;   // the whole wave leaves; no block barrier follows
    const bool two = t + 1 < n_tiles;
    u32x4 a0, b0, a1, b1;
    rowsum_load(items, t, lane, &a0, &b0);
    rowsum_load(items, two ? t + 1 : t, lane, &a1, &b1);   // in bounds either way
    float g0[8], g1[8];
    rowsum_gather(a0, b0, imask, part, g0);
    rowsum_gather(a1, b1, imask, part, g1);
    rowsum_reduce(a0, b0, g0, ib, trow[t], trow[t + 1], lane, stg_all[wid], y);
    if (two) rowsum_reduce(a1, b1, g1, ib, trow[t + 1], trow[t + 2], lane, stg_all[wid], y);
}

// rows with more partials than a row-sum tile holds (their tile entry is a
// bare end marker, so the row-sum kernel wrote 0 there first): one workgroup
// per row, a fixed-order sum (lane-strided, xor tree in the wave, the four
// waves added in order)
__global__ __launch_bounds__(256) void panel_long_rows_kernel(const int32_t* __restrict__ lrow,
                                                              const int32_t* __restrict__ rs,
                                                              const int32_t* __restrict__ gidx,
                                                              const float* __restrict__ part, float* __restrict__ y)
{
    __shared__ float ws[4];
    const int32_t r = lrow[blockIdx.x];
    const int32_t b = rs[r], e = rs[r + 1];
    float acc = 0.0f;
    for (int32_t j = b + (int32_t)threadIdx.x; j < e; j += 256) acc += part[gidx[j]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) y[r] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

template <typename VT, int S, int W, bool NT>
hipError_t launch_panel(const PanelView& v, const void* x, void* part, void* y, hipStream_t s)
{
    const size_t lds = panel_lds_bytes(S, W);
    auto kern = panel_spmv_kernel<VT, S, W, NT>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (v.n_wg > 0) {
        kern<<<(unsigned)v.n_wg, 64 * W, lds, s>>>(v.pcol, static_cast<const VT*>(v.pval), v.tseg, v.jobs, v.wg,
                                                   static_cast<const float*>(x), v.n_cols, static_cast<float*>(part));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (v.n_rtiles > 0) {
        // two tiles per wave, four waves per block, a multiple of 8 blocks
        int64_t g = ((v.n_rtiles + 1) / 2 + 3) / 4;
        g = (g + 7) / 8 * 8;
        panel_rowsum_kernel<<<(unsigned)g, 256, 0, s>>>(v.items, v.trow, v.n_rtiles, v.rs_ib,
                                                        static_cast<const float*>(part), static_cast<float*>(y));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (v.n_long > 0) {
        panel_long_rows_kernel<<<(unsigned)v.n_long, 256, 0, s>>>(v.lrow, v.rs, v.gidx, static_cast<const float*>(part),
                                                                  static_cast<float*>(y));
        e = hipGetLastError();
    }
    return e;
}

// ---- plan build --------------------------------------------------------------
unsigned grid_of(int64_t work, int64_t cap = 65536)
{
    int64_t g = (work + 255) / 256;
    if (g > cap) g = cap;
    return (unsigned)(g < 1 ? 1 : g);
}

// F[k] = 1 where a segment starts: a row's first nonzero, or a nonzero whose
// panel differs from its predecessor's in the row (one wave per row)
__global__ __launch_bounds__(256) void pn_seg_flags_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci, int32_t* __restrict__ F)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n_rows; r += waves) {
        const int32_t b = rp[r], e = rp[r + 1];
        for (int32_t k = b + lane; k < e; k += 64)
            F[k] = (k == b || (ci[k] >> kPanelBits) != (ci[k - 1] >> kPanelBits)) ? 1 : 0;
    }
}

// per segment s (row-major order): first nonzero, panel key, identity payload
__global__ void pn_seg_info_kernel(int64_t nnz, const int32_t* __restrict__ ci, const int32_t* __restrict__ F,
                                   const int32_t* __restrict__ sid, int32_t* __restrict__ k0,
                                   uint32_t* __restrict__ key, uint32_t* __restrict__ ident)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        if (!F[k]) continue;
        const int32_t s = sid[k];
        k0[s] = (int32_t)k;
        key[s] = (uint32_t)(ci[k] >> kPanelBits);
        ident[s] = (uint32_t)s;
    }
}

__global__ void pn_set_tail_kernel(int32_t* a, int64_t at, const int64_t* total, int64_t fixed)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a[at] = (int32_t)(total ? *total : fixed);
}

// len[i] = length of the i-th segment in panel order
__global__ void pn_sorted_len_kernel(int64_t n_seg, const uint32_t* __restrict__ order, const int32_t* __restrict__ k0,
                                     int32_t* __restrict__ len)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_seg; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = order[i];
        len[i] = k0[s + 1] - k0[s];
    }
}

// first segment (in panel order) and first unpadded item of every panel
__global__ void pn_panel_starts_kernel(int64_t n_seg, const uint32_t* __restrict__ skey, const int32_t* __restrict__ X,
                                       int32_t* __restrict__ pfirst, int32_t* __restrict__ pstart)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_seg; i += (int64_t)gridDim.x * blockDim.x) {
        if (i == 0 || skey[i] != skey[i - 1]) {
            pfirst[skey[i]] = (int32_t)i;
            pstart[skey[i]] = X[i];
        }
    }
}

// position of every segment's first item on the padded tile layout
__global__ void pn_seg_pos_kernel(int64_t n_seg, const uint32_t* __restrict__ order, const uint32_t* __restrict__ skey,
                                  const int32_t* __restrict__ X, const int32_t* __restrict__ padd,
                                  int32_t* __restrict__ pos)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_seg; i += (int64_t)gridDim.x * blockDim.x)
        pos[order[i]] = X[i] + padd[skey[i]];
}

__global__ void pn_zero_kernel(int64_t n, uint16_t* __restrict__ pcol, uint16_t* __restrict__ pv16,
                               uint32_t* __restrict__ pv32, int32_t* __restrict__ flag)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        pcol[i] = 0;
        if (pv16) pv16[i] = 0;
        if (pv32) pv32[i] = 0u;
        flag[i] = 0;
    }
}

// items: 16-bit panel-local column (bit 15 = last item of a segment or of a
// tile), value, and the end flag for the partial numbering
template <typename VT>
__global__ void pn_scatter_kernel(int64_t nnz, int tile, const int32_t* __restrict__ ci, const VT* __restrict__ val,
                                  const int32_t* __restrict__ F, const int32_t* __restrict__ sid,
                                  const int32_t* __restrict__ k0, const int32_t* __restrict__ pos,
                                  uint16_t* __restrict__ pcol, VT* __restrict__ pval, int32_t* __restrict__ flag)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        // sid is the exclusive count of segment starts: the segment holding k
        // is one less than the inclusive count
        const int32_t s = sid[k] + F[k] - 1;
        const int64_t q = (int64_t)pos[s] + (k - k0[s]);
        const bool end = (k == (int64_t)k0[s + 1] - 1) || ((q % tile) == tile - 1);
        flag[q] = end ? 1 : 0;
        // logical item i of a tile belongs to lane i / S at slot jj = i % S;
        // lane l's 16-byte chunk c is stored at chunk row c: c * 64 + l
        const int S = tile / 64;
        const int64_t tb = q - q % tile;
        const int i = (int)(q - tb), l = i / S, jj = i % S;
        constexpr int EC = 8;                       // 16-bit columns per chunk
        constexpr int EV = 16 / (int)sizeof(VT);    // values per chunk
        pcol[tb + (int64_t)((jj / EC) * 64 + l) * EC + jj % EC] =
            (uint16_t)((ci[k] & (kPanelCols - 1)) | (end ? 0x8000 : 0));
        pval[tb + (int64_t)((jj / EV) * 64 + l) * EV + jj % EV] = val[k];
    }
}

__global__ void pn_tile_base_kernel(int64_t n_tiles, int tile, const int32_t* __restrict__ pidx,
                                    int32_t* __restrict__ tseg)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_tiles; t += (int64_t)gridDim.x * blockDim.x)
        tseg[t] = pidx[t * tile];
}

__global__ void pn_pieces_kernel(int64_t n_seg, int tile, const int32_t* __restrict__ pos,
                                 const int32_t* __restrict__ k0, int32_t* __restrict__ npieces)
{
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n_seg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q0 = pos[s], len = k0[s + 1] - k0[s];
        npieces[s] = (int32_t)((q0 + len - 1) / tile - q0 / tile + 1);
    }
}

__global__ void pn_row_start_kernel(int64_t n_rows, const int32_t* __restrict__ rp, const int32_t* __restrict__ sid,
                                    const int32_t* __restrict__ goff, int32_t* __restrict__ rs)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n_rows; r += (int64_t)gridDim.x * blockDim.x)
        rs[r] = goff[sid[rp[r]]];
}

__global__ void pn_gidx_kernel(int64_t n_seg, int tile, const int32_t* __restrict__ pos, const int32_t* __restrict__ k0,
                               const int32_t* __restrict__ goff, const int32_t* __restrict__ pidx,
                               int32_t* __restrict__ gidx)
{
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n_seg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q0 = pos[s], last = q0 + (k0[s + 1] - k0[s]) - 1;
        int32_t o = goff[s];
        for (int64_t tb = q0 / tile; tb <= last / tile; ++tb) {
            const int64_t e = ((tb + 1) * tile - 1) < last ? ((tb + 1) * tile - 1) : last;
            gidx[o++] = pidx[e];
        }
    }
}

// the row-sum stream: pads everywhere, then every row's partial indices and
// end marker at its packed position rpos[r] (logical item i of a tile
// belongs to lane i / 8, slot i % 8; lane l's chunk c at chunk row c)
__global__ void pn_rowsum_fill_kernel(int64_t n, int32_t* __restrict__ items)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        items[i] = kRowPad;
}

__device__ __forceinline__ int64_t rowsum_slot(int64_t q)
{
    const int64_t tb = q - q % kRowSumTile;
    const int i = (int)(q - tb), l = i / 8, jj = i % 8;
    return tb + (int64_t)((jj / 4) * 64 + l) * 4 + jj % 4;
}

__global__ void pn_rowsum_scatter_kernel(int64_t n_rows, int ib, const int32_t* __restrict__ rs,
                                         const int32_t* __restrict__ gidx, const int32_t* __restrict__ rpos,
                                         const int32_t* __restrict__ rcnt, int32_t* __restrict__ items)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = rcnt[r];
        if (c == 0) continue;   // an empty row closed by an earlier end
        const int32_t b = rs[r], len = rs[r + 1] - b;
        const int64_t q = rpos[r];
        if (len == 0 || len > kRowSumTile) {   // no partials, or a long row its own kernel adds
            items[rowsum_slot(q)] = -c;
            continue;
        }
        for (int32_t i = 0; i + 1 < len; ++i) items[rowsum_slot(q + i)] = gidx[b + i];
        items[rowsum_slot(q + len - 1)] = (int32_t)((uint32_t)gidx[b + len - 1] | ((uint32_t)c << ib));
    }
}

}  // namespace

hipError_t panel_seg_flags(int64_t n_rows, const int32_t* rp, const int32_t* ci, int32_t* F, hipStream_t s)
{
    if (n_rows <= 0) return hipSuccess;
    int64_t g = (n_rows + 3) / 4;
    if (g > 65536) g = 65536;
    pn_seg_flags_kernel<<<(unsigned)g, 256, 0, s>>>(n_rows, rp, ci, F);
    return hipGetLastError();
}

hipError_t panel_seg_info(int64_t nnz, const int32_t* ci, const int32_t* F, const int32_t* sid, int32_t* k0,
                          uint32_t* key, uint32_t* ident, hipStream_t s)
{
    if (nnz <= 0) return hipSuccess;
    pn_seg_info_kernel<<<grid_of(nnz), 256, 0, s>>>(nnz, ci, F, sid, k0, key, ident);
    return hipGetLastError();
}

hipError_t panel_set_tail(int32_t* a, int64_t at, const int64_t* total, int64_t fixed, hipStream_t s)
{
    pn_set_tail_kernel<<<1, 64, 0, s>>>(a, at, total, fixed);
    return hipGetLastError();
}

hipError_t panel_sorted_len(int64_t n_seg, const uint32_t* order, const int32_t* k0, int32_t* len, hipStream_t s)
{
    if (n_seg <= 0) return hipSuccess;
    pn_sorted_len_kernel<<<grid_of(n_seg), 256, 0, s>>>(n_seg, order, k0, len);
    return hipGetLastError();
}

hipError_t panel_starts(int64_t n_seg, const uint32_t* skey, const int32_t* X, int32_t* pfirst, int32_t* pstart,
                        hipStream_t s)
{
    if (n_seg <= 0) return hipSuccess;
    pn_panel_starts_kernel<<<grid_of(n_seg), 256, 0, s>>>(n_seg, skey, X, pfirst, pstart);
    return hipGetLastError();
}

hipError_t panel_seg_pos(int64_t n_seg, const uint32_t* order, const uint32_t* skey, const int32_t* X,
                         const int32_t* padd, int32_t* pos, hipStream_t s)
{
    if (n_seg <= 0) return hipSuccess;
    pn_seg_pos_kernel<<<grid_of(n_seg), 256, 0, s>>>(n_seg, order, skey, X, padd, pos);
    return hipGetLastError();
}

hipError_t panel_scatter(int dt, int64_t n_items, int64_t nnz, int tile, const int32_t* ci, const void* val,
                         const int32_t* F, const int32_t* sid, const int32_t* k0, const int32_t* pos, uint16_t* pcol,
                         void* pval, int32_t* flag, hipStream_t s)
{
    if (dt == DT_F64) return hipErrorInvalidValue;
    if (n_items > 0) {
        pn_zero_kernel<<<grid_of(n_items), 256, 0, s>>>(n_items, pcol, dt == DT_F32 ? nullptr : static_cast<uint16_t*>(pval),
                                                         dt == DT_F32 ? static_cast<uint32_t*>(pval) : nullptr, flag);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nnz <= 0) return hipSuccess;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        pn_scatter_kernel<VT><<<grid_of(nnz), 256, 0, s>>>(nnz, tile, ci, static_cast<const VT*>(val), F, sid, k0, pos,
                                                           pcol, static_cast<VT*>(pval), flag);
        return hipGetLastError();
    });
}

hipError_t panel_tile_base(int64_t n_tiles, int tile, const int32_t* pidx, int32_t* tseg, hipStream_t s)
{
    if (n_tiles <= 0) return hipSuccess;
    pn_tile_base_kernel<<<grid_of(n_tiles), 256, 0, s>>>(n_tiles, tile, pidx, tseg);
    return hipGetLastError();
}

hipError_t panel_pieces(int64_t n_seg, int tile, const int32_t* pos, const int32_t* k0, int32_t* npieces,
                        hipStream_t s)
{
    if (n_seg <= 0) return hipSuccess;
    pn_pieces_kernel<<<grid_of(n_seg), 256, 0, s>>>(n_seg, tile, pos, k0, npieces);
    return hipGetLastError();
}

hipError_t panel_row_starts(int64_t n_rows, const int32_t* rp, const int32_t* sid, const int32_t* goff, int32_t* rs,
                            hipStream_t s)
{
    pn_row_start_kernel<<<grid_of(n_rows + 1), 256, 0, s>>>(n_rows, rp, sid, goff, rs);
    return hipGetLastError();
}

hipError_t panel_gidx(int64_t n_seg, int tile, const int32_t* pos, const int32_t* k0, const int32_t* goff,
                      const int32_t* pidx, int32_t* gidx, hipStream_t s)
{
    if (n_seg <= 0) return hipSuccess;
    pn_gidx_kernel<<<grid_of(n_seg), 256, 0, s>>>(n_seg, tile, pos, k0, goff, pidx, gidx);
    return hipGetLastError();
}

hipError_t panel_rowsum_build(int64_t n_rows, int64_t n_items, int ib, const int32_t* rs, const int32_t* gidx,
                              const int32_t* rpos, const int32_t* rcnt, int32_t* items, hipStream_t s)
{
    if (n_items > 0) {
        pn_rowsum_fill_kernel<<<grid_of(n_items), 256, 0, s>>>(n_items, items);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_rows <= 0) return hipSuccess;
    pn_rowsum_scatter_kernel<<<grid_of(n_rows), 256, 0, s>>>(n_rows, ib, rs, gidx, rpos, rcnt, items);
    return hipGetLastError();
}

size_t panel_lds_bytes(int S, int W) { return sizeof(float) * ((size_t)kPanelCols + (size_t)W * 64 * S); }

hipError_t panel_spmv(int dt, const PanelView& v, const void* x, void* part, void* y, hipStream_t s)
{
    if (v.n_rows <= 0) return hipSuccess;
    if (dt == DT_F64) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto tag) -> hipError_t {
        using VT = decltype(tag);
        if constexpr (sizeof(VT) == 8) {
            return hipErrorInvalidValue;
        } else {
            auto go = [&](auto stag, auto wtag) -> hipError_t {
                constexpr int SS = decltype(stag)::value;
                constexpr int WW = decltype(wtag)::value;
                if constexpr ((SS / 2) % 4 != 0 || (SS * sizeof(VT)) % 16 != 0) {
                    return hipErrorInvalidValue;
                } else {
                    return v.nt ? launch_panel<VT, SS, WW, true>(v, x, part, y, s)
                                : launch_panel<VT, SS, WW, false>(v, x, part, y, s);
                }
            };
            using I8 = std::integral_constant<int, 8>;
            using I16 = std::integral_constant<int, 16>;
            if (v.S == 8 && v.W == 16) return go(I8{}, I16{});
            if (v.S == 8 && v.W == 8) return go(I8{}, I8{});
            if (v.S == 16 && v.W == 8) return go(I16{}, I8{});
            return hipErrorInvalidValue;
        }
    });
}

}  // namespace k
}  // namespace mr
