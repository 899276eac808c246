// merbit_lp.hip -- MERBIT SpMV on lane-padded wave tiles with an LDS cache of
// the hottest columns of x (FP32 accumulation: F32, F16 and BF16 storage).
//
// The merge path is the one of merbit.hip (include/merbit.cuh:7-155 in the
// reference): item p of the n + nnz items ends a row when p = row_ptr[r+1] + r
// and consumes a nonzero otherwise.  This variant stores the path itself:
//
//  * pc[p] / pv[p] hold, for every merge item p, the column and value of its
//    nonzero, or the marker -1 (value 0) when p ends a row; the stream is padded
//    with -2 to a whole number of tiles.  That costs one slot per row (4 + |val|
//    bytes), and buys a fixed layout: a tile is 64 lanes x S consecutive items,
//    lane l owns items [t*T + l*S, t*T + l*S + S), and loads them with 16-byte
//    vector loads;
//  * a tile belongs to one wave.  The lane walk runs on registers with
//    compile-time indices (no LDS staging, no row-end bitmap), row numbers come
//    from a wave scan of the row-end counts, row tails are joined by a wave
//    segmented scan, finished rows are stored from registers;
//  * workgroups are persistent (one per CU) and the whole LDS holds the `xc`
//    columns of x with the most nonzeros; such a column is stored in pc as
//    (slot | 1 << 31).  After pGOrder on an R-MAT these columns carry more than
//    half of the nonzeros, whose gathers become LDS reads: the kernel is bound
//    by outstanding L1 -> L2 gather requests, not by HBM bytes;
//  * the open row at a tile end goes to carry[t] / carry_row[t]; a wave-parallel
//    fix-up adds each run of carries to its row in a fixed order, so results
//    are bit-reproducible (no floating-point atomics).
#include <algorithm>

#include "dtypes.hpp"
#include "launch.hpp"
#include "ldst.hpp"

namespace mr {
namespace k {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int32_t kRowEnd = -1;
constexpr int32_t kPadItem = -2;

// the S values of a lane from `words` 32-bit words (S * sizeof(VT) bytes)
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

template <bool NT> __device__ __forceinline__ u32x4 ld_vec(const u32x4* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// the column and value words of lane `lane` in tile t
template <typename VT, int S, bool NT>
__device__ __forceinline__ void load_tile(const int32_t* __restrict__ pc, const VT* __restrict__ pv, int64_t t,
                                          int lane, uint32_t* cw, uint32_t* vw)
{
    constexpr int T = 64 * S;
    constexpr int CW = S;
    constexpr int VW = S * (int)sizeof(VT) / 4;
    const int64_t base = t * (int64_t)T + (int64_t)lane * S;
    const u32x4* cp = reinterpret_cast<const u32x4*>(pc + base);
    const u32x4* vp = reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(pv) + base * (int64_t)sizeof(VT));
#pragma unroll
    for (int q = 0; q < CW / 4; ++q) {
        const u32x4 a = ld_vec<NT>(cp + q);
        cw[4 * q] = a.x; cw[4 * q + 1] = a.y; cw[4 * q + 2] = a.z; cw[4 * q + 3] = a.w;
    }
#pragma unroll
    for (int q = 0; q < VW / 4; ++q) {
        const u32x4 b = ld_vec<NT>(vp + q);
        vw[4 * q] = b.x; vw[4 * q + 1] = b.y; vw[4 * q + 2] = b.z; vw[4 * q + 3] = b.w;
    }
}

template <typename VT, int S, int W, bool NT>
__global__ __launch_bounds__(64 * W) void merbit_lp_kernel(int64_t n_tiles, const int32_t* __restrict__ pc,
                                                           const VT* __restrict__ pv, const int32_t* __restrict__ ty,
                                                           const float* __restrict__ x, const int32_t* __restrict__ hot,
                                                           int xc, float* __restrict__ y, float* __restrict__ carry,
                                                           int32_t* __restrict__ carry_row)
{
    constexpr int CW = S;                          // column words per lane
    constexpr int VW = S * (int)sizeof(VT) / 4;    // value words per lane
    static_assert(CW % 4 == 0 && VW % 4 == 0, "lanes load whole 16-byte vectors");
    extern __shared__ float xcache[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

    for (int i = threadIdx.x; i < xc; i += 64 * W) xcache[i] = x[hot[i]];
    __syncthreads();

    for (int64_t t = (int64_t)blockIdx.x * W + wid; t < n_tiles; t += (int64_t)gridDim.x * W) {
        const int32_t y0 = ty[t], y1 = ty[t + 1];
        uint32_t cw[CW], vw[VW];
        load_tile<VT, S, NT>(pc, pv, t, lane, cw, vw);
        // x gathers into registers of their own for HBM and for LDS (one
        // register per item would make every global load wait for the LDS
        // read that last wrote it); the global ones are issued first
        float xg[S], xl[S];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            xg[j] = 0.0f;
            if ((int32_t)cw[j] >= 0) xg[j] = x[(int32_t)cw[j]];
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            xl[j] = 0.0f;
            if ((int32_t)cw[j] < kPadItem) xl[j] = xcache[cw[j] & 0x7fffffffu];
        }
        float xv[S];
#pragma unroll
        for (int j = 0; j < S; ++j) xv[j] = (int32_t)cw[j] >= 0 ? xg[j] : xl[j];
        float p[S];
#pragma unroll
        for (int j = 0; j < S; ++j) p[j] = value_at<VT>(vw, j) * xv[j];

        if (y0 == y1) {
            // every item consumes a nonzero of row y0 (fast track)
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < S; ++j) acc += p[j];
            acc = wave_sum(acc);
            if (lane == 0) {
                carry[t] = acc;
                carry_row[t] = y0;
            }
            continue;
        }

        int cnt = 0;
#pragma unroll
        for (int j = 0; j < S; ++j) cnt += ((int32_t)cw[j] == kRowEnd) ? 1 : 0;
        int inc = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        int row = inc - cnt;
        float acc = 0.0f, head = 0.0f;
        bool has_end = false;
        int head_row = 0;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if ((int32_t)cw[j] == kRowEnd) {
                if (!has_end) {
                    head = acc;
                    head_row = row;
                    has_end = true;
                } else {
                    st_stream<NT>(y + y0 + row, acc);
                }
                acc = 0.0f;
                ++row;
            } else {
                acc += p[j];
            }
        }

        // segmented scan of (tail, has_end) over the wave
        float sv = acc;
        int fl = has_end ? 1 : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const float vu = __shfl_up(sv, d, 64);
            const int fu = __shfl_up(fl, d, 64);
            if (lane >= d) {
                if (!fl) sv += vu;
                fl |= fu;
            }
        }
        const float up = __shfl_up(sv, 1, 64);
        const float carry_in = lane == 0 ? 0.0f : up;
        if (has_end) st_stream<NT>(y + y0 + head_row, carry_in + head);
        if (lane == 63) {
            carry[t] = sv;
            carry_row[t] = y1;
        }
    }
}

// one wave per window of 64 tiles: carries of one row form a run of
// consecutive tiles; the wave holding a run's first tile adds the run's sum
// (wave segmented scan inside the window, whole-window wave sums after it)
__global__ __launch_bounds__(256) void tile_fixup_kernel(int64_t n_tiles, int64_t n_rows,
                                                         const float* __restrict__ carry,
                                                         const int32_t* __restrict__ carry_row,
                                                         float* __restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t base = w * 64;
    if (base >= n_tiles) return;
    const int64_t t = base + lane;
    const bool in = t < n_tiles;
    const int32_t r = in ? carry_row[t] : -1;
    const float c = in ? carry[t] : 0.0f;
    int32_t rprev = __shfl_up(r, 1, 64);
    if (lane == 0) rprev = base > 0 ? carry_row[base - 1] : -2;
    int32_t rnext = __shfl_down(r, 1, 64);
    if (lane == 63) rnext = t + 1 < n_tiles ? carry_row[t + 1] : -3;
    if (!in) rnext = -3;

    float s = c;
    int fl = (in && r != rprev) ? 1 : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float vu = __shfl_up(s, d, 64);
        const int fu = __shfl_up(fl, d, 64);
        if (lane >= d) {
            if (!fl) s += vu;
            fl |= fu;
        }
    }
    // a run whose first tile is in this window and which continues past it
    const int32_t r63 = __shfl(r, 63, 64);
    const int fl63 = __shfl(fl, 63, 64);
    const int32_t rn63 = __shfl(rnext, 63, 64);
    float extra = 0.0f;
    if (fl63 && r63 >= 0 && r63 < n_rows && rn63 == r63) {
        for (int64_t b2 = base + 64; b2 < n_tiles; b2 += 64) {
            const int64_t t2 = b2 + lane;
            const bool same = t2 < n_tiles && carry_row[t2] == r63;
            extra += wave_sum(same ? carry[t2] : 0.0f);
            if (__ballot(same) != ~0ull) break;
        }
    }
    if (in && fl && r >= 0 && r < n_rows && r != rnext) y[r] = y[r] + s;
    if (lane == 63 && in && fl && r >= 0 && r < n_rows && r == rnext) y[r] = y[r] + s + extra;
}

// ---- SpMM on the same lane-padded tiles ------------------------------------
// Y = A X for kk <= 4*G right-hand sides.  The column stream here holds plain
// column ids (-1 row end, -2 padding).  The tiles are the SpMV tiles (64 * S
// items); a group of G lanes walks G*S consecutive items together and lane q
// of the group owns columns [4q, 4q+4): a nonzero's X row is read by the G
// lanes of its group with one 16-byte load each, which the L1 coalesces into
// one line request.  X rows are ldx floats (ldx % 4 == 0, ldx >= 4*ceil(kk/4));
// the first xr rows of X (the hub columns after pGOrder) are cached in LDS.
template <typename VT, int S, int W, int G, bool NT>
__global__ __launch_bounds__(64 * W) void merbit_lp_mm_kernel(int64_t n_tiles, const int32_t* __restrict__ pcm,
                                                              const VT* __restrict__ pv, const int32_t* __restrict__ ty,
                                                              const float* __restrict__ X, int64_t ldx, int kk, int xr,
                                                              float* __restrict__ Y, int64_t ldy,
                                                              float* __restrict__ carry, int32_t* __restrict__ carry_row)
{
    constexpr int T = 64 * S;
    constexpr int K = 4 * G;            // cached / carried columns
    constexpr int SG = S * G;           // items per group
    constexpr int VW = SG * (int)sizeof(VT) / 4;
    static_assert(SG % 4 == 0 && VW % 4 == 0, "groups load whole 16-byte vectors");
    extern __shared__ float xrows[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int grp = lane / G, q = lane % G;
    const bool own = 4 * q < kk;        // this lane has columns to compute

    for (int64_t i = threadIdx.x; i < (int64_t)xr * K; i += 64 * W) {
        const int64_t r = i / K;
        const int c = (int)(i - r * K);
        xrows[i] = c < 4 * ((kk + 3) / 4) ? X[r * ldx + c] : 0.0f;
    }
    __syncthreads();

    for (int64_t t = (int64_t)blockIdx.x * W + wid; t < n_tiles; t += (int64_t)gridDim.x * W) {
        const int32_t y0 = ty[t], y1 = ty[t + 1];
        const int64_t base = t * (int64_t)T + (int64_t)grp * SG;
        uint32_t cw[SG], vw[VW];
        const u32x4* cp = reinterpret_cast<const u32x4*>(pcm + base);
        const u32x4* vp = reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(pv) + base * (int64_t)sizeof(VT));
#pragma unroll
        for (int i = 0; i < SG / 4; ++i) {
            const u32x4 a = ld_vec<NT>(cp + i);
            cw[4 * i] = a.x; cw[4 * i + 1] = a.y; cw[4 * i + 2] = a.z; cw[4 * i + 3] = a.w;
        }
#pragma unroll
        for (int i = 0; i < VW / 4; ++i) {
            const u32x4 b = ld_vec<NT>(vp + i);
            vw[4 * i] = b.x; vw[4 * i + 1] = b.y; vw[4 * i + 2] = b.z; vw[4 * i + 3] = b.w;
        }

        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, head[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < SG; ++j) cnt += ((int32_t)cw[j] == kRowEnd) ? 1 : 0;
        int inc = cnt;
#pragma unroll
        for (int d = G; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        int row = inc - cnt;
        bool has_end = false;
        int head_row = 0;
        // chunks of CH items: the chunk's X rows are gathered before it is
        // walked -- global rows first, into registers of their own, then the
        // LDS rows -- so the gathers of a chunk are in flight together
        // instead of one per item between the row-end branches (k = 16 at
        // s22: 859 -> 834 us).  Launches of 16 waves cap a lane at 128
        // registers: their longest walks (S * G = 64) keep one item per chunk.
        constexpr int CH = (SG <= 32 || W <= 8) ? (SG < 4 ? SG : 4) : 1;
#pragma unroll
        for (int j0 = 0; j0 < SG; j0 += CH) {
            float4 xg[CH], xl[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int32_t col = (int32_t)cw[j0 + c];
                xg[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (own && col >= xr) xg[c] = *reinterpret_cast<const float4*>(X + (int64_t)col * ldx + 4 * q);
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int32_t col = (int32_t)cw[j0 + c];
                xl[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (own && col >= 0 && col < xr)
                    xl[c] = *reinterpret_cast<const float4*>(xrows + (int64_t)col * K + 4 * q);
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = j0 + c;
                const int32_t col = (int32_t)cw[j];
                if (col == kRowEnd) {
                    if (!has_end) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) head[e] = acc[e];
                        head_row = row;
                        has_end = true;
                    } else if (own) {
                        float* yr = Y + (int64_t)(y0 + row) * ldy + 4 * q;
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (4 * q + e < kk) yr[e] = acc[e];
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[e] = 0.0f;
                    ++row;
                } else if (col >= 0 && own) {
                    const float v = value_at<VT>(vw, j);
                    const float4 xv = col < xr ? xl[c] : xg[c];
                    acc[0] += v * xv.x; acc[1] += v * xv.y; acc[2] += v * xv.z; acc[3] += v * xv.w;
                }
            }
        }

        if (y0 == y1) {
            // every item consumes a nonzero of row y0 (fast track): sum the
            // groups, lanes with the same q hold the same columns
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = acc[e];
#pragma unroll
                for (int off = 32; off >= G; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane < G && 4 * q + e < kk) carry[t * K + 4 * q + e] = v;
            }
            if (lane == 0) carry_row[t] = y0;
            continue;
        }
        // segmented scan of (tails, has_end) over the groups of the wave
        int fl = has_end ? 1 : 0;
#pragma unroll
        for (int d = G; d < 64; d <<= 1) {
            const int fu = __shfl_up(fl, d, 64);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float vu = __shfl_up(acc[e], d, 64);
                if (lane >= d && !fl) acc[e] += vu;
            }
            if (lane >= d) fl |= fu;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float up = __shfl_up(acc[e], G, 64);
            head[e] += lane < G ? 0.0f : up;
        }
        if (has_end && own) {
            float* yr = Y + (int64_t)(y0 + head_row) * ldy + 4 * q;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (4 * q + e < kk) yr[e] = head[e];
        }
        if (lane >= 64 - G) {
#pragma unroll
            for (int e = 0; e < 4; ++e) carry[t * K + 4 * q + e] = acc[e];
            if (lane == 63) carry_row[t] = y1;
        }
    }
}

// k-wide version of tile_fixup_kernel: carry[t * K + c], columns c < kk
template <int K>
__global__ __launch_bounds__(256) void tile_fixup_mm_kernel(int64_t n_tiles, int64_t n_rows, int kk,
                                                            const float* __restrict__ carry,
                                                            const int32_t* __restrict__ carry_row,
                                                            float* __restrict__ Y, int64_t ldy)
{
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t base = w * 64;
    if (base >= n_tiles) return;
    const int64_t t = base + lane;
    const bool in = t < n_tiles;
    const int32_t r = in ? carry_row[t] : -1;
    int32_t rprev = __shfl_up(r, 1, 64);
    if (lane == 0) rprev = base > 0 ? carry_row[base - 1] : -2;
    int32_t rnext = __shfl_down(r, 1, 64);
    if (lane == 63) rnext = t + 1 < n_tiles ? carry_row[t + 1] : -3;
    if (!in) rnext = -3;
    const int32_t r63 = __shfl(r, 63, 64);
    const int32_t rn63 = __shfl(rnext, 63, 64);
    for (int c = 0; c < K && c < kk; ++c) {
        float s = in ? carry[t * K + c] : 0.0f;
        int fl = (in && r != rprev) ? 1 : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const float vu = __shfl_up(s, d, 64);
            const int fu = __shfl_up(fl, d, 64);
            if (lane >= d) {
                if (!fl) s += vu;
                fl |= fu;
            }
        }
        const int fl63 = __shfl(fl, 63, 64);
        float extra = 0.0f;
        if (fl63 && r63 >= 0 && r63 < n_rows && rn63 == r63) {
            for (int64_t b2 = base + 64; b2 < n_tiles; b2 += 64) {
                const int64_t t2 = b2 + lane;
                const bool same = t2 < n_tiles && carry_row[t2] == r63;
                extra += wave_sum(same ? carry[t2 * K + c] : 0.0f);
                if (__ballot(same) != ~0ull) break;
            }
        }
        if (in && fl && r >= 0 && r < n_rows && r != rnext) Y[(int64_t)r * ldy + c] = Y[(int64_t)r * ldy + c] + s;
        if (lane == 63 && in && fl && r >= 0 && r < n_rows && r == rnext)
            Y[(int64_t)r * ldy + c] = Y[(int64_t)r * ldy + c] + s + extra;
    }
}

template <typename VT, int S, int W, int G>
hipError_t launch_lp_mm(int nt, int64_t n_rows, int64_t n_tiles, const int32_t* pcm, const void* pv, const int32_t* ty,
                        const void* X, int64_t ldx, int kk, int xr, void* Y, int64_t ldy, void* carry,
                        int32_t* carry_row, int grid, hipStream_t s)
{
    constexpr int K = 4 * G;
    const size_t lds = sizeof(float) * (size_t)xr * K;
    auto kern = nt ? merbit_lp_mm_kernel<VT, S, W, G, true> : merbit_lp_mm_kernel<VT, S, W, G, false>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    kern<<<(unsigned)grid, 64 * W, lds, s>>>(n_tiles, pcm, static_cast<const VT*>(pv), ty,
                                             static_cast<const float*>(X), ldx, kk, xr, static_cast<float*>(Y), ldy,
                                             static_cast<float*>(carry), carry_row);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t waves = (n_tiles + 63) / 64;
    tile_fixup_mm_kernel<K><<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(n_tiles, n_rows, kk,
                                                                        static_cast<const float*>(carry), carry_row,
                                                                        static_cast<float*>(Y), ldy);
    return hipGetLastError();
}

// ---- plan build: the merge path as padded (column, value) streams ----------
__global__ void lp_fill_kernel(int64_t from, int64_t to, int32_t* __restrict__ pc, uint16_t* __restrict__ pv16,
                               uint32_t* __restrict__ pv32)
{
    for (int64_t i = from + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < to;
         i += (int64_t)gridDim.x * blockDim.x) {
        pc[i] = kPadItem;
        if (pv16) pv16[i] = 0;
        if (pv32) pv32[i] = 0u;
    }
}

// one wave per row: the row's nonzeros land at p = k + r, its end at rp[r+1] + r
template <typename VT>
__global__ __launch_bounds__(256) void lp_scatter_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci, const VT* __restrict__ val,
                                                         const int32_t* __restrict__ map, int32_t* __restrict__ pc,
                                                         VT* __restrict__ pv)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n_rows; r += waves) {
        const int32_t b = rp[r], e = rp[r + 1];
        for (int32_t k = b + lane; k < e; k += 64) {
            pc[(int64_t)k + r] = map ? map[ci[k]] : ci[k];
            if (pv) pv[(int64_t)k + r] = val[k];
        }
        if (lane == 0) {
            pc[(int64_t)e + r] = kRowEnd;
            if (pv) pv[(int64_t)e + r] = VT(0);
        }
    }
}

template <typename VT, int S, int W, bool NT>
hipError_t launch_lp(int64_t n_rows, int64_t n_tiles, const int32_t* pc, const void* pv, const int32_t* ty,
                     const void* x, const int32_t* hot, int xc, void* y, void* carry, int32_t* carry_row, int grid,
                     hipStream_t s)
{
    const size_t lds = merbit_lp_lds_bytes(xc);
    auto kern = merbit_lp_kernel<VT, S, W, NT>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    // persistent workgroups: as many per CU as the LDS cache and the wave
    // slots allow (one with the full cache; more when fewer columns are cached)
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 64 * W, lds) !=
            hipSuccess || per_cu < 1)
        per_cu = 1;
    const int64_t want = (int64_t)grid * per_cu;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(want, (n_tiles + W - 1) / W));
    kern<<<(unsigned)blocks, 64 * W, lds, s>>>(n_tiles, pc, static_cast<const VT*>(pv), ty,
                                             static_cast<const float*>(x), hot, xc, static_cast<float*>(y),
                                             static_cast<float*>(carry), carry_row);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t waves = (n_tiles + 63) / 64;
    tile_fixup_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(n_tiles, n_rows, static_cast<const float*>(carry),
                                                                  carry_row, static_cast<float*>(y));
    return hipGetLastError();
}

}  // namespace

int merbit_lp_mm_chunk(int kk) { return kk <= 4 ? 4 : kk <= 8 ? 8 : 16; }

// shapes whose group registers fit without spilling (16 items per lane with
// 16 waves leaves 128 VGPRs: only one lane per item fits)
// waves per workgroup of an SpMM launch on tiles of S items per lane: the
// SpMV's W, except that 16-item tiles with more than 8 columns (64-item walks
// per lane group) launch 8 waves, whose lanes hold twice the registers
int merbit_lp_mm_waves(int S, int W, int kk) { return (S == 16 && kk > 8) ? 8 : W; }
bool merbit_lp_mm_fits(int S, int W, int kk)
{
    W = merbit_lp_mm_waves(S, W, kk);
    return (S == 8 || S == 16) && (W == 8 || W == 16) && kk >= 1 && kk <= 16;
}

hipError_t merbit_lp_spmm(int dt, int S, int W, int nt, int64_t n_rows, int64_t n_tiles, const int32_t* pcm,
                          const void* pv, const int32_t* ty, const void* X, int64_t ldx, int kk, int xr, void* Y,
                          int64_t ldy, void* carry, int32_t* carry_row, int grid, hipStream_t s)
{
    if (n_tiles <= 0 || n_rows <= 0 || kk <= 0) return hipSuccess;
    const int K = merbit_lp_mm_chunk(kk);
    if (!merbit_lp_mm_fits(S, W, kk)) return hipErrorInvalidValue;
    W = merbit_lp_mm_waves(S, W, kk);
    if (dt == DT_F64 || grid <= 0 || kk > 16 || ldx < (kk + 3) / 4 * 4 || (ldx & 3) || ldy < kk ||
        (reinterpret_cast<uintptr_t>(X) & 15))
        return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        if constexpr (sizeof(typename Acc<VT>::type) == 8) {
            return hipErrorInvalidValue;
        } else {
            auto go = [&](auto stag, auto wtag) -> hipError_t {
                constexpr int SS = decltype(stag)::value;
                constexpr int WW = decltype(wtag)::value;
                if constexpr (SS * sizeof(VT) % 16 != 0 || SS % 4 != 0) {
                    return hipErrorInvalidValue;
                } else {
                    switch (K) {
                    case 4: return launch_lp_mm<VT, SS, WW, 1>(nt, n_rows, n_tiles, pcm, pv, ty, X, ldx, kk, xr, Y, ldy,
                                                               carry, carry_row, grid, s);
                    case 8: return launch_lp_mm<VT, SS, WW, 2>(nt, n_rows, n_tiles, pcm, pv, ty, X, ldx, kk, xr, Y, ldy,
                                                               carry, carry_row, grid, s);
                    default: return launch_lp_mm<VT, SS, WW, 4>(nt, n_rows, n_tiles, pcm, pv, ty, X, ldx, kk, xr, Y,
                                                                 ldy, carry, carry_row, grid, s);
                    }
                }
            };
            using I8 = std::integral_constant<int, 8>;
            using I16 = std::integral_constant<int, 16>;
            if (S == 8 && W == 16) return go(I8{}, I16{});
            if (S == 16 && W == 16) return go(I16{}, I16{});
            if (S == 8 && W == 8) return go(I8{}, I8{});
            if (S == 16 && W == 8) return go(I16{}, I8{});
            return hipErrorInvalidValue;
        }
    });
}

namespace {

__global__ void lp_map_identity_kernel(int64_t nc, int32_t* __restrict__ map)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x)
        map[i] = (int32_t)i;
}

__global__ void lp_map_hot_kernel(const int32_t* __restrict__ hot, int xc, int32_t* __restrict__ map)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < xc; i += gridDim.x * blockDim.x)
        map[hot[i]] = (int32_t)(0x80000000u | (uint32_t)i);
}

}  // namespace

hipError_t merbit_lp_map(int64_t nc, const int32_t* hot, int xc, int32_t* map, hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    int64_t g = (nc + 255) / 256;
    if (g > 16384) g = 16384;
    lp_map_identity_kernel<<<(unsigned)g, 256, 0, s>>>(nc, map);
    if (xc > 0) lp_map_hot_kernel<<<(unsigned)((xc + 255) / 256), 256, 0, s>>>(hot, xc, map);
    return hipGetLastError();
}

size_t merbit_lp_lds_bytes(int xc) { return sizeof(float) * (size_t)((xc + 63) / 64 * 64 + 64); }

hipError_t merbit_lp_build(int dt, int64_t n_rows, const int32_t* rp, const int32_t* ci, const void* val,
                           const int32_t* map, int64_t items, int64_t padded, int32_t* pc, void* pv, hipStream_t s)
{
    if (dt == DT_F64) return hipErrorInvalidValue;
    if (padded > items) {
        int64_t g = (padded - items + 255) / 256;
        if (g > 4096) g = 4096;
        lp_fill_kernel<<<(unsigned)g, 256, 0, s>>>(items, padded, pc,
                                                   (dt == DT_F32 || !pv) ? nullptr : static_cast<uint16_t*>(pv),
                                                   (dt == DT_F32 && pv) ? static_cast<uint32_t*>(pv) : nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_rows <= 0) return hipSuccess;
    int64_t g = (n_rows + 3) / 4;
    if (g > 65536) g = 65536;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        lp_scatter_kernel<VT><<<(unsigned)g, 256, 0, s>>>(n_rows, rp, ci, static_cast<const VT*>(val), map, pc,
                                                          static_cast<VT*>(pv));
        return hipGetLastError();
    });
}

hipError_t merbit_lp_spmv(int dt, int S, int W, int nt, int64_t n_rows, int64_t n_tiles, const int32_t* pc,
                          const void* pv, const int32_t* ty, const void* x, const int32_t* hot, int xc, void* y,
                          void* carry, int32_t* carry_row, int grid, hipStream_t s)
{
    if (n_tiles <= 0 || n_rows <= 0) return hipSuccess;
    if (dt == DT_F64 || grid <= 0) return hipErrorInvalidValue;
    return dispatch_dt(dt, [&](auto v) -> hipError_t {
        using VT = decltype(v);
        if constexpr (sizeof(typename Acc<VT>::type) == 8) {
            return hipErrorInvalidValue;
        } else {
            auto go = [&](auto stag, auto wtag) -> hipError_t {
                constexpr int SS = decltype(stag)::value;
                constexpr int WW = decltype(wtag)::value;
                if constexpr (SS * sizeof(VT) % 16 != 0) {
                    return hipErrorInvalidValue;
                } else {
                    return nt ? launch_lp<VT, SS, WW, true>(n_rows, n_tiles, pc, pv, ty, x, hot, xc, y, carry,
                                                            carry_row, grid, s)
                              : launch_lp<VT, SS, WW, false>(n_rows, n_tiles, pc, pv, ty, x, hot, xc, y, carry,
                                                             carry_row, grid, s);
                }
            };
            using I4 = std::integral_constant<int, 4>;
            using I8 = std::integral_constant<int, 8>;
            using I16 = std::integral_constant<int, 16>;
            if (S == 8 && W == 16) return go(I8{}, I16{});
            if (S == 16 && W == 16) return go(I16{}, I16{});
            if (S == 8 && W == 8) return go(I8{}, I8{});
            if (S == 16 && W == 8) return go(I16{}, I8{});
            if (S == 4 && W == 16) return go(I4{}, I16{});
            return hipErrorInvalidValue;
        }
    });
}

}  // namespace k
}  // namespace mr
