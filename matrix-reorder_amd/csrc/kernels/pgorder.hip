// pgorder.hip -- the GPU greedy-placement loop of pGOrder on gfx950.
//
// Reference: GRAPH::gorder6 GPU part (src/class.cu:3033-3062) with
// gorder_gpu_cpu_kernel1 (:2357-2369), gorder6_kernel2 (:2897-2913),
// gorder6_kernel3 (:2917-2953) and a full CUB stable descending radix sort of
// all n (score, id) pairs every batch (:3046, 3052).
//
// score[v] is a u32: bit 31 = not yet placed, low bits = count.  Batch b places
// the |P(b)| unplaced vertices that come first by (count descending, id
// ascending) -- the order the stable radix sort gives over identity ids -- then
// every score is reset to its status bit and the window O(b) adds one per
// 2-hop path x -> y -> z (y in cols(x), z in rows(y)) to every unplaced z.
// gorder6 skips rows x and columns y of 10000 or more entries (hub_cap =
// 10000); gorder5 (:2727-2873, kernel3 :2392-2420) has no cap (hub_cap =
// INT32_MAX) and starts from column counts instead of row lengths.
//
// The reference's per-batch full sort and full reset are O(n) per batch
// (quadratic overall).  After a reset only vertices touched by the window have
// a non-zero count, so this build keeps the touched set T of the last window
// (the unplaced vertices whose count it made non-zero, listed by one scan of
// the counts) and selects exactly:
//   |T| <= 4096       : one workgroup sorts T (a register bitonic network);
//   otherwise         : a count histogram (exact below 65536) finds the
//                       threshold bin, candidates above it are collected, and
//                       the tie bin is resolved by an id histogram when it is
//                       too large to sort;
//   |T| < |P(b)|      : the rest comes from untouched unplaced vertices in
//                       ascending id order (count 0 ties), found by a
//                       workgroup scan of per-1024-id unplaced counters.
// Only T is reset between batches.  A batch is five dependent launches:
// candidates (with the threshold search), selection + placement (one
// workgroup), the window's column pass (with the reset of T's counts), its row
// pass, and the next touched list (with its count histogram).  Batch 0 starts
// from count = the start lengths (row lengths for gorder6, column counts for
// gorder5) with T = all vertices whose start length is non-zero.  The placement is identical to
// the reference's for every input (the oracle restates the full-sort loop and
// the tests compare the two bit for bit).
#include <algorithm>

#include "dtypes.hpp"
#include "launch_perm.hpp"
#include "launch_pg.hpp"
#include "launch_radix.hpp"

namespace mr {
namespace k {

namespace {

constexpr uint32_t kUnplaced = 0x80000000u;
constexpr int kExactBins = 65536;
constexpr int kBins = kExactBins + 16 * 64;
constexpr int kCoarseShift = 9;                    // 512 count bins per coarse bin
constexpr int kCoarse = (kBins + (1 << kCoarseShift) - 1) >> kCoarseShift;   // 130
constexpr int kSmallT = 4096;     // sort T directly at or below this size
constexpr int kSmallIds = 8192;   // ids sorted at once (32-bit keys in the same LDS)
constexpr int kChunk = 1024;      // ids per unplaced counter
constexpr int kGridStride = 1024; // blocks of the grid-stride kernels
constexpr int kWindowGrid = 2048; // blocks of the window-score kernel (8192 waves)
constexpr int kSureCap = 4096;    // capacity of the candidates-above-threshold list
constexpr int kTouchedPer = 16;   // ids per thread of the touched-list kernel
constexpr int kItem = 64;         // row-list entries per work item of the window's row pass

__device__ __forceinline__ int count_bin(uint32_t c)
{
    if (c < (uint32_t)kExactBins) return (int)c;
    const int e = 31 - __clz(c);              // 16..30
    const int mant = (int)((c >> (e - 6)) & 63u);
    return kExactBins + (e - 16) * 64 + mant;
}

// ascending order of this key == (count descending, id ascending)
__device__ __forceinline__ uint64_t sel_key(uint32_t count, uint32_t id)
{
    return ((uint64_t)(~count) << 32) | id;
}

// wave-aggregated append of `v` when `pred`; entries past `cap` are dropped
__device__ __forceinline__ void wave_append(bool pred, int32_t v, int32_t* list, int32_t* counter,
                                            int64_t cap = INT64_MAX)
{
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    const int64_t at = (int64_t)base + __popcll(m & ((1ull << lane) - 1ull));
    if (pred && at < cap) list[at] = v;
}

// Ascending bitonic sort of 256 * R keys held in registers by a 256-thread
// block: wave w holds elements e = w * 64R + 64r + lane in v[r].  A
// compare-exchange at distance j < 64 is a lane shuffle, 64 <= j < 64R
// swaps two registers of the same lane, and only j >= 64R (the last
// log2(4) merge levels) goes through LDS with a block barrier -- three
// barrier stages instead of one per stage of the network.  K is uint64_t
// (count, id) selection keys or uint32_t ids.
template <typename K>
__device__ __forceinline__ K cx_keep(K x, K y, bool keep_min)
{
    const bool lt = x < y;
    return keep_min ? (lt ? x : y) : (lt ? y : x);
}

template <typename K, int R, int JJ>
__device__ __forceinline__ void sort_lane_stage(K (&v)[R], int k2, int ebase)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r & JJ) continue;
        const bool up = ((ebase + 64 * r) & k2) == 0;
        const K a = v[r], b = v[r + JJ];
        const bool sw = (a > b) == up;
        v[r] = sw ? b : a;
        v[r + JJ] = sw ? a : b;
    }
}

template <typename K, int R>
__device__ void block_sort_regs(K (&v)[R], K* x)
{
    constexpr int S = 64 * R, N = 256 * R;
    const int lane = threadIdx.x & 63;
    const int ebase = (threadIdx.x >> 6) * S + lane;
    for (int k2 = 2; k2 <= N; k2 <<= 1) {
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            if (j >= S) {
                __syncthreads();   // earlier reads of x are done
#pragma unroll
                for (int r = 0; r < R; ++r) x[ebase + 64 * r] = v[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = ebase + 64 * r;
                    v[r] = cx_keep(v[r], x[e ^ j], ((e & j) == 0) == ((e & k2) == 0));
                }
            } else if (j >= 64) {
                const int jj = j >> 6;
                if constexpr (R > 1) if (jj == 1) sort_lane_stage<K, R, 1>(v, k2, ebase);
                if constexpr (R > 2) if (jj == 2) sort_lane_stage<K, R, 2>(v, k2, ebase);
                if constexpr (R > 4) if (jj == 4) sort_lane_stage<K, R, 4>(v, k2, ebase);
                if constexpr (R > 8) if (jj == 8) sort_lane_stage<K, R, 8>(v, k2, ebase);
                if constexpr (R > 16) if (jj == 16) sort_lane_stage<K, R, 16>(v, k2, ebase);
            } else {
                const bool lower = (lane & j) == 0;
                // all shuffles first: they are independent of each other
                K y[R];
#pragma unroll
                for (int r = 0; r < R; ++r) y[r] = __shfl_xor(v[r], j, 64);
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = cx_keep(v[r], y[r], lower == (((ebase + 64 * r) & k2) == 0));
            }
        }
    }
}

// sort keys[0, cnt) (cnt <= 256 R) and write the low 32 bits of the `take`
// smallest to sel[base, base + take); keys[] is the LDS exchange buffer
template <typename K, int R>
__device__ void sort_take_regs(K* keys, int cnt, int take, int32_t* sel, int base)
{
    const int ebase = (threadIdx.x >> 6) * (64 * R) + (threadIdx.x & 63);
    K v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        v[r] = e < cnt ? keys[e] : ~(K)0;
    }
    block_sort_regs<K, R>(v, keys);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        if (e < take) sel[base + e] = (int32_t)(uint32_t)v[r];
    }
}

// block-wide exclusive prefix count of a bool (256 threads); returns total
__device__ __forceinline__ int block_rank(bool f, int& rank, int* s_w)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t m = __ballot(f);
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_w[wid] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        off += (w < wid) ? s_w[w] : 0;
        tot += s_w[w];
    }
    __syncthreads();
    rank = off + in_wave;
    return tot;
}

__global__ void init_kernel(int64_t n, const int32_t* __restrict__ rp, uint32_t* __restrict__ score,
                            int32_t* __restrict__ t0, PgState* __restrict__ st, int32_t* __restrict__ chunk_left)
{
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < ((n + 63) / 64) * 64;
         v += (int64_t)gridDim.x * blockDim.x) {
        bool pred = false;
        if (v < n) {
            const int32_t len = rp[v + 1] - rp[v];
            score[v] = (uint32_t)len | kUnplaced;
            pred = len > 0;
            if ((v & (kChunk - 1)) == 0) {
                const int64_t left = n - v;
                chunk_left[v / kChunk] = (int32_t)(left < kChunk ? left : kChunk);
            }
        }
        wave_append(pred, (int32_t)v, t0, &st->tcount[0]);
    }
}

// count histogram of the first touched list: hist[0, kBins) per count bin,
// hist[kBins + c] per coarse bin c of 512 count bins (the threshold search
// reads the coarse bins first); later lists get theirs from touched_kernel
__global__ __launch_bounds__(256) void hist_kernel(const int32_t* __restrict__ T, const PgState* __restrict__ st,
                                                   int cur, const uint32_t* __restrict__ score,
                                                   int32_t* __restrict__ hist, int32_t small_t)
{
    const int32_t tn = st->tcount[cur];
    if (tn <= small_t) return;
    __shared__ int32_t h[4096];
    __shared__ int32_t hc[kCoarse];
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) h[i] = 0;
    for (int i = threadIdx.x; i < kCoarse; i += blockDim.x) hc[i] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tn; i += (int64_t)gridDim.x * blockDim.x) {
        const int b = count_bin(score[T[i]] & ~kUnplaced);
        if (b < 4096) atomicAdd(&h[b], 1);
        else atomicAdd(&hist[b], 1);
        atomicAdd(&hc[b >> kCoarseShift], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4096; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
    for (int i = threadIdx.x; i < kCoarse; i += blockDim.x)
        if (hc[i]) atomicAdd(&hist[kBins + i], hc[i]);
}

// inclusive suffix sums of one value per thread over a 256-thread block
// (thread t gets the sum over threads >= t)
__device__ __forceinline__ int64_t block_suffix(int64_t v, int64_t* s_tot)
{
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t o = __shfl_down(v, d, 64);
        if (lane + d < 64) v += o;
    }
    if (lane == 0) s_tot[wid] = v;     // wave total
    __syncthreads();
    int64_t higher = 0;
    for (int w = wid + 1; w < 4; ++w) higher += s_tot[w];
    __syncthreads();
    return v + higher;
}

// The threshold bin tb: the bins above it hold fewer than pl vertices, with it
// at least pl.  Two coalesced passes over the histogram: the coarse bins,
// then the 512 bins of the coarse bin that holds the threshold.  Every
// thread of the block returns the same (tb, above, need).
struct Thresh {
    int32_t tb, above, need;
};

__device__ Thresh find_threshold(const int32_t* __restrict__ hist, int32_t tn, int32_t pl)
{
    __shared__ int64_t s_tot[4];
    __shared__ int s_c;
    __shared__ int64_t s_above_c;
    __shared__ Thresh s_th;
    const int t = threadIdx.x;
    if (pl <= 0 || tn < pl) return Thresh{-1, tn, 0};
    if (t == 0) {
        s_c = 0;
        s_th = Thresh{-1, tn, 0};
    }
    // coarse: the highest coarse bin c with suffix(c) >= pl
    const int64_t c_cnt = t < kCoarse ? hist[kBins + t] : 0;
    const int64_t c_suf = block_suffix(c_cnt, s_tot);
    if (c_suf >= pl && c_suf - c_cnt < pl) {
        s_c = t;
        s_above_c = c_suf - c_cnt;
    }
    __syncthreads();
    const int c = s_c;
    const int64_t above_c = s_above_c;
    // fine: two bins per thread inside coarse bin c, thread t owning bins
    // lo + 2t and lo + 2t + 1
    const int lo = c << kCoarseShift;
    const int b0 = lo + 2 * t, b1 = b0 + 1;
    const int64_t h0 = b0 < kBins ? hist[b0] : 0, h1 = b1 < kBins ? hist[b1] : 0;
    const int64_t suf = block_suffix(h0 + h1, s_tot) + above_c;   // bins >= b0
    // bin b1 first (the higher count), then b0
    const int64_t above1 = suf - h0 - h1;
    if (above1 < pl && above1 + h1 >= pl)
        s_th = Thresh{b1, (int32_t)above1, (int32_t)(pl - above1)};
    else if (above1 + h1 < pl && suf >= pl)
        s_th = Thresh{b0, (int32_t)(above1 + h1), (int32_t)(pl - above1 - h1)};
    __syncthreads();
    return s_th;
}

// candidates of a touched list longer than small_t: every block finds the
// threshold bin itself from the histogram (a few KB read per block, no launch
// of its own), block 0 records it for the selection, and the vertices above
// it (sure, at most sure_cap) and in it (tie) are listed.  A list shorter than
// |P| has no threshold (tb = -1): all of it is sure.
__global__ __launch_bounds__(256) void collect_kernel(const int32_t* __restrict__ T, PgState* __restrict__ st,
                                                      int cur, const uint32_t* __restrict__ score,
                                                      int32_t* __restrict__ sure, int32_t* __restrict__ tie,
                                                      int32_t pl, const int32_t* __restrict__ hist, int32_t small_t,
                                                      int32_t sure_cap)
{
    const int32_t tn = st->tcount[cur];
    if (tn <= small_t || pl <= 0) return;
    const Thresh th = find_threshold(hist, tn, pl);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->tb = th.tb;
        st->above = th.above;
        st->need = th.need;
        // the wide tie selection starts from the count when the bin is exact
        st->kneed = th.need;
        st->khi = (th.tb >= 0 && th.tb < kExactBins) ? ~(uint32_t)th.tb : 0u;
        st->klo = 0u;
    }
    const int tb = th.tb;
    const int64_t lim = ((int64_t)tn + 63) / 64 * 64;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = 0;
        int b = -2;
        if (i < tn) {
            v = T[i];
            b = count_bin(score[v] & ~kUnplaced);
        }
        wave_append(i < tn && b > tb, v, sure, &st->nsure, sure_cap);
        wave_append(i < tn && b == tb, v, tie, &st->ntie);
    }
}

// The ids a[0, na) then b[0, cnt - na) (cnt <= 256 R), keyed by (count
// descending, id ascending) from score[], sorted, and the `take` first written
// to sel[base, base + take).  Every global load is issued unconditionally at a
// clamped, in-bounds index (a[0], b[0] and score[0] exist), so the R loads of
// each kind go out together: one workgroup waits two memory latencies, not
// two per element.
// An all-ones / all-zero mask the optimiser cannot see through.  Gating a
// loaded value with it keeps the load unconditional: a plain select on the
// value lets the compiler branch around the load instead, one memory round
// trip per element.
__device__ __forceinline__ uint32_t opaque_mask(bool f)
{
    uint32_t m = f ? 0xffffffffu : 0u;
    asm volatile("" : "+v"(m));
    return m;
}

template <int R>
__device__ void select_sorted_r(int cnt, int take, const int32_t* __restrict__ a, int na,
                                const int32_t* __restrict__ b, const uint32_t* __restrict__ score, uint64_t* x,
                                int32_t* sel, int base)
{
    const int ebase = (threadIdx.x >> 6) * (64 * R) + (threadIdx.x & 63);
    // the masks first: the empty asm statements bound the scheduler's regions
    uint32_t am[R], okm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        am[r] = opaque_mask(e < na);
        okm[r] = opaque_mask(e < cnt);
    }
    uint32_t id[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        const uint32_t ia = (uint32_t)a[e < na ? e : 0];
        const uint32_t ib = (uint32_t)b[(e >= na && e < cnt) ? e - na : 0];
        id[r] = ((ia & am[r]) | (ib & ~am[r])) & okm[r];   // 0 for padding
    }
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t sc = score[id[r]];
        // padding: count 0 and id ~0, the key ~0 (sorts last)
        v[r] = sel_key((sc & ~kUnplaced) & okm[r], id[r] | ~okm[r]);
    }
    block_sort_regs<uint64_t, R>(v, x);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        if (e < take) sel[base + e] = (int32_t)(v[r] & 0xffffffffu);
    }
}

// The `take` smallest of the ids a[0, cnt) (cnt <= 256 R), ascending, to
// sel[base, base + take): ties of one exact count are ordered by id alone, so
// 32-bit keys do (half the shuffles and compares of the 64-bit keys).
template <int R>
__device__ void select_ids_r(int cnt, int take, const int32_t* __restrict__ a, uint32_t* x, int32_t* sel, int base)
{
    const int ebase = (threadIdx.x >> 6) * (64 * R) + (threadIdx.x & 63);
    uint32_t okm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) okm[r] = opaque_mask(ebase + 64 * r < cnt);
    uint32_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        v[r] = ((uint32_t)a[e < cnt ? e : 0] & okm[r]) | ~okm[r];   // padding ~0 sorts last
    }
    block_sort_regs<uint32_t, R>(v, x);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = ebase + 64 * r;
        if (e < take) sel[base + e] = (int32_t)v[r];
    }
}

// One workgroup: final selection, fill with untouched vertices, placement,
// clearing of the per-batch scratch and the prefix of the window's row
// lengths.  Each phase issues its global loads together (a single workgroup
// is latency-bound: a chain of dependent loads per element would cost one
// memory round trip each).
__global__ __launch_bounds__(256) void final_kernel(int64_t n, const int32_t* __restrict__ T, int cur,
                                                    PgState* __restrict__ st, uint32_t* __restrict__ score,
                                                    const int32_t* __restrict__ sure, const int32_t* __restrict__ tie,
                                                    int32_t* __restrict__ chunk_left,
                                                    uint32_t* __restrict__ placement, int32_t ps, int32_t pl,
                                                    const int32_t* __restrict__ rp, int32_t os, int32_t ol,
                                                    int32_t* __restrict__ wprefix, int32_t hub_cap)
{
    __shared__ uint64_t keys[kSmallT];
    __shared__ int32_t sel[1024];
    __shared__ int32_t s_w[8];
    __shared__ int32_t s_nsel;
    __shared__ int32_t s_bins[4096];
    __shared__ int32_t s_misc[4];
    __shared__ uint64_t s_ball[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // every bookkeeping field this kernel reads, loaded once (tb is
    // meaningful only after collect_kernel ran, i.e. for a large list)
    const int32_t tn = st->tcount[cur];
    const int32_t nsure = st->nsure, ntie = st->ntie, tb = st->tb, lo_chunk0 = st->lo_chunk;
    const int32_t want = pl > 0 ? (pl < 1024 ? pl : 1024) : 0;
    if (tid == 0) {
        s_nsel = 0;
        s_misc[0] = 0;
        s_misc[1] = 0;
    }
    __syncthreads();

    // the `take` best of a[0, na) ++ b[0, cnt - na) appended to sel
    auto select_sorted = [&](int cnt, int take, const int32_t* a, int na, const int32_t* b) {
        const int base = s_nsel;
        if (cnt <= 256) select_sorted_r<1>(cnt, take, a, na, b, score, keys, sel, base);
        else if (cnt <= 512) select_sorted_r<2>(cnt, take, a, na, b, score, keys, sel, base);
        else if (cnt <= 1024) select_sorted_r<4>(cnt, take, a, na, b, score, keys, sel, base);
        else if (cnt <= 2048) select_sorted_r<8>(cnt, take, a, na, b, score, keys, sel, base);
        else select_sorted_r<16>(cnt, take, a, na, b, score, keys, sel, base);
        __syncthreads();
        if (tid == 0) s_nsel = base + take;
        __syncthreads();
    };
    // the `take` smallest ids among ids32[0, cnt) (cnt <= kSmallIds, ids
    // staged in LDS) appended to sel
    uint32_t* const ids32 = reinterpret_cast<uint32_t*>(keys);
    auto sort_take_ids = [&](int cnt, int take) {
        __syncthreads();
        const int base = s_nsel;
        if (cnt <= 256) sort_take_regs<uint32_t, 1>(ids32, cnt, take, sel, base);
        else if (cnt <= 512) sort_take_regs<uint32_t, 2>(ids32, cnt, take, sel, base);
        else if (cnt <= 1024) sort_take_regs<uint32_t, 4>(ids32, cnt, take, sel, base);
        else if (cnt <= 2048) sort_take_regs<uint32_t, 8>(ids32, cnt, take, sel, base);
        else if (cnt <= 4096) sort_take_regs<uint32_t, 16>(ids32, cnt, take, sel, base);
        else sort_take_regs<uint32_t, 32>(ids32, cnt, take, sel, base);
        __syncthreads();
        if (tid == 0) s_nsel = base + take;
        __syncthreads();
    };
    // the `take` smallest of the ids a[0, cnt) (cnt <= kSmallIds) appended to sel
    auto select_ids = [&](int cnt, int take, const int32_t* a) {
        const int base = s_nsel;
        if (cnt <= 256) select_ids_r<1>(cnt, take, a, ids32, sel, base);
        else if (cnt <= 512) select_ids_r<2>(cnt, take, a, ids32, sel, base);
        else if (cnt <= 1024) select_ids_r<4>(cnt, take, a, ids32, sel, base);
        else if (cnt <= 2048) select_ids_r<8>(cnt, take, a, ids32, sel, base);
        else if (cnt <= 4096) select_ids_r<16>(cnt, take, a, ids32, sel, base);
        else select_ids_r<32>(cnt, take, a, ids32, sel, base);
        __syncthreads();
        if (tid == 0) s_nsel = base + take;
        __syncthreads();
    };

    if (want > 0) {
        if (tn <= kSmallT) {
            select_sorted(tn, tn < want ? tn : want, T, tn, T);
        } else {
            const int32_t ns = min(nsure, kSureCap), nt = ntie;
            // every sure candidate beats every tie, so the ties fill what is
            // left: take_t = need of the threshold search
            const int32_t take_t = min(nt, max(0, want - ns));
            bool general = false;
            if (tb < kExactBins) {
                // an exact threshold bin: the ties share one count and come
                // in id order after the sorted sure candidates
                select_sorted(ns, ns, sure, ns, sure);
                if (take_t > 0 && nt <= kSmallIds) {
                    select_ids(nt, take_t, tie);
                } else if (take_t > 0) {
                    // the take_t smallest of many ids: an id histogram finds
                    // the bin holding the take_t-th, the ids up to it are
                    // sorted
                    int sh = 0;
                    while (((n - 1) >> sh) >= 4096) ++sh;
                    for (int i = tid; i < 4096; i += blockDim.x) s_bins[i] = 0;
                    __syncthreads();
                    for (int i0 = 0; i0 < nt; i0 += 8 * 256) {
                        int32_t v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int i = i0 + u * 256 + tid;
                            v[u] = tie[i < nt ? i : 0];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (i0 + u * 256 + tid < nt) atomicAdd(&s_bins[v[u] >> sh], 1);
                    }
                    __syncthreads();
                    // thread t owns bins [16t, 16t + 16): block scan of its sum
                    int own = 0;
#pragma unroll
                    for (int q = 0; q < 16; ++q) own += s_bins[16 * tid + q];
                    int incl = own;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        const int o = __shfl_up(incl, d, 64);
                        if (lane >= d) incl += o;
                    }
                    if (lane == 63) s_w[wid] = incl;
                    if (tid == 0) s_misc[1] = 0;
                    __syncthreads();
                    int before = incl - own;
                    for (int w = 0; w < wid; ++w) before += s_w[w];
                    if (before < take_t && before + own >= take_t) {
                        int run = before;
                        for (int q = 0; q < 16; ++q) {
                            if (run + s_bins[16 * tid + q] >= take_t) {
                                s_misc[0] = 16 * tid + q;
                                break;
                            }
                            run += s_bins[16 * tid + q];
                        }
                    }
                    __syncthreads();
                    const int tbin = s_misc[0];
                    // ids in bins <= tbin: fewer than take_t + 2^sh
                    for (int i0 = 0; i0 < nt; i0 += 8 * 256) {
                        int32_t v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int i = i0 + u * 256 + tid;
                            v[u] = tie[i < nt ? i : 0];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            if (i0 + u * 256 + tid < nt && (v[u] >> sh) <= tbin) {
                                const int pos = atomicAdd(&s_misc[1], 1);
                                if (pos < kSmallIds) ids32[pos] = (uint32_t)v[u];
                            }
                        }
                    }
                    __syncthreads();
                    const int cnt = s_misc[1];
                    if (cnt <= kSmallIds) sort_take_ids(cnt, take_t);
                    else general = true;
                }
            } else if (ns + nt <= kSmallT) {
                select_sorted(ns + nt, ns + nt < want ? ns + nt : want, sure, ns, tie);
            } else {
                select_sorted(ns, ns, sure, ns, sure);
                general = true;
            }
            if (general) {
                // rare: repeated minimum of the (count, id) keys over the tie list
                uint64_t last = 0;
                for (int r = 0; r < take_t; ++r) {
                    uint64_t best = ~0ull;
                    for (int i = tid; i < nt; i += blockDim.x) {
                        const int32_t v = tie[i];
                        const uint64_t kk = sel_key(score[v] & ~kUnplaced, (uint32_t)v);
                        if ((r == 0 || kk > last) && kk < best) best = kk;
                    }
                    for (int off = 32; off > 0; off >>= 1) {
                        const uint64_t o = __shfl_xor(best, off, 64);
                        best = o < best ? o : best;
                    }
                    if ((tid & 63) == 0) keys[tid >> 6] = best;
                    __syncthreads();
                    if (tid == 0) {
                        uint64_t bm = keys[0];
                        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) bm = keys[w] < bm ? keys[w] : bm;
                        sel[s_nsel] = (int32_t)(bm & 0xffffffffu);
                        s_nsel = s_nsel + 1;
                        keys[8] = bm;
                    }
                    __syncthreads();
                    last = keys[8];
                    __syncthreads();
                }
            }
        }
    }
    __syncthreads();
    // place the selected touched vertices
    const int nsel_t = s_nsel;
    for (int i = tid; i < nsel_t; i += blockDim.x) {
        const int32_t v = sel[i];
        score[v] = 0u;
        placement[ps + i] = (uint32_t)v;
        atomicSub(&chunk_left[v / kChunk], 1);
    }
    __syncthreads();
    // fill with untouched unplaced vertices in ascending id order.  Each pass
    // reads the counters of the next 256 chunks at once and scans the first
    // four that still hold unplaced vertices, one wave per chunk (16
    // consecutive ids per lane), ranked in chunk and id order.
    int missing = want - nsel_t;
    int filled = 0;
    if (missing > 0) {
        const int64_t nchunks = (n + kChunk - 1) / kChunk;
        int64_t c = lo_chunk0;
        int64_t lo_new = -1;   // first chunk with unplaced vertices
        while (missing > 0 && c < nchunks) {
            const int64_t cc = c + tid;
            const bool nz = cc < nchunks && chunk_left[cc] > 0;
            const uint64_t m = __ballot(nz);
            if (lane == 0) s_ball[wid] = m;
            __syncthreads();
            int pos[4] = {0, 0, 0, 0};
            int np = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint64_t bb = s_ball[q];
#pragma unroll
                for (int k2 = 0; k2 < 4; ++k2) {
                    const bool take_bit = bb != 0 && np < 4;
                    const int at = q * 64 + __ffsll((unsigned long long)bb) - 1;
                    if (take_bit) {
                        pos[0] = np == 0 ? at : pos[0];
                        pos[1] = np == 1 ? at : pos[1];
                        pos[2] = np == 2 ? at : pos[2];
                        pos[3] = np == 3 ? at : pos[3];
                        ++np;
                        bb &= bb - 1;
                    }
                }
            }
            if (lo_new < 0 && np > 0) lo_new = c + pos[0];
            const int my = wid == 0 ? pos[0] : wid == 1 ? pos[1] : wid == 2 ? pos[2] : pos[3];
            uint32_t flags = 0;
            int64_t idb = 0;
            if (wid < np) {
                idb = (c + my) * kChunk + lane * 16;
                uint32_t sc[16];
                if (idb + 16 <= n) {
                    const uint4* p4 = reinterpret_cast<const uint4*>(score + idb);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint4 u = p4[q];
                        sc[4 * q] = u.x;
                        sc[4 * q + 1] = u.y;
                        sc[4 * q + 2] = u.z;
                        sc[4 * q + 3] = u.w;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; ++j) sc[j] = idb + j < n ? score[idb + j] : 0u;
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) flags |= (sc[j] == kUnplaced) ? (1u << j) : 0u;
            }
            const int cnt = __popc(flags);
            int inc = cnt;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(inc, d, 64);
                if (lane >= d) inc += o;
            }
            if (lane == 63) s_w[wid] = inc;
            __syncthreads();
            int off = inc - cnt, tot = 0;
            for (int w = 0; w < 4; ++w) {
                off += w < wid ? s_w[w] : 0;
                tot += s_w[w];
            }
            int r = off;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (flags & (1u << j)) {
                    if (r < missing) sel[nsel_t + filled + r] = (int32_t)(idb + j);
                    ++r;
                }
            }
            __syncthreads();   // s_ball and s_w are rewritten by the next pass
            const int got = tot < missing ? tot : missing;
            filled += got;
            missing -= got;
            c = np == 4 ? c + pos[3] + 1 : c + 256;
        }
        if (lo_new < 0) lo_new = c < nchunks ? c : nchunks;
        if (tid == 0 && lo_new != lo_chunk0) st->lo_chunk = (int32_t)lo_new;
        for (int i = tid; i < filled; i += blockDim.x) {
            const int32_t v = sel[nsel_t + i];
            score[v] = 0u;
            placement[ps + nsel_t + i] = (uint32_t)v;
            atomicSub(&chunk_left[v / kChunk], 1);
        }
    }
    __syncthreads();
    // clear per-batch scratch for the next batch (the histogram is cleared by
    // window_y_kernel)
    if (tid == 0) {
        st->nsure = 0;
        st->ntie = 0;
        st->ycount = 0;
        st->tcount[cur ^ 1] = 0;
    }
    // prefix sums of the (x, y) pair counts of the window O(b): wprefix[w] =
    // pairs of window entries before w (rows at or above the hub cap count 0).
    // ol <= 1024: four consecutive entries per thread.  When the window is the
    // batch just placed (the steady state of the gorder6 schedule) its
    // vertices are still in sel[].
    if (ol > 0) {
        const bool from_sel = os == ps && ol <= nsel_t + filled;
        int32_t xs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = 4 * tid + j;
            const int wc = w < ol ? w : 0;
            xs[j] = from_sel ? sel[wc] : (int32_t)placement[os + wc];
        }
        int32_t mine[4];
        int32_t sum = 0;
        uint32_t okm[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)   // unplaced positions hold ~0
            okm[j] = opaque_mask(4 * tid + j < ol && (uint32_t)xs[j] < (uint32_t)n);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int32_t xi = (int32_t)((uint32_t)xs[j] & okm[j]);
            const int32_t r0 = rp[xi], r1 = rp[xi + 1];
            int32_t len = (int32_t)((uint32_t)(r1 - r0) & okm[j]);
            if (len >= hub_cap) len = 0;
            mine[j] = len;
            sum += len;
        }
        int32_t inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) s_w[wid] = inc;
        __syncthreads();
        int32_t run = inc - sum;
        for (int w = 0; w < wid; ++w) run += s_w[w];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = 4 * tid + j;
            if (w < ol) wprefix[w] = run;
            run += mine[j];
        }
        if (4 * tid < ol && 4 * tid + 4 >= ol) wprefix[ol] = run;   // the thread holding entry ol - 1
    } else if (tid == 0) {
        wprefix[0] = 0;
    }
}

// window score (gorder6_kernel3, class.cu:2917-2953).  The reference adds one
// to every unplaced z for every 2-hop path x -> y -> z (x in the window with a
// row below the hub cap, y in cols(x) with a column below the cap, z in
// rows(y)), walking rows(y) once per (x, y) pair.  The sum regroups exactly as
// count[z] += h[y] * [z in rows(y)] with h[y] = #window rows holding y, so the
// work is done in two passes: (1) every (x, y) pair adds one to h[y] -- pairs
// spread over a fixed grid by a prefix sum of the window's row lengths -- and
// when h[y] leaves 0, rows(y) is listed as work items of at most kItem
// entries; (2) one 16-lane group per item adds h[y] over its entries.
// Integer sums: identical counts, far fewer updates when window rows share
// columns (hubs).  Splitting long lists into items spreads them over the
// whole grid: a batch lists a few thousand columns, most of whose entries sit
// in lists of hundreds, so one group per list would leave the walk bound by
// the longest lists.
//
// A work item is (chunk << 32) | y: entries [chunk * kItem, (chunk + 1) * kItem)
// of the live part of rows(y).  kLds: the window's prefix sums (ol <= 1024) are
// staged in LDS; a wider window searches them in global memory.
template <bool kLds>
__global__ __launch_bounds__(256) void window_y_kernel(int64_t n, const uint32_t* __restrict__ window, int32_t ol,
                                                       const int32_t* __restrict__ wprefix,
                                                       const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                       const int32_t* __restrict__ tp, int32_t* __restrict__ h,
                                                       const int32_t* __restrict__ live, uint64_t* __restrict__ items,
                                                       PgState* __restrict__ st, int32_t hub_cap,
                                                       const int32_t* __restrict__ T, int cur,
                                                       uint32_t* __restrict__ score, int32_t* __restrict__ hist)
{
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t threads = (int64_t)gridDim.x * blockDim.x;
    const int32_t tn = st->tcount[cur];
    for (int64_t i = gtid; i < tn; i += threads) {
        const int32_t v = T[i];
        if (score[v] != 0u) score[v] = kUnplaced;
    }
    for (int64_t i = gtid; i < kBins + kCoarse; i += threads) hist[i] = 0;
    __shared__ int32_t s_lds[kLds ? 1025 : 1];
    const int32_t* s_pre = wprefix;
    if constexpr (kLds) {
        for (int i = threadIdx.x; i <= ol; i += blockDim.x) s_lds[i] = wprefix[i];
        __syncthreads();
        s_pre = s_lds;
    }
    const int32_t total = s_pre[ol];
    const int64_t lim = ((int64_t)total + 63) / 64 * 64;
    for (int64_t p = gtid; p < lim; p += threads) {
        bool first = false;
        int32_t y = 0;
        if (p < total) {
            // window entry owning pair p: last w with s_pre[w] <= p
            int lo = 0, hi = ol - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_pre[mid] <= p) lo = mid; else hi = mid - 1;
            }
            const uint32_t xu = window[lo];
            if (xu < (uint32_t)n) {   // every position of the window is placed
                y = ci[rp[xu] + (int32_t)(p - s_pre[lo])];
                if (tp[y + 1] - tp[y] < hub_cap) first = atomicAdd(&h[y], 1) == 0;
            }
        }
        // the first pair of a column lists its row list as work items (one
        // at least: the item of chunk 0 also resets h[y] after the pass)
        const int32_t nit = first ? max(1, (live[y] + kItem - 1) / kItem) : 0;
        int32_t inc = nit;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t o = __shfl_up(inc, d, 64);
            if ((int)(threadIdx.x & 63) >= d) inc += o;
        }
        const int32_t wtot = __shfl(inc, 63, 64);
        if (wtot > 0) {
            int32_t base = 0;
            if ((threadIdx.x & 63) == 0) base = atomicAdd(&st->ycount, wtot);
            base = __shfl(base, 0, 64);
            for (int32_t j = 0; j < nit; ++j)
                items[base + inc - nit + j] = ((uint64_t)(uint32_t)j << 32) | (uint32_t)y;
        }
    }
}

__global__ __launch_bounds__(256) void window_z_kernel(const int32_t* __restrict__ tp, const int32_t* __restrict__ ti,
                                                       const int32_t* __restrict__ live, const int32_t* __restrict__ h,
                                                       const uint64_t* __restrict__ items,
                                                       uint32_t* __restrict__ score, const PgState* __restrict__ st)
{
    // groups of 16 lanes, four work items per wave; rows(y) holds only
    // vertices still unplaced at the last compaction (live[y] entries)
    const int32_t nit = st->ycount;
    const int lane = threadIdx.x & 63, li = lane & 15;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 4);
    for (int64_t q0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 4; q0 < nit; q0 += groups) {
        const int64_t q = q0 + (lane >> 4);
        int32_t zb = 0, ze = 0;
        uint32_t w = 0;
        if (q < nit) {
            const uint64_t it = items[q];
            const int32_t y = (int32_t)(uint32_t)it, chunk = (int32_t)(it >> 32);
            w = (uint32_t)h[y];
            const int32_t b0 = tp[y], e0 = b0 + live[y];
            zb = b0 + chunk * kItem;
            ze = zb + kItem < e0 ? zb + kItem : e0;
        }
        for (int32_t c0 = zb; c0 < ze; c0 += 16) {
            const int32_t c = c0 + li;
            if (c < ze) {
                const int32_t z = ti[c];
                // the old value is not needed: a non-returning atomic, so the
                // wave does not wait for it (the touched list is rebuilt by
                // touched_kernel)
                if (score[z] & kUnplaced) atomicAdd(&score[z], w);
            }
        }
    }
}

// The touched list of the next batch: every unplaced vertex whose count the
// window made non-zero (all counts were 0 before the window), and its count
// histogram for the next selection (as hist_kernel builds it).  Each thread
// owns kTouchedPer consecutive ids, read once as 16-byte vectors and kept in
// registers; a block scan orders the block's entries, one atomic reserves the
// block's slice of the list (a counter shared by every wave would serialise
// tens of thousands of atomics per batch).
__device__ __forceinline__ bool is_touched(uint32_t sc) { return (sc & kUnplaced) && sc != kUnplaced; }

__global__ __launch_bounds__(256) void touched_kernel(int64_t n, const uint32_t* __restrict__ score,
                                                      int32_t* __restrict__ T, PgState* __restrict__ st, int nxt,
                                                      int32_t* __restrict__ hist, int32_t* __restrict__ hcol,
                                                      const uint64_t* __restrict__ items)
{
    // the window's column weights go back to 0 (the first item of each
    // listed column resets it; the row pass that read them has ended)
    const int32_t nit = st->ycount;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nit; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t it = items[i];
        if ((it >> 32) == 0) hcol[(int32_t)(uint32_t)it] = 0;
    }
    // a window that listed nothing changed no count: every count is back at
    // zero after the reset, so the next list is empty (its size was cleared
    // by final_kernel) and its histogram stays clear -- no scan of n counts
    if (nit == 0) return;
    __shared__ int s_w[4];
    __shared__ int s_base, s_tot;
    __shared__ int32_t h[4096];
    __shared__ int32_t hc[kCoarse];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t lo = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kTouchedPer;
    uint32_t sc[kTouchedPer];
    if (lo + kTouchedPer <= n) {
        const uint4* p4 = reinterpret_cast<const uint4*>(score + lo);   // lo is a multiple of 16 ids
#pragma unroll
        for (int q = 0; q < kTouchedPer / 4; ++q) {
            const uint4 u = p4[q];
            sc[4 * q] = u.x;
            sc[4 * q + 1] = u.y;
            sc[4 * q + 2] = u.z;
            sc[4 * q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int k2 = 0; k2 < kTouchedPer; ++k2) sc[k2] = lo + k2 < n ? score[lo + k2] : 0u;
    }
    uint32_t flags = 0;
#pragma unroll
    for (int k2 = 0; k2 < kTouchedPer; ++k2) flags |= is_touched(sc[k2]) ? (1u << k2) : 0u;
    const int cnt = __popc(flags);
    // block exclusive scan of the per-thread counts
    int inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    int off = inc - cnt;
    for (int w = 0; w < wid; ++w) off += s_w[w];
    if (threadIdx.x == 0) {
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        s_base = tot ? atomicAdd(&st->tcount[nxt], tot) : 0;
        s_tot = tot;
    }
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) h[i] = 0;
    for (int i = threadIdx.x; i < kCoarse; i += blockDim.x) hc[i] = 0;
    __syncthreads();
    if (s_tot == 0) return;   // uniform over the block
    int at = s_base + off;
#pragma unroll
    for (int k2 = 0; k2 < kTouchedPer; ++k2) {
        if (flags & (1u << k2)) {
            T[at++] = (int32_t)(lo + k2);
            const int b = count_bin(sc[k2] & ~kUnplaced);
            if (b < 4096) atomicAdd(&h[b], 1);
            else atomicAdd(&hist[b], 1);
            atomicAdd(&hc[b >> kCoarseShift], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4096; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
    for (int i = threadIdx.x; i < kCoarse; i += blockDim.x)
        if (hc[i]) atomicAdd(&hist[kBins + i], hc[i]);
}

__global__ void live_init_kernel(int64_t n, const int32_t* __restrict__ tp, int32_t* __restrict__ live)
{
    for (int64_t y = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; y < n; y += (int64_t)gridDim.x * blockDim.x)
        live[y] = tp[y + 1] - tp[y];
}

// drop placed vertices from every column's row list (a placed vertex is never
// counted again, so the counts do not change); one wave per column, order kept
__global__ __launch_bounds__(256) void compact_kernel(int64_t n, const int32_t* __restrict__ tp, int32_t* __restrict__ ti,
                                                      int32_t* __restrict__ live, const uint32_t* __restrict__ score)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t y = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); y < n; y += waves) {
        const int32_t zb = tp[y], ze = zb + live[y];
        int32_t wpos = zb;
        for (int32_t c0 = zb; c0 < ze; c0 += 64) {
            const int32_t c = c0 + lane;
            int32_t z = 0;
            bool keep = false;
            if (c < ze) {
                z = ti[c];
                keep = (score[z] & kUnplaced) != 0;
            }
            const uint64_t m = __ballot(keep);
            if (keep) ti[wpos + __popcll(m & ((1ull << lane) - 1ull))] = z;
            wpos += __popcll(m);
        }
        if (lane == 0) live[y] = wpos - zb;
    }
}

// ---- wide batches: |P| or the window above 1024 ----
//
// The selection of the |P| first (count descending, id ascending) unplaced
// vertices, spread over the grid:
//   1. collect_kernel lists the touched vertices above the threshold bin
//      (sure, fewer than |P|) and in it (tie);
//   2. when only some ties are taken, the key of the need-th smallest tie is
//      found digit by digit (11-bit digits of the 64-bit (~count, id) key,
//      high to low: a histogram pass over the ties whose key matches the
//      digits found so far, then one workgroup picks the digit).  An exact
//      threshold bin fixes the count, so only the id digits are searched;
//   3. the ties up to that key join the sure list: exactly min(|T|, |P|)
//      vertices;
//   4. two stable radix sorts (by id, then by count descending) order them;
//   5. they are placed, then the untouched rest (every unplaced vertex left,
//      all at count 0) fills the batch in ascending id order: one workgroup
//      scans the per-1024-id unplaced counters for each chunk's first slot,
//      one workgroup per chunk ranks and places its ids.
// The window's prefix sums come from a device scan of its row lengths.

constexpr int kDigitBins = 2048;   // 11-bit digits

__device__ __forceinline__ bool tie_search_runs(const PgState* st, int shift)
{
    const int32_t need = st->need, nt = st->ntie;
    if (need <= 0 || need >= nt) return false;   // no tie, or every tie, is taken
    return !(shift >= 32 && st->tb < kExactBins);  // the count of an exact bin is known
}

__global__ __launch_bounds__(256) void tie_hist_kernel(const PgState* __restrict__ st, const int32_t* __restrict__ tie,
                                                       const uint32_t* __restrict__ score, int shift, int bits,
                                                       int32_t* __restrict__ hist2)
{
    if (!tie_search_runs(st, shift)) return;
    const int32_t nt = st->ntie;
    if ((int64_t)blockIdx.x * blockDim.x >= nt) return;   // uniform over the block
    const uint64_t prefix = ((uint64_t)st->khi << 32) | st->klo;
    const int top = shift + bits;
    const int nb = 1 << bits;
    __shared__ int32_t h[kDigitBins];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nt; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = tie[i];
        const uint64_t kk = sel_key(score[v] & ~kUnplaced, (uint32_t)v);
        if (top >= 64 || (kk >> top) == (prefix >> top)) atomicAdd(&h[(int)((kk >> shift) & (uint64_t)(nb - 1))], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (h[i]) atomicAdd(&hist2[i], h[i]);
}

// one workgroup: the digit holding the kneed-th key; clears the bins
__global__ __launch_bounds__(256) void tie_pick_kernel(PgState* __restrict__ st, int shift, int bits,
                                                       int32_t* __restrict__ hist2)
{
    if (!tie_search_runs(st, shift)) return;
    __shared__ int s_w[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int nb = 1 << bits;
    const int per = (nb + 255) / 256;   // <= 8
    const int32_t kneed = st->kneed;
    int own[8];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int b = tid * per + q;
        own[q] = (q < per && b < nb) ? hist2[b] : 0;
        sum += own[q];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int b = tid * per + q;
        if (q < per && b < nb) hist2[b] = 0;
    }
    int inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();   // also: every thread has read kneed
    int before = inc - sum;
    for (int w = 0; w < wid; ++w) before += s_w[w];
    if (before < kneed && before + sum >= kneed) {
        // the first bin of this thread where the count reaches kneed
        int run = before, d = tid * per;
        bool found = false;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (!found && q < per) {
                if (run + own[q] >= kneed) {
                    found = true;
                    d = tid * per + q;
                } else {
                    run += own[q];
                }
            }
        }
        // run = keys under the prefix with a smaller digit than d
        st->kneed = kneed - run;
        if (shift >= 32) st->khi |= (uint32_t)d << (shift - 32);
        else st->klo |= (uint32_t)d << shift;
    }
}

// the ties taken join the sure list (every tie, or those up to the key found)
__global__ __launch_bounds__(256) void tie_take_kernel(PgState* __restrict__ st, const int32_t* __restrict__ tie,
                                                       const uint32_t* __restrict__ score, int32_t* __restrict__ sure,
                                                       int32_t cap)
{
    const int32_t need = st->need, nt = st->ntie;
    if (need <= 0) return;
    const uint64_t kt = need >= nt ? ~0ull : (((uint64_t)st->khi << 32) | st->klo);
    const int64_t lim = ((int64_t)nt + 63) / 64 * 64;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = 0;
        bool take = false;
        if (i < nt) {
            v = tie[i];
            take = sel_key(score[v] & ~kUnplaced, (uint32_t)v) <= kt;
        }
        wave_append(take, v, sure, &st->nsure, cap);
    }
}

// sort 1 input: (id, id) for the m selected, (pad, pad) up to pl (pad = n
// sorts after every id)
__global__ void wide_keys1_kernel(const PgState* __restrict__ st, const int32_t* __restrict__ sel, int32_t pl,
                                  uint32_t pad, uint32_t* __restrict__ k, uint32_t* __restrict__ v)
{
    const int32_t m = st->nsure;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < pl; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t id = i < m ? (uint32_t)sel[i] : pad;
        k[i] = id;
        v[i] = id;
    }
}

// sort 2 keys over the id-sorted ids: 0x7fffffff - count (31 bits, count >= 1
// for a touched vertex), the pads 0x7fffffff
__global__ void wide_keys2_kernel(const PgState* __restrict__ st, const uint32_t* __restrict__ ids, int32_t pl,
                                  const uint32_t* __restrict__ score, uint32_t* __restrict__ k)
{
    const int32_t m = st->nsure;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < pl; i += (int64_t)gridDim.x * blockDim.x)
        k[i] = i < m ? (~(score[ids[i]] & ~kUnplaced)) & 0x7fffffffu : 0x7fffffffu;
}

__global__ void wide_place_kernel(const PgState* __restrict__ st, const uint32_t* __restrict__ ids, int32_t ps,
                                  int32_t pl, uint32_t* __restrict__ score, uint32_t* __restrict__ placement,
                                  int32_t* __restrict__ chunk_left)
{
    const int32_t m = st->nsure;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m && i < pl;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = ids[i];
        score[v] = 0u;
        placement[ps + i] = v;
        atomicSub(&chunk_left[v / kChunk], 1);
    }
}

// One workgroup: records how many touched vertices were placed, clears the
// per-batch counters, and when the batch is not full, scans the unplaced
// counters from lo_chunk: coff[c - lo] = fill slot of chunk c's first unplaced
// id, up to the chunk holding the last one needed.
__global__ __launch_bounds__(256) void wide_fill_scan_kernel(int64_t n, PgState* __restrict__ st, int nxt,
                                                             int32_t pl, const int32_t* __restrict__ chunk_left,
                                                             int32_t* __restrict__ coff)
{
    __shared__ int s_w[4];
    __shared__ int64_t s_run;
    __shared__ int64_t s_last;
    __shared__ int64_t s_first;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t m = st->nsure;
    const int64_t need = (int64_t)pl - m;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    const int64_t lo = st->lo_chunk;
    if (tid == 0) {
        s_run = 0;
        s_last = -1;
        s_first = nchunks;
    }
    __syncthreads();   // every thread has read the state
    if (tid == 0) {
        if (m > pl) st->overflow = 1;   // never expected: the host reports it
        st->msel = m;
        st->nsure = 0;
        st->ntie = 0;
        st->ycount = 0;
        st->tcount[nxt] = 0;
        st->fill_lo = (int32_t)lo;
        st->fill_n = 0;
    }
    if (need <= 0) return;
    for (int64_t c0 = lo; c0 < nchunks; c0 += 256 * 16) {
        int v[16];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int64_t c = c0 + 16 * tid + j;
            v[j] = c < nchunks ? chunk_left[c] : 0;
            sum += v[j];
        }
        int inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) s_w[wid] = inc;
        __syncthreads();
        int64_t run = s_run + inc - sum;
        for (int w = 0; w < wid; ++w) run += s_w[w];
        int64_t first = nchunks;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int64_t c = c0 + 16 * tid + j;
            if (c < nchunks && run < need) {
                coff[c - lo] = (int32_t)run;
                if (run + v[j] >= need) s_last = c;
            }
            if (c < nchunks && v[j] > 0 && first == nchunks) first = c;
            run += v[j];
        }
        if (first < nchunks) atomicMin(reinterpret_cast<unsigned long long*>(&s_first), (unsigned long long)first);
        __syncthreads();
        if (tid == 0) s_run += s_w[0] + s_w[1] + s_w[2] + s_w[3];
        __syncthreads();
        if (s_last >= 0) break;   // uniform
    }
    if (tid == 0) {
        const int64_t last = s_last >= 0 ? s_last : nchunks - 1;
        st->fill_n = (int32_t)(last - lo + 1);
        st->lo_chunk = (int32_t)s_first;
    }
}

// one workgroup per chunk of the fill: its unplaced ids in ascending order to
// the slots from coff
__global__ __launch_bounds__(256) void wide_fill_place_kernel(int64_t n, const PgState* __restrict__ st, int32_t ps,
                                                              int32_t pl, const int32_t* __restrict__ coff,
                                                              uint32_t* __restrict__ score,
                                                              uint32_t* __restrict__ placement,
                                                              int32_t* __restrict__ chunk_left)
{
    __shared__ int s_w[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int32_t m = st->msel;
    const int64_t need = (int64_t)pl - m;
    if (need <= 0) return;
    const int64_t lo = st->fill_lo;
    const int32_t nc = st->fill_n;
    for (int32_t cc = blockIdx.x; cc < nc; cc += gridDim.x) {
        const int64_t base = coff[cc];   // uniform over the block
        if (base >= need) continue;
        const int64_t id0 = (lo + cc) * kChunk + 4 * tid;
        uint32_t f = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (id0 + j < n && score[id0 + j] == kUnplaced) f |= 1u << j;
        const int cnt = __popc(f);
        int inc = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) s_w[wid] = inc;
        __syncthreads();
        int64_t r = base + inc - cnt;
        for (int w = 0; w < wid; ++w) r += s_w[w];
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (f & (1u << j)) {
                if (r < need) {
                    score[id0 + j] = 0u;
                    placement[ps + m + r] = (uint32_t)(id0 + j);
                }
                ++r;
            }
        }
        if (tid == 0) {
            const int64_t placed = tot < need - base ? tot : need - base;
            atomicSub(&chunk_left[lo + cc], (int32_t)placed);
        }
        __syncthreads();   // s_w is rewritten by the next chunk
    }
}

// row lengths of the window's vertices (0 at or above the hub cap, 0 for an
// unplaced position), lens[ol] = 0 so an exclusive scan of ol + 1 entries ends
// with the total
__global__ void window_lens_kernel(int64_t n, const uint32_t* __restrict__ window, int32_t ol,
                                   const int32_t* __restrict__ rp, int32_t hub_cap, int32_t* __restrict__ lens)
{
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w <= ol; w += (int64_t)gridDim.x * blockDim.x) {
        int32_t len = 0;
        if (w < ol) {
            const uint32_t x = window[w];
            if (x < (uint32_t)n) {
                len = rp[x + 1] - rp[x];
                if (len >= hub_cap) len = 0;
            }
        }
        lens[w] = len;
    }
}

}  // namespace

hipError_t pg_init(int64_t n, const int32_t* rp, uint32_t* score, int32_t* t0, PgState* st, int32_t* chunk_left,
                   int32_t* hist, bool always_hist, hipStream_t s)
{
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    init_kernel<<<(unsigned)g, 256, 0, s>>>(n, rp, score, t0, st, chunk_left);
    hist_kernel<<<kGridStride, 256, 0, s>>>(t0, st, 0, score, hist, always_hist ? -1 : kSmallT);
    return hipGetLastError();
}

// the selection and placement of a wide batch (steps 1-5 above)
static hipError_t pg_select_wide(int64_t n, int cur, uint32_t* score, int32_t* const Tl[2], PgState* st,
                                 int32_t* sure, int32_t* tie, int32_t* hist, int32_t* chunk_left,
                                 uint32_t* placement, int32_t ps, int32_t pl, const PgWide& w, hipStream_t s)
{
    const int nxt = cur ^ 1;
    if (pl > w.cap) return hipErrorInvalidValue;
    collect_kernel<<<kGridStride, 256, 0, s>>>(Tl[cur], st, cur, score, sure, tie, pl, hist, -1, w.cap);
    // digits of the tie key, high to low: the count word (searched only for an
    // inexact bin), then the id bits below n
    int id_bits = 1;
    while (id_bits < 31 && ((n - 1) >> id_bits) != 0) ++id_bits;
    int digits[8][2], nd = 0;
    digits[nd][0] = 53, digits[nd][1] = 11, ++nd;
    digits[nd][0] = 42, digits[nd][1] = 11, ++nd;
    digits[nd][0] = 32, digits[nd][1] = 10, ++nd;
    for (int left = id_bits; left > 0;) {
        const int b = left < 11 ? left : 11;
        left -= b;
        digits[nd][0] = left, digits[nd][1] = b, ++nd;
    }
    for (int d = 0; d < nd; ++d) {
        tie_hist_kernel<<<kGridStride, 256, 0, s>>>(st, tie, score, digits[d][0], digits[d][1], w.hist2);
        tie_pick_kernel<<<1, 256, 0, s>>>(st, digits[d][0], digits[d][1], w.hist2);
    }
    tie_take_kernel<<<kGridStride, 256, 0, s>>>(st, tie, score, sure, w.cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // order: stable by id, then stable by count descending
    const unsigned gp = (unsigned)std::min<int64_t>(((int64_t)pl + 255) / 256, 4096);
    int key_bits = 1;
    while (key_bits < 32 && ((uint64_t)n >> key_bits) != 0) ++key_bits;   // ids and the pad n
    wide_keys1_kernel<<<gp, 256, 0, s>>>(st, sure, pl, (uint32_t)n, w.ka, w.va);
    e = radix_sort_pairs(w.ka, w.va, w.kb, w.vb, pl, key_bits, w.rws, w.rws_bytes, s);
    if (e != hipSuccess) return e;
    wide_keys2_kernel<<<gp, 256, 0, s>>>(st, w.vb, pl, score, w.ka);
    e = radix_sort_pairs(w.ka, w.vb, w.kb, w.va, pl, 31, w.rws, w.rws_bytes, s);
    if (e != hipSuccess) return e;
    wide_place_kernel<<<gp, 256, 0, s>>>(st, w.va, ps, pl, score, placement, chunk_left);
    wide_fill_scan_kernel<<<1, 256, 0, s>>>(n, st, nxt, pl, chunk_left, w.coff);
    wide_fill_place_kernel<<<kGridStride, 256, 0, s>>>(n, st, ps, pl, w.coff, score, placement, chunk_left);
    return hipGetLastError();
}

// Five dependent launches per batch: candidates (with the threshold search),
// the one-workgroup selection and placement, the window's column pass (with
// the reset of the batch's counts), its row pass, and the next touched list
// (with its histogram).
hipError_t pg_batch(int64_t n, int cur, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                    uint32_t* score, int32_t* const Tl[2], PgState* st, int32_t* sure, int32_t* tie, int32_t* hist,
                    int32_t* chunk_left, uint32_t* placement, int32_t* wprefix, int32_t* h, uint64_t* items,
                    int32_t* live, int32_t ps, int32_t pl, int32_t os, int32_t ol, int32_t hub_cap,
                    const PgWide* wide, hipStream_t s)
{
    const int nxt = cur ^ 1;
    if (pl <= kPgNarrow && ol <= kPgNarrow) {
        collect_kernel<<<kGridStride, 256, 0, s>>>(Tl[cur], st, cur, score, sure, tie, pl, hist, kSmallT, kSureCap);
        final_kernel<<<1, 256, 0, s>>>(n, Tl[cur], cur, st, score, sure, tie, chunk_left, placement, ps, pl, rp, os,
                                       ol, wprefix, hub_cap);
        window_y_kernel<true><<<kWindowGrid, 256, 0, s>>>(n, placement + os, ol, wprefix, rp, ci, tp, h, live, items,
                                                          st, hub_cap, Tl[cur], cur, score, hist);
    } else {
        if (!wide) return hipErrorInvalidValue;
        hipError_t e = pg_select_wide(n, cur, score, Tl, st, sure, tie, hist, chunk_left, placement, ps, pl, *wide, s);
        if (e != hipSuccess) return e;
        const unsigned gl = (unsigned)std::min<int64_t>(((int64_t)ol + 256) / 256, 4096);
        window_lens_kernel<<<gl, 256, 0, s>>>(n, placement + os, ol, rp, hub_cap, wide->lens);
        e = exclusive_scan_i32(wide->lens, (int64_t)ol + 1, wprefix, nullptr, wide->sws, s);
        if (e != hipSuccess) return e;
        window_y_kernel<false><<<kWindowGrid, 256, 0, s>>>(n, placement + os, ol, wprefix, rp, ci, tp, h, live, items,
                                                           st, hub_cap, Tl[cur], cur, score, hist);
    }
    if (ol > 0) {
        window_z_kernel<<<kWindowGrid, 256, 0, s>>>(tp, ti, live, h, items, score, st);
        const int64_t per_block = 256 * (int64_t)kTouchedPer;
        const int64_t g = n > 0 ? (n + per_block - 1) / per_block : 1;
        touched_kernel<<<(unsigned)g, 256, 0, s>>>(n, score, Tl[nxt], st, nxt, hist, h, items);
    }
    return hipGetLastError();
}

hipError_t pg_compact(int64_t n, const int32_t* tp, int32_t* ti, int32_t* live, const uint32_t* score,
                      hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 3) / 4;
    if (g > 8192) g = 8192;
    compact_kernel<<<(unsigned)g, 256, 0, s>>>(n, tp, ti, live, score);
    return hipGetLastError();
}

hipError_t pg_live_init(int64_t n, const int32_t* tp, int32_t* live, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    live_init_kernel<<<(unsigned)g, 256, 0, s>>>(n, tp, live);
    return hipGetLastError();
}

int pg_hist_bins() { return kBins + kCoarse; }

int64_t pg_item_capacity(int64_t n, int64_t nnz) { return n + nnz / kItem + 64; }

}  // namespace k
}  // namespace mr
