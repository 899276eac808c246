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
// (appended on the 0 -> 1 transition of a count) and selects exactly:
//   |T| <= 4096       : one workgroup sorts T in LDS;
//   otherwise         : a count histogram (exact below 65536) finds the
//                       threshold bin, candidates above it are collected, and
//                       the tie bin is resolved by an id histogram when it is
//                       too large to sort;
//   |T| < |P(b)|      : the rest comes from untouched unplaced vertices in
//                       ascending id order (count 0 ties), found by a
//                       workgroup scan of per-1024-id unplaced counters.
// Only T is reset between batches.  Batch 0 starts from count = the start
// lengths (row lengths for gorder6, column counts for gorder5) with T = all
// vertices whose start length is non-zero.  The placement is identical to
// the reference's for every input (the oracle restates the full-sort loop and
// the tests compare the two bit for bit).
#include "dtypes.hpp"
#include "launch_pg.hpp"

namespace mr {
namespace k {

namespace {

constexpr uint32_t kUnplaced = 0x80000000u;
constexpr int kExactBins = 65536;
constexpr int kBins = kExactBins + 16 * 64;
constexpr int kCoarseShift = 9;                    // 512 count bins per coarse bin
constexpr int kCoarse = (kBins + (1 << kCoarseShift) - 1) >> kCoarseShift;   // 130
constexpr int kSmallT = 4096;     // sort T directly at or below this size
constexpr int kChunk = 1024;      // ids per unplaced counter
constexpr int kGridStride = 1024; // blocks of the grid-stride kernels
constexpr int kWindowGrid = 2048; // blocks of the window-score kernel (8192 waves)
constexpr int kSureCap = 4096;    // capacity of the candidates-above-threshold list

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

__device__ void lds_bitonic_asc(uint64_t* a, int P)
{
    for (int k2 = 2; k2 <= P; k2 <<= 1) {
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P / 2; i += blockDim.x) {
                const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const int hi = lo | j;
                const uint64_t x = a[lo], y = a[hi];
                const bool up = (lo & k2) == 0;
                if ((x > y) == up) { a[lo] = y; a[hi] = x; }
            }
            __syncthreads();
        }
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

// count histogram of T: hist[0, kBins) per count bin, hist[kBins + c] per
// coarse bin c of 512 count bins (the threshold search reads the coarse bins
// first)
__global__ __launch_bounds__(256) void hist_kernel(const int32_t* __restrict__ T, const PgState* __restrict__ st,
                                                   int cur, const uint32_t* __restrict__ score,
                                                   int32_t* __restrict__ hist)
{
    const int32_t tn = st->tcount[cur];
    if (tn <= kSmallT) return;
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

// the threshold bin tb: the bins above it hold fewer than pl vertices, with it
// at least pl.  Two coalesced passes: the coarse bins, then the 512 bins of
// the coarse bin that holds the threshold.
__global__ __launch_bounds__(256) void thresh_kernel(const int32_t* __restrict__ hist, PgState* __restrict__ st,
                                                     int cur, int32_t pl)
{
    const int32_t tn = st->tcount[cur];
    if (tn <= kSmallT) return;
    __shared__ int64_t s_tot[4];
    __shared__ int s_c;
    __shared__ int64_t s_above_c;
    const int t = threadIdx.x;
    if (pl <= 0 || tn < pl) {
        if (t == 0) { st->tb = -1; st->above = tn; st->need = 0; }
        return;
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
    if (above1 < pl && above1 + h1 >= pl) {
        st->tb = b1;
        st->above = (int32_t)above1;
        st->need = (int32_t)(pl - above1);
    } else if (above1 + h1 < pl && suf >= pl) {
        st->tb = b0;
        st->above = (int32_t)(above1 + h1);
        st->need = (int32_t)(pl - above1 - h1);
    }
}

__global__ __launch_bounds__(256) void collect_kernel(const int32_t* __restrict__ T, PgState* __restrict__ st,
                                                      int cur, const uint32_t* __restrict__ score,
                                                      int32_t* __restrict__ sure, int32_t* __restrict__ tie,
                                                      int32_t pl, int32_t* __restrict__ hist)
{
    const int32_t tn = st->tcount[cur];
    if (tn <= kSmallT) return;
    // the histogram was consumed by thresh_kernel: clear it for the next batch
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < kBins + kCoarse;
         i += (int64_t)gridDim.x * blockDim.x)
        hist[i] = 0;
    if (pl <= 0) return;
    const int tb = st->tb;
    const int64_t lim = ((int64_t)tn + 63) / 64 * 64;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = 0;
        int b = -2;
        if (i < tn) {
            v = T[i];
            b = count_bin(score[v] & ~kUnplaced);
        }
        wave_append(i < tn && b > tb, v, sure, &st->nsure, kSureCap);
        wave_append(i < tn && b == tb, v, tie, &st->ntie);
    }
}

// one workgroup: final selection, fill with untouched vertices, placement,
// and clearing of the per-batch scratch
__global__ __launch_bounds__(256) void final_kernel(int64_t n, const int32_t* __restrict__ T, int cur,
                                                    PgState* __restrict__ st, uint32_t* __restrict__ score,
                                                    const int32_t* __restrict__ sure, const int32_t* __restrict__ tie,
                                                    int32_t* __restrict__ hist, int32_t* __restrict__ chunk_left,
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
    const int tid = threadIdx.x;
    const int32_t tn = st->tcount[cur];
    const int32_t want = pl > 0 ? (pl < 1024 ? pl : 1024) : 0;
    if (tid == 0) {
        s_nsel = 0;
        s_misc[0] = 0;
        s_misc[1] = 0;
    }
    __syncthreads();

    auto sort_take = [&](int cnt, int take) {
        int P = 2;
        while (P < cnt) P <<= 1;
        for (int i = cnt + tid; i < P; i += blockDim.x) keys[i] = ~0ull;
        __syncthreads();
        lds_bitonic_asc(keys, P);
        const int base = s_nsel;
        for (int i = tid; i < take; i += blockDim.x) sel[base + i] = (int32_t)(keys[i] & 0xffffffffu);
        __syncthreads();
        if (tid == 0) s_nsel = base + take;
        __syncthreads();
    };

    if (want > 0) {
        if (tn <= kSmallT) {
            for (int i = tid; i < tn; i += blockDim.x) {
                const int32_t v = T[i];
                keys[i] = sel_key(score[v] & ~kUnplaced, (uint32_t)v);
            }
            __syncthreads();
            sort_take(tn, tn < want ? tn : want);
        } else {
            const int32_t ns = min(st->nsure, kSureCap), nt = st->ntie, need = st->need;
            if (ns + nt <= kSmallT) {
                for (int i = tid; i < ns + nt; i += blockDim.x) {
                    const int32_t v = i < ns ? sure[i] : tie[i - ns];
                    keys[i] = sel_key(score[v] & ~kUnplaced, (uint32_t)v);
                }
                __syncthreads();
                const int take = ns + nt < want ? ns + nt : want;
                sort_take(ns + nt, take);
            } else {
                // all sure candidates first (they beat every tie)
                for (int i = tid; i < ns; i += blockDim.x) {
                    const int32_t v = sure[i];
                    keys[i] = sel_key(score[v] & ~kUnplaced, (uint32_t)v);
                }
                __syncthreads();
                sort_take(ns, ns);
                // the `need` best ties
                if (st->tb < kExactBins) {
                    // equal counts: the `need` smallest ids, via an id histogram
                    int sh = 0;
                    while (((n - 1) >> sh) >= 4096) ++sh;
                    for (int i = tid; i < 4096; i += blockDim.x) s_bins[i] = 0;
                    __syncthreads();
                    for (int i = tid; i < nt; i += blockDim.x) atomicAdd(&s_bins[tie[i] >> sh], 1);
                    __syncthreads();
                    if (tid == 0) {
                        int run = 0, tbin = 4095;
                        for (int bb = 0; bb < 4096; ++bb) {
                            if (run + s_bins[bb] >= need) { tbin = bb; break; }
                            run += s_bins[bb];
                        }
                        s_misc[0] = tbin;
                        s_misc[1] = 0;
                    }
                    __syncthreads();
                    const int tbin = s_misc[0];
                    // ids in bins <= tbin: fewer than need + 2^sh; sort them by id
                    for (int i = tid; i < nt; i += blockDim.x) {
                        const int32_t v = tie[i];
                        if ((v >> sh) <= tbin) {
                            const int pos = atomicAdd(&s_misc[1], 1);
                            if (pos < kSmallT) keys[pos] = (uint64_t)(uint32_t)v;
                        }
                    }
                    __syncthreads();
                    const int cnt = s_misc[1];
                    if (cnt <= kSmallT) sort_take(cnt, need);
                }
                if (st->tb >= kExactBins || (st->tb < kExactBins && s_misc[1] > kSmallT)) {
                    // rare general case: repeated maximum over the tie list
                    uint64_t last = 0;
                    for (int r = 0; r < need; ++r) {
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
    // fill with untouched unplaced vertices in ascending id order
    int missing = want - nsel_t;
    if (missing > 0) {
        const int64_t nchunks = (n + kChunk - 1) / kChunk;
        int64_t c = st->lo_chunk;
        int filled = 0;
        while (missing > 0 && c < nchunks) {
            const int left = chunk_left[c];
            __syncthreads();
            if (left == 0) {
                if (tid == 0 && st->lo_chunk == c) st->lo_chunk = (int32_t)(c + 1);
                __syncthreads();
                ++c;
                continue;
            }
            const int64_t base = c * kChunk;
            for (int sub = 0; sub < kChunk && missing > 0; sub += 256) {
                const int64_t id = base + sub + tid;
                const bool f = id < n && score[id] == kUnplaced;
                int rank = 0;
                const int tot = block_rank(f, rank, s_w);
                if (f && rank < missing) sel[nsel_t + filled + rank] = (int32_t)id;
                const int got = tot < missing ? tot : missing;
                filled += got;
                missing -= got;
            }
            ++c;
        }
        __syncthreads();
        for (int i = tid; i < filled; i += blockDim.x) {
            const int32_t v = sel[nsel_t + i];
            score[v] = 0u;
            placement[ps + nsel_t + i] = (uint32_t)v;
            atomicSub(&chunk_left[v / kChunk], 1);
        }
    }
    __syncthreads();
    // clear per-batch scratch for the next batch (the histogram is cleared by
    // collect_kernel)
    if (tid == 0) {
        st->nsure = 0;
        st->ntie = 0;
        st->ycount = 0;
        st->tcount[cur ^ 1] = 0;
    }
    // prefix sums of the (x, y) pair counts of the window O(b): wprefix[w] =
    // pairs of window entries before w (rows at or above the hub cap count 0)
    __syncthreads();
    if (ol > 0) {
        int32_t* s_len = s_bins;   // reuse: ol <= 1024 entries
        for (int w = tid; w < ol; w += blockDim.x) {
            const uint32_t xu = placement[os + w];
            int32_t len = 0;
            if (xu < (uint32_t)n) {
                len = rp[xu + 1] - rp[xu];
                if (len >= hub_cap) len = 0;
            }
            s_len[w] = len;
        }
        __syncthreads();
        // exclusive prefix over ol <= 1024 lengths: four consecutive entries
        // per thread, then a block scan of the per-thread sums
        int32_t mine[4];
        int32_t sum = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = 4 * tid + j;
            mine[j] = w < ol ? s_len[w] : 0;
            sum += mine[j];
        }
        const int lane = tid & 63, wid = tid >> 6;
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

__global__ void reset_kernel(const int32_t* __restrict__ T, const PgState* __restrict__ st, int cur,
                             uint32_t* __restrict__ score)
{
    const int32_t tn = st->tcount[cur];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tn; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = T[i];
        if (score[v] != 0u) score[v] = kUnplaced;
    }
}

// window score (gorder6_kernel3, class.cu:2917-2953).  The reference adds one
// to every unplaced z for every 2-hop path x -> y -> z (x in the window with a
// row below the hub cap, y in cols(x) with a column below the cap, z in
// rows(y)), walking rows(y) once per (x, y) pair.  The sum regroups exactly as
// count[z] += h[y] * [z in rows(y)] with h[y] = #window rows holding y, so the
// work is done in two passes: (1) every (x, y) pair adds one to h[y] -- pairs
// spread over a fixed grid by a prefix sum of the window's row lengths -- and
// lists y when h[y] leaves 0; (2) one wave per listed y walks rows(y) once and
// adds h[y].  Integer sums: identical counts, far fewer updates when window
// rows share columns (hubs).
__global__ __launch_bounds__(256) void window_y_kernel(int64_t n, const uint32_t* __restrict__ window, int32_t ol,
                                                       const int32_t* __restrict__ wprefix,
                                                       const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                       const int32_t* __restrict__ tp, int32_t* __restrict__ h,
                                                       int32_t* __restrict__ ylist, PgState* __restrict__ st,
                                                       int32_t hub_cap)
{
    __shared__ int32_t s_pre[1025];
    for (int i = threadIdx.x; i <= ol; i += blockDim.x) s_pre[i] = wprefix[i];
    __syncthreads();
    const int32_t total = s_pre[ol];
    const int64_t threads = (int64_t)gridDim.x * blockDim.x;
    const int64_t lim = ((int64_t)total + 63) / 64 * 64;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < lim; p += threads) {
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
        wave_append(first, y, ylist, &st->ycount);
    }
}

__global__ __launch_bounds__(256) void window_z_kernel(const int32_t* __restrict__ tp, const int32_t* __restrict__ ti,
                                                       const int32_t* __restrict__ live,
                                                       int32_t* __restrict__ h, const int32_t* __restrict__ ylist,
                                                       uint32_t* __restrict__ score, const PgState* __restrict__ st)
{
    // groups of 16 lanes, four listed columns per wave; rows(y) holds only
    // vertices still unplaced at the last compaction (live[y] entries)
    const int32_t ny = st->ycount;
    const int lane = threadIdx.x & 63, li = lane & 15;
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x >> 4);
    for (int64_t q0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 4; q0 < ny; q0 += groups) {
        const int64_t q = q0 + (lane >> 4);
        int32_t zb = 0, ze = 0, y = 0;
        uint32_t w = 0;
        if (q < ny) {
            y = ylist[q];
            w = (uint32_t)h[y];
            zb = tp[y];
            ze = zb + live[y];
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
        if (q < ny && li == 0) h[y] = 0;   // y is listed once: back to 0 for the next window
    }
}

// the touched list of the next batch: every unplaced vertex whose count the
// window made non-zero (all counts were 0 before the window).  Each block
// owns a contiguous id range, counts it, reserves its slice of the list with
// one atomic, then writes it in id order (a counter shared by every wave
// would serialise tens of thousands of atomics per batch).
__device__ __forceinline__ bool is_touched(uint32_t sc) { return (sc & kUnplaced) && sc != kUnplaced; }

__global__ __launch_bounds__(256) void touched_kernel(int64_t n, const uint32_t* __restrict__ score,
                                                      int32_t* __restrict__ T, PgState* __restrict__ st, int nxt)
{
    __shared__ int s_w[4];
    __shared__ int s_base;
    const int64_t per = ((n + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < n ? lo + per : n;
    if (lo >= hi) return;                       // uniform over the block
    int cnt = 0;
    for (int64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) cnt += is_touched(score[v]) ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        s_base = tot ? atomicAdd(&st->tcount[nxt], tot) : 0;
    }
    __syncthreads();
    int run = s_base;
    for (int64_t c = lo; c < hi; c += blockDim.x) {
        const int64_t v = c + threadIdx.x;
        const bool f = v < hi && is_touched(score[v]);
        int rank = 0;
        const int tot = block_rank(f, rank, s_w);
        if (f) T[run + rank] = (int32_t)v;
        run += tot;
    }
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

}  // namespace

hipError_t pg_init(int64_t n, const int32_t* rp, uint32_t* score, int32_t* t0, PgState* st, int32_t* chunk_left,
                   hipStream_t s)
{
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    init_kernel<<<(unsigned)g, 256, 0, s>>>(n, rp, score, t0, st, chunk_left);
    return hipGetLastError();
}

hipError_t pg_batch(int64_t n, int cur, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                    uint32_t* score, int32_t* const Tl[2], PgState* st, int32_t* sure, int32_t* tie, int32_t* hist,
                    int32_t* chunk_left, uint32_t* placement, int32_t* wprefix, int32_t* h, int32_t* ylist,
                    int32_t* live, int32_t ps, int32_t pl, int32_t os, int32_t ol, int32_t hub_cap, hipStream_t s)
{
    const int nxt = cur ^ 1;
    hist_kernel<<<kGridStride, 256, 0, s>>>(Tl[cur], st, cur, score, hist);
    thresh_kernel<<<1, 256, 0, s>>>(hist, st, cur, pl);
    collect_kernel<<<kGridStride, 256, 0, s>>>(Tl[cur], st, cur, score, sure, tie, pl, hist);
    final_kernel<<<1, 256, 0, s>>>(n, Tl[cur], cur, st, score, sure, tie, hist, chunk_left, placement, ps, pl, rp,
                                   os, ol, wprefix, hub_cap);
    reset_kernel<<<kGridStride, 256, 0, s>>>(Tl[cur], st, cur, score);
    if (ol > 0) {
        window_y_kernel<<<kWindowGrid, 256, 0, s>>>(n, placement + os, ol, wprefix, rp, ci, tp, h, ylist, st,
                                                    hub_cap);
        window_z_kernel<<<kWindowGrid, 256, 0, s>>>(tp, ti, live, h, ylist, score, st);
        touched_kernel<<<kGridStride, 256, 0, s>>>(n, score, Tl[nxt], st, nxt);
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

}  // namespace k
}  // namespace mr
