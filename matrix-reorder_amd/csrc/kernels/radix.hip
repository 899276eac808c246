// radix.hip -- stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Each pass sorts by one digit of up to 8 bits:
//   1. count:  per tile of 4096 elements, a 256-bin digit histogram, stored
//              digit-major (hist[d * n_tiles + tile]);
//   2. scan:   one exclusive prefix sum over that array gives every (digit,
//              tile) its first output slot;
//   3. scatter: each tile ranks its elements stably -- 16 rounds of 256
//              consecutive elements, ranks inside a wave from ballots over the
//              digit bits, wave and round offsets from LDS counters -- and
//              writes them to slot + rank.
// Equal digits keep their input order, so the passes compose into a stable
// sort by the whole key.  Integer counts only: the output does not depend on
// scheduling.
#include "dtypes.hpp"
#include "launch_perm.hpp"
#include "launch_radix.hpp"

namespace mr {
namespace k {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 16;
constexpr int kTileElems = kThreads * kRounds;   // 4096
constexpr int kBins = 256;

__global__ __launch_bounds__(kThreads) void radix_count_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                               int shift, uint32_t mask, int64_t n_tiles,
                                                               int32_t* __restrict__ hist)
{
    __shared__ int32_t cnt[kBins];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = blockIdx.x * (int64_t)kTileElems;
    for (int i = threadIdx.x; i < kTileElems; i += kThreads) {
        const int64_t q = base + i;
        if (q < n) atomicAdd(&cnt[(keys[q] >> shift) & mask], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(const uint32_t* __restrict__ kin,
                                                                 const uint32_t* __restrict__ vin, int64_t n,
                                                                 int shift, int bits, int64_t n_tiles,
                                                                 const int32_t* __restrict__ offs,
                                                                 uint32_t* __restrict__ kout,
                                                                 uint32_t* __restrict__ vout)
{
    __shared__ int32_t base_of[kBins];        // first output slot of each digit for this tile
    __shared__ int32_t wcnt[kWaves][kBins];   // per-wave digit counts of the current round
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    base_of[threadIdx.x] = offs[(int64_t)threadIdx.x * n_tiles + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t tile0 = blockIdx.x * (int64_t)kTileElems;
    for (int r = 0; r < kRounds; ++r) {
        for (int d = threadIdx.x; d < kWaves * kBins; d += kThreads) (&wcnt[0][0])[d] = 0;
        __syncthreads();   // also orders base_of updates of the previous round
        const int64_t q = tile0 + (int64_t)r * kThreads + threadIdx.x;
        const bool in = q < n;
        const uint32_t key = in ? kin[q] : 0u;
        const uint32_t val = in ? vin[q] : 0u;
        const uint32_t dig = (key >> shift) & mask;
        // lanes of this wave holding the same digit
        uint64_t same = __ballot(in);
        for (int b = 0; b < bits; ++b) {
            const uint64_t on = __ballot(in && ((dig >> b) & 1u));
            same &= ((dig >> b) & 1u) ? on : ~on;
        }
        const int rank_w = __popcll(same & lt);
        if (in && rank_w == 0) wcnt[wid][dig] = __popcll(same);
        __syncthreads();
        if (in) {
            int before = 0;
            for (int w = 0; w < wid; ++w) before += wcnt[w][dig];
            const int64_t dst = (int64_t)base_of[dig] + before + rank_w;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        // advance each digit's base by this round's count
        {
            const int d = threadIdx.x;
            int c = 0;
            for (int w = 0; w < kWaves; ++w) c += wcnt[w][d];
            base_of[d] += c;
        }
    }
}

}  // namespace

size_t radix_sort_ws_bytes(int64_t n)
{
    const int64_t n_tiles = (n + kTileElems - 1) / kTileElems;
    const int64_t h = kBins * (n_tiles > 0 ? n_tiles : 1);
    size_t b = 0;
    b += 2 * sizeof(uint32_t) * (size_t)(n > 0 ? n : 1);        // ping-pong keys / values
    b = (b + 255) / 256 * 256;
    b += sizeof(int32_t) * (size_t)h * 2;                        // histogram and offsets
    b = (b + 255) / 256 * 256;
    b += scan_workspace_bytes(h);
    return b + 256;
}

hipError_t radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, int64_t n, int key_bits, void* ws, size_t ws_bytes, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (key_bits < 1) key_bits = 1;
    if (key_bits > 32 || n > ((int64_t)1 << 31) - 1) return hipErrorInvalidValue;
    if (ws_bytes < radix_sort_ws_bytes(n)) return hipErrorInvalidValue;
    const int64_t n_tiles = (n + kTileElems - 1) / kTileElems;
    const int64_t h = kBins * n_tiles;
    char* w = static_cast<char*>(ws);
    uint32_t* tk = reinterpret_cast<uint32_t*>(w);
    uint32_t* tv = tk + n;
    size_t off = (2 * sizeof(uint32_t) * (size_t)n + 255) / 256 * 256;
    int32_t* hist = reinterpret_cast<int32_t*>(w + off);
    int32_t* offs = hist + h;
    off += (sizeof(int32_t) * (size_t)h * 2 + 255) / 256 * 256;
    void* scan_ws = w + off;
    const int passes = (key_bits + 7) / 8;
    // the last pass lands in the caller's output: odd pass counts start there
    const uint32_t* ck = keys_in;
    const uint32_t* cv = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        const int bits = (key_bits - shift) < 8 ? (key_bits - shift) : 8;
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        uint32_t* ok = to_out ? keys_out : tk;
        uint32_t* ov = to_out ? vals_out : tv;
        radix_count_kernel<<<(unsigned)n_tiles, kThreads, 0, s>>>(ck, n, shift, (1u << bits) - 1u, n_tiles, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = exclusive_scan_i32(hist, h, offs, nullptr, scan_ws, s);
        if (e != hipSuccess) return e;
        radix_scatter_kernel<<<(unsigned)n_tiles, kThreads, 0, s>>>(ck, cv, n, shift, bits, n_tiles, offs, ok, ov);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        ck = ok;
        cv = ov;
    }
    return hipSuccess;
}

}  // namespace k
}  // namespace mr
