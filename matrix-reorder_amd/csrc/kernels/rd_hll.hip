// rd_hll.hip -- sampled reuse distance with HyperLogLog registers in LDS.
//
// Reference: sampled_hll_rd_block_kernel (src/class.cu:1014-1131).  One
// workgroup per sampled position sp = 1 + hash(seed ^ (sid*0x9e3779b9 + 17)) %
// (nnz-1) of the x-access stream; it scans backwards for the previous access
// to the same sector and inserts every intervening sector into 2^p HLL
// registers (register = low p hash bits, rank = leading zeros of the rest).
// Status: 0 found, 1 cold (reached position 0), 2 estimate above 1.03*rd_cap at
// a check point (every 16 chunks of 256 positions), 3 max_scan reached.
//
// gfx950 shape: 256 lanes scan 4096 positions per step (16 per lane, loads
// issued together), which is exactly the reference's 16-chunk check interval,
// so the check points coincide; the inserted set (positions after the nearest
// previous occurrence) does not depend on the chunking.  The reference
// recomputes the float register sum serially on one thread at every check;
// here the exact sum S = sum 2^(R - r_i) (R = max rank) is maintained with a
// 64-bit LDS atomic on every register increase, the check is decided from S
// when it is farther than 1e-3 (relative) from the threshold -- the serial
// float sum of 2^p terms differs from the exact one by < 2^p * 2^-24 -- and
// only otherwise is the serial float sum formed.  The register file, the
// found position and the status are integers and identical to the
// reference's; the serial float sum of the final registers is returned so the
// host forms the estimate with the reference's float formula.
#include "dtypes.hpp"
#include "launch_rd.hpp"

namespace mr {
namespace k {

namespace {

__device__ __forceinline__ uint32_t hll_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ int hll_rank(uint32_t h, int p)
{
    const uint32_t w = h >> p;
    const int max_rank = 32 - p + 1;
    if (w == 0u) return max_rank;
    int r = __clz(w) - p + 1;
    if (r < 1) r = 1;
    if (r > max_rank) r = max_rank;
    return r;
}

// correctly rounded float quotient: the double quotient of two floats rounds
// to the same float as an exact IEEE float division (53 >= 2*24 + 2)
__device__ __forceinline__ float div_rn(float a, float b) { return (float)((double)a / (double)b); }

__device__ __forceinline__ float hll_alpha(int m)
{
    if (m == 16) return 0.673f;
    if (m == 32) return 0.697f;
    if (m == 64) return 0.709f;
    return div_rn(0.7213f, 1.0f + div_rn(1.079f, (float)m));
}

constexpr int kThreads = 256;
constexpr int kPerLane = 16;
constexpr int kStep = kThreads * kPerLane;   // 4096 positions per step

__global__ __launch_bounds__(kThreads) void hll_kernel(const int32_t* __restrict__ stream, int64_t nnz,
                                                       int32_t samples, int32_t sector, uint32_t seed,
                                                       int32_t rd_cap, int32_t max_scan, int32_t p,
                                                       float* __restrict__ out_sum, int32_t* __restrict__ out_zeros,
                                                       float* __restrict__ out_est, int32_t* __restrict__ out_status,
                                                       int32_t* __restrict__ out_pos)
{
#pragma clang fp contract(off)
    extern __shared__ int32_t regs[];
    __shared__ unsigned long long s_sum;
    __shared__ int s_found;
    __shared__ int s_flag;
    __shared__ float s_est;
    const int sid = blockIdx.x;
    if (sid >= samples || nnz <= 1) return;
    const int m = 1 << p;
    const int R = 32 - p + 1;
    const int tid = threadIdx.x;
    for (int i = tid; i < m; i += kThreads) regs[i] = 0;
    if (tid == 0) s_sum = (unsigned long long)m << R;
    const uint32_t r0 = hll_hash(seed ^ (uint32_t)((uint32_t)sid * 0x9e3779b9U + 17U));
    const int64_t sp = 1 + (int64_t)(r0 % (uint32_t)(nnz - 1));
    const int64_t lo = (max_scan > 0) ? (sp - max_scan > 0 ? sp - max_scan : 0) : 0;
    const int32_t target = stream[sp] / sector;
    const float thr = (float)rd_cap * 1.03f;
    const float alpha = hll_alpha(m);
    const float fm = (float)m;
    __syncthreads();

    // serial float estimate exactly as the reference forms it
    auto serial = [&](float& sum, int& zeros) {
        sum = 0.0f;
        zeros = 0;
        for (int i = 0; i < m; ++i) {
            const int r = regs[i];
            if (r == 0) ++zeros;
            sum += __builtin_ldexpf(1.0f, -r);
        }
    };
    int status = -99;
    float est = 0.0f;
    int64_t chunks_done = 0;   // reference chunks of 256 positions
    for (int64_t hi = sp - 1; hi >= lo && status == -99; hi -= kStep) {
        if (tid == 0) s_found = -1;
        __syncthreads();
        int32_t secs[kPerLane];
        int64_t idxs[kPerLane];
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            idxs[u] = hi - (int64_t)(u * kThreads + tid);
            secs[u] = idxs[u] >= lo ? stream[idxs[u]] / sector : -1;
        }
        int local_found = -1;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u)
            if (idxs[u] >= lo && secs[u] == target && (int)(hi - idxs[u]) >= 0) {
                const int off = (int)(hi - idxs[u]);   // distance from hi: smaller = nearer
                if (local_found < 0 || off < local_found) local_found = off;
            }
        if (local_found >= 0) atomicMin((unsigned int*)&s_found, (unsigned int)local_found);
        __syncthreads();
        const int found_off = s_found;   // -1 as unsigned is max: min() keeps real offsets
        const bool found = found_off >= 0;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            const int off = (int)(hi - idxs[u]);
            if (idxs[u] >= lo && secs[u] != target && (!found || off < found_off)) {
                const uint32_t h = hll_hash((uint32_t)secs[u]);
                const int rid = (int)(h & (uint32_t)(m - 1));
                const int rk = hll_rank(h, p);
                if (regs[rid] < rk) {
                    const int old = atomicMax(&regs[rid], rk);
                    if (old < rk)
                        atomicAdd(&s_sum, (1ull << (R - rk)) - (1ull << (R - old)));
                }
            }
        }
        __syncthreads();
        // chunks of 256 the reference has executed by the end of this step
        int64_t last_small = (hi - lo) / kThreads;   // index of the last small chunk inside [lo, hi]
        if (last_small > kPerLane - 1) last_small = kPerLane - 1;
        chunks_done += last_small + 1;
        if (found) {
            status = 0;
        } else if (rd_cap > 0 && last_small == kPerLane - 1) {
            // check point: chunks_done is a multiple of 16 here
            const double exact = (double)s_sum / (double)(1ull << R);
            const double rough = (double)alpha * (double)fm * (double)fm / exact;
            if (rough > (double)thr * 1.001) {
                status = 2;
            } else if (rough >= (double)thr * 0.999) {
                if (tid == 0) {
                    float sum;
                    int zeros;
                    serial(sum, zeros);
                    float e = div_rn(alpha * fm * fm, sum);
                    if (e <= 2.5f * fm && zeros > 0) e = fm * logf(div_rn(fm, (float)zeros));
                    s_flag = e > thr ? 1 : 0;
                }
                __syncthreads();
                if (s_flag) status = 2;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (status == -99) status = (lo == 0) ? 1 : 3;
        float sum;
        int zeros;
        serial(sum, zeros);
        if (status == 2) {
            est = div_rn(alpha * fm * fm, sum);
            if (est <= 2.5f * fm && zeros > 0) est = fm * logf(div_rn(fm, (float)zeros));
        }
        s_est = est;
        out_sum[sid] = sum;
        out_zeros[sid] = zeros;
        out_est[sid] = est;
        out_status[sid] = status;
        out_pos[sid] = (int32_t)sp;
    }
    (void)chunks_done;
}

}  // namespace

hipError_t hll_sample(const int32_t* stream, int64_t nnz, int32_t samples, int32_t sector, uint32_t seed,
                      int32_t rd_cap, int32_t max_scan, int32_t p, float* out_sum, int32_t* out_zeros,
                      float* out_est, int32_t* out_status, int32_t* out_pos, hipStream_t s)
{
    if (samples <= 0 || nnz <= 1) return hipSuccess;
    const size_t smem = sizeof(int32_t) * ((size_t)1 << p);
    hll_kernel<<<(unsigned)samples, kThreads, smem, s>>>(stream, nnz, samples, sector, seed, rd_cap, max_scan, p,
                                                         out_sum, out_zeros, out_est, out_status, out_pos);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
