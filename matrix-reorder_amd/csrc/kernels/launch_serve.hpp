// launch_serve.hpp -- device helpers of the serving loop (serve.hip) and the
// host/device formulas they share.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace mr {
namespace k {

constexpr int kMaxServeReqs = 1024;

// one request of a batch: its slice of the X / Y arenas (row-major with
// leading dimension ldx / ldy, k columns) and its seed
struct ServeReq {
    void* x;
    void* y;
    int64_t ldx, ldy;
    int64_t n_cols, n_rows;
    int32_t k, pad;
    uint64_t seed;
};

// right-hand side element j = i * k + c of a request: a SplitMix64 output of
// (seed, j), top 24 bits scaled to [-1, 1) (exact in float, so the host and
// the device produce the same bits)
__host__ __device__ inline float serve_rhs_value(uint64_t seed, uint64_t j)
{
    uint64_t z = seed + (j + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)(int32_t)(z >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

// hash term of result element j with bit pattern `bits`; a request's hash is
// the wrapping sum of its terms
__host__ __device__ inline uint64_t serve_hash_term(uint64_t j, uint32_t bits)
{
    uint64_t z = (j * 0xD1B54A32D192ED03ull) ^ ((uint64_t)bits * 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// X[i * ldx + c] = serve_rhs_value(seed, i * k + c) for every request
hipError_t serve_rhs(const ServeReq* reqs, const int64_t* begin, int n, int64_t total, hipStream_t s);
// out[r] = hash of request r's Y slice (out is cleared first)
hipError_t serve_hash(const ServeReq* reqs, const int64_t* begin, int n, int64_t total, uint64_t* out,
                      hipStream_t s);

}  // namespace k
}  // namespace mr
