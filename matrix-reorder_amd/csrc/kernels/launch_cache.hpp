// launch_cache.hpp -- the pattern fingerprint of the plan cache: formulas
// shared by the host and the device, and the device reduction.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace mr {
namespace k {

// SplitMix64 finaliser (Steele, Lea and Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014)
__host__ __device__ inline uint64_t fp_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// term of element i with value v; an array's fingerprint is the wrapping sum
// of its terms, so it does not depend on the order of the reduction but does
// depend on the position of every value
__host__ __device__ inline uint64_t fingerprint_term(uint64_t i, uint32_t v, uint64_t salt)
{
    return fp_mix64(salt + i * 0x9E3779B97F4A7C15ull + fp_mix64((uint64_t)v ^ salt));
}

constexpr uint64_t kSaltRowPtr = 0x726f775f70747273ull;
constexpr uint64_t kSaltColIdx = 0x636f6c5f69647873ull;
constexpr uint64_t kSaltPerm = 0x7065726d75746573ull;

// *out += sum over i < n of fingerprint_term(i, a[i], salt)
hipError_t fingerprint_sum(const int32_t* a, int64_t n, uint64_t salt, uint64_t* out, int cu_count,
                           hipStream_t s);

}  // namespace k
}  // namespace mr
