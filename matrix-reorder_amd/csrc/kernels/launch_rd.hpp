// launch_rd.hpp -- host launcher of the sampled HyperLogLog reuse-distance kernel.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mr {
namespace k {

// per sample: serial float register sum, zero-register count, estimate (set
// for status 2 only), status and sampled position
hipError_t hll_sample(const int32_t* stream, int64_t nnz, int32_t samples, int32_t sector, uint32_t seed,
                      int32_t rd_cap, int32_t max_scan, int32_t p, float* out_sum, int32_t* out_zeros,
                      float* out_est, int32_t* out_status, int32_t* out_pos, hipStream_t s);

}  // namespace k
}  // namespace mr
