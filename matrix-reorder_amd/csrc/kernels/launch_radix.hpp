// launch_radix.hpp -- stable LSD radix sort of (u32 key, u32 value) pairs
// (radix.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>

namespace mr {
namespace k {

// workspace bytes for radix_sort_pairs over n pairs
size_t radix_sort_ws_bytes(int64_t n);
// sort n pairs by the low key_bits bits of the key (1..32), stably: pairs with
// equal keys keep their input order.  Outputs must not alias the inputs.
hipError_t radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, int64_t n, int key_bits, void* ws, size_t ws_bytes, hipStream_t s);

}  // namespace k
}  // namespace mr
