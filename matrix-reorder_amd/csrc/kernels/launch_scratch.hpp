// launch_scratch.hpp -- the static device scratch buffer (scratch.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>

namespace mr {
namespace k {

// bytes of the per-device static scratch array (64 MiB)
constexpr size_t kScratchBytes = (size_t)64 << 20;
// device address of that array on `device`, which must be current; looked up
// once per device and cached
hipError_t scratch_init(int device, void** ptr);

}  // namespace k
}  // namespace mr
