// scratch.hip -- the library's static device scratch buffer for the entry
// point that allocates nothing (mr_spmv_noalloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy of kScratchBytes for the life of the process.  In host code the
// array's name is only a host-side stand-in, so its device address is looked
// up once per device (hipGetSymbolAddress, with that device current) and only
// that address is handed to kernels, memsets and copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch_scratch.hpp"

namespace mr {
namespace k {

namespace {

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t scratch_init(int device, void** ptr)
{
    if (device < 0 || device >= kMaxDevices || !ptr) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // the caller makes the device current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

}  // namespace k
}  // namespace mr
