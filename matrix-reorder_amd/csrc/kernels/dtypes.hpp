// dtypes.hpp -- value-type traits shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>

namespace mr {
namespace k {

// Storage type -> accumulation / vector type.  F64 matrices use F64 x, y and
// accumulation; every other storage type accumulates in F32 with F32 x, y.
template <typename VT> struct Acc { using type = float; };
template <> struct Acc<double> { using type = double; };

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ double to_f(double v) { return v; }
__device__ __forceinline__ float to_f(__half v) { return __half2float(v); }
// BF16 -> F32 is an exact widening: the 16 stored bits become the high half
__device__ __forceinline__ float to_f(__hip_bfloat16 v)
{
    return __uint_as_float((uint32_t)__bfloat16_as_ushort(v) << 16);
}

// dtype codes match mr_dtype
enum { DT_F64 = 0, DT_F32 = 1, DT_F16 = 2, DT_BF16 = 3 };

// dispatch a functor templated on the storage type
template <typename F> hipError_t dispatch_dt(int dt, F&& f)
{
    switch (dt) {
    case DT_F64: return f(double{});
    case DT_F32: return f(float{});
    case DT_F16: return f(__half{});
    case DT_BF16: return f(__hip_bfloat16{});
    default: return hipErrorInvalidValue;
    }
}

}  // namespace k
}  // namespace mr
