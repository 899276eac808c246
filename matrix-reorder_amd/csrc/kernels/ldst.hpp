// ldst.hpp -- wave64 helpers shared by the MERBIT kernels: wave reductions and
// streaming loads / stores (the non-temporal hint keeps the col / val streams
// and the y stores from displacing x in the caches).
#pragma once
#include "dtypes.hpp"

namespace mr {
namespace k {

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <bool NT, typename T> __device__ __forceinline__ T ld_stream(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT> __device__ __forceinline__ __half ld_stream(const __half* p)
{
    return __ushort_as_half(ld_stream<NT>(reinterpret_cast<const unsigned short*>(p)));
}
template <bool NT> __device__ __forceinline__ __hip_bfloat16 ld_stream(const __hip_bfloat16* p)
{
    return __ushort_as_bfloat16(ld_stream<NT>(reinterpret_cast<const unsigned short*>(p)));
}
template <bool NT, typename T> __device__ __forceinline__ void st_stream(T* p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

}  // namespace k
}  // namespace mr
