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

// ---- wave64 scans on DPP ----------------------------------------------------
// The lanes of a wave exchange values through DPP operand modifiers, which
// read another lane's register within the VALU instruction itself: no LDS
// round trip, unlike __shfl_up (ds_bpermute).  Controls (GFX9):
//   row_shr:n (0x110 + n)  lane i of a 16-lane row reads lane i - n of the row,
//   row_bcast:15 (0x142)   rows 1 and 3 (row mask 0xa) read lane 15 of the row
//                          before them,
//   row_bcast:31 (0x143)   rows 2 and 3 (row mask 0xc) read lane 31,
//   wave_shr:1 (0x138)     lane i reads lane i - 1 of the wave.
// A lane whose source is out of range, or whose row the mask leaves out, reads
// 0: the identity of every scan below.
template <int CTRL, int ROWS> __device__ __forceinline__ int dpp_read(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}

// inclusive prefix sum over the 64 lanes: four shift-and-add steps inside
// each row, then the rows' totals carried over by the two broadcasts
__device__ __forceinline__ int wave_incl_sum(int v)
{
    v += dpp_read<0x111, 0xf>(v);
    v += dpp_read<0x112, 0xf>(v);
    v += dpp_read<0x114, 0xf>(v);
    v += dpp_read<0x118, 0xf>(v);
    v += dpp_read<0x142, 0xa>(v);
    v += dpp_read<0x143, 0xc>(v);
    return v;
}

// one step of a segmented inclusive scan over (head flag, sum): a lane whose
// span holds a segment head keeps its sum, any other adds the sum before it
template <int CTRL, int ROWS> __device__ __forceinline__ void seg_step(float& v, int& f)
{
    const float vu = __int_as_float(dpp_read<CTRL, ROWS>(__float_as_int(v)));
    const int fu = dpp_read<CTRL, ROWS>(f);
    v = f ? v : v + vu;
    f |= fu;
}

// segmented inclusive scan: on return v = the sum of the lanes from the last
// lane at or before this one with f set (lane 0 if none) up to this one
__device__ __forceinline__ void wave_seg_scan(float& v, int& f)
{
    seg_step<0x111, 0xf>(v, f);
    seg_step<0x112, 0xf>(v, f);
    seg_step<0x114, 0xf>(v, f);
    seg_step<0x118, 0xf>(v, f);
    seg_step<0x142, 0xa>(v, f);
    seg_step<0x143, 0xc>(v, f);
}

// the value of lane - 1 (0 in lane 0)
__device__ __forceinline__ float wave_shr1(float v)
{
    return __int_as_float(dpp_read<0x138, 0xf>(__float_as_int(v)));
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
