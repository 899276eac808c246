// xcd_order.hpp -- the XCD-aware tile order of the batched MERBIT launches.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace mr {
namespace k {

// Workgroups are dispatched round-robin over the 8 XCDs (block b to XCD b % 8),
// and every XCD has its own L2.  Dealing consecutive tiles to one XCD in chunks
// of `chunk` keeps the x vectors those tiles gather in that XCD's L2 instead of
// every XCD caching all of them.  The grid is padded to whole rounds of 8
// chunks; blocks mapped past the last tile have none.
// Tiles per chunk in the batched launches (0: block b runs tile b).  On the
// 64-matrix set of R-MAT and banded matrices, 16 or 64 was 0-2% faster than
// block-order tiles (k = 8 most) and 256 was 4-13% slower.
constexpr int kBatchChunk = 64;
inline int64_t xcd_chunk_grid(int64_t total, int chunk)
{
    const int64_t chunks = (total + chunk - 1) / chunk;
    return (chunks + 7) / 8 * 8 * chunk;
}
__device__ __forceinline__ int64_t xcd_chunk_tile(int64_t b, int chunk)
{
    const int64_t x = b % 8, i = b / 8;
    return ((i / chunk) * 8 + x) * chunk + i % chunk;
}

}  // namespace k
}  // namespace mr
