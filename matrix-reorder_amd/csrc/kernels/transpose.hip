// transpose.hip -- CSR -> CSC with ascending row ids inside every column.
//
// The entries of a CSR are stored row by row, so a stable sort of the
// (column, row) pairs by column alone leaves each column's rows ascending:
// one stable LSD radix sort (radix.hip) over the column bits, then the column starts read
// off the sorted keys.  No atomics and no per-column sort, and the result does
// not depend on scheduling.  Used for the row lists rows(y) = { x : A[x][y] }
// that the pGOrder window score and the host orderers walk.
#include <hip/hip_runtime.h>

#include "launch_radix.hpp"
#include "launch_transpose.hpp"

namespace mr {
namespace k {

namespace {

// rows[q] = the row holding entry q; eight lanes per row
__global__ __launch_bounds__(256) void row_ids_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                                                      int32_t* __restrict__ rows)
{
    const int64_t groups = (int64_t)gridDim.x * (blockDim.x / 8);
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 8; r < n_rows; r += groups) {
        const int32_t b = rp[r], e = rp[r + 1];
        for (int32_t q = b + (int32_t)(threadIdx.x & 7); q < e; q += 8) rows[q] = (int32_t)r;
    }
}

// tp[c] = first position of a key >= c in the sorted keys; tp[n_cols] = nnz
__global__ void col_starts_kernel(int64_t nnz, int64_t n_cols, const uint32_t* __restrict__ keys,
                                  int32_t* __restrict__ tp)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q <= nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t prev = q > 0 ? (int64_t)keys[q - 1] : -1;
        const int64_t cur = q < nnz ? (int64_t)keys[q] : n_cols;
        for (int64_t c = prev + 1; c <= cur; ++c) tp[c] = (int32_t)q;   // columns (prev, cur] start at q
    }
}

unsigned grid_for(int64_t work, int64_t cap)
{
    int64_t g = (work + 255) / 256;
    if (g > cap) g = cap;
    return (unsigned)(g < 1 ? 1 : g);
}

unsigned key_bits(int64_t n_cols)
{
    unsigned b = 1;
    while (b < 32 && ((int64_t)1 << b) < n_cols) ++b;
    return b;
}

}  // namespace

size_t transpose_ws_bytes(int64_t nnz, int64_t n_cols)
{
    (void)n_cols;
    const size_t pairs = 2 * sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1);
    return pairs + radix_sort_ws_bytes(nnz) + 256;
}

hipError_t transpose_sorted(int64_t n_rows, int64_t n_cols, int64_t nnz, const int32_t* rp, const int32_t* ci,
                            int32_t* tp, int32_t* ti, void* ws, size_t ws_bytes, hipStream_t s)
{
    if (nnz <= 0) return hipMemsetAsync(tp, 0, sizeof(int32_t) * (size_t)(n_cols + 1), s);
    const unsigned bits = key_bits(n_cols);
    uint32_t* keys = static_cast<uint32_t*>(ws);
    int32_t* rows = reinterpret_cast<int32_t*>(keys + nnz);
    void* tmp = rows + nnz;
    const size_t used = 2 * sizeof(int32_t) * (size_t)nnz;
    const size_t tmp_bytes = radix_sort_ws_bytes(nnz);
    if (ws_bytes < used + tmp_bytes) return hipErrorInvalidValue;
    row_ids_kernel<<<grid_for(n_rows * 8, 16384), 256, 0, s>>>(n_rows, rp, rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(reinterpret_cast<const uint32_t*>(ci), reinterpret_cast<const uint32_t*>(rows), keys,
                         reinterpret_cast<uint32_t*>(ti), nnz, (int)bits, tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    col_starts_kernel<<<grid_for(nnz + 1, 65536), 256, 0, s>>>(nnz, n_cols, keys, tp);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
