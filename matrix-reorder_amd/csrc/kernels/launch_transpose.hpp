// launch_transpose.hpp -- CSR -> CSC on the device (transpose.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>

namespace mr {
namespace k {

// workspace of transpose_sorted for nnz entries and n_cols columns
size_t transpose_ws_bytes(int64_t nnz, int64_t n_cols);
// tp[n_cols + 1], ti[nnz]: the rows of each column, ascending (column
// indices of ci must lie in [0, n_cols))
hipError_t transpose_sorted(int64_t n_rows, int64_t n_cols, int64_t nnz, const int32_t* rp, const int32_t* ci,
                            int32_t* tp, int32_t* ti, void* ws, size_t ws_bytes, hipStream_t s);

}  // namespace k
}  // namespace mr
