// launch_perm.hpp -- host launchers of the relabel (P A P^T) kernels, the
// device scan, and the pGOrder GPU loop kernels.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mr {
namespace k {

// ---- scan.hip ----
size_t scan_workspace_bytes(int64_t n);
hipError_t exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, int64_t* total, void* ws,
                              hipStream_t s);

// ---- permute.hip ----
constexpr int kPermClasses = 9;   // 0: len<=1, 1..6: len<=2^c, 7: len<=4096, 8: longer
// new_len[p] = old length of row inv[p]
hipError_t perm_row_lengths(int64_t n, const int32_t* rp, const int32_t* inv, int32_t* new_len,
                            hipStream_t s);
// per-class row counts (counts[kPermClasses], zeroed by the caller)
hipError_t perm_class_count(int64_t n, const int32_t* new_len, int32_t* counts, hipStream_t s);
// scatter row ids into class lists starting at offs[c]; cursor[kPermClasses] zeroed by caller
hipError_t perm_class_fill(int64_t n, const int32_t* new_len, const int32_t* offs, int32_t* cursor,
                           int32_t* lists, hipStream_t s);
// relabel + sort the rows of one class below the last
hipError_t perm_sort_class(int cls, int32_t count, const int32_t* rows, const int32_t* rp,
                           const int32_t* ci, const void* val, int val_bytes, const int32_t* perm,
                           const int32_t* inv, const int32_t* new_rp, int32_t* new_ci, void* new_val,
                           hipStream_t s);
// the last class (rows longer than 4096): out[i] = new_len[rows[i]]; then,
// with boff the exclusive prefix of those lengths (total entries), one
// stable radix sort of all their entries
hipError_t perm_long_lengths(int32_t count, const int32_t* rows, const int32_t* new_len, int32_t* out, hipStream_t s);
size_t perm_long_ws_bytes(int64_t total, int32_t count, int64_t n);
hipError_t perm_sort_long(int32_t count, const int32_t* rows, const int32_t* rp, const int32_t* ci, const void* val,
                          int val_bytes, const int32_t* perm, const int32_t* inv, const int32_t* new_rp,
                          int32_t* new_ci, void* new_val, const int64_t* boff, int64_t total, int64_t n, void* ws,
                          size_t ws_bytes, hipStream_t s);

// inv[perm[i]] = i; *bad becomes 1 unless perm is a permutation of [0, n)
// (out-of-range entries are not written)
hipError_t invert_perm(int64_t n, const int32_t* perm, int32_t* inv, int32_t* bad, hipStream_t s);
// column counts of a CSR (cnt zeroed by the caller)
hipError_t transpose_count(int64_t nnz, const int32_t* ci, int32_t* cnt, hipStream_t s);

}  // namespace k
}  // namespace mr
