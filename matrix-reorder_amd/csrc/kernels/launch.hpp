// launch.hpp -- host launchers of the gfx950 kernels (implemented in *.hip).
// Each returns hipSuccess or the launch error; callers map it to a status.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mr {
namespace k {

// ---- utility (util.hip) ----
hipError_t fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
// convert n values from src dtype to dst dtype (MR_F64/F32/F16/BF16)
hipError_t convert_values(const void* src, int src_dt, void* dst, int dst_dt, int64_t n,
                          hipStream_t s);
// fill n values of dtype with 1
hipError_t fill_ones(void* dst, int dt, int64_t n, hipStream_t s);
// dir 0: out[perm[j]] = in[j]; dir 1: out[i] = in[perm[i]]; elem 4 or 8 bytes
hipError_t permute_vec(const int32_t* perm, int64_t n, int dir, const void* in, void* out,
                       int elem, hipStream_t s);
// 16-byte-word copy (bytes rounded down to a multiple of 16)
hipError_t copy_bytes(void* dst, const void* src, int64_t bytes, int cu_count, hipStream_t s);
// CSR validation: returns via flag (0 ok, 1 non-monotone, 2 column out of range)
hipError_t validate_csr(int64_t n_rows, int64_t n_cols, const int32_t* rp, const int32_t* ci,
                        int64_t nnz, int32_t* flag, hipStream_t s);

// ---- CSR SpMV / SpMM (spmv_csr.hip) ----
// row-length classes of the CSR kernels: <= 1, <= 2, 4, 8, 16, 32, 64,
// <= 4096, longer
constexpr int kCsrClasses = 9;
// per-class row counts (counts[kCsrClasses] zeroed by the caller)
hipError_t csr_bin_count(int64_t n_rows, const int32_t* rp, int32_t* counts, hipStream_t s);
// rows of class c to list[offs[c] ...] (cursor[kCsrClasses] zeroed by the caller)
hipError_t csr_bin_fill(int64_t n_rows, const int32_t* rp, const int32_t* offs, int32_t* cursor, int32_t* list,
                        hipStream_t s);
// y = A x over the binned rows; counts are the host copy of the class counts
hipError_t csr_spmv(int dt, int vector, const int32_t* list, const int32_t* counts, const int32_t* rp,
                    const int32_t* ci, const void* val, const void* x, void* y, hipStream_t s);
hipError_t csr_spmm(int dt, int64_t n_rows, int k, const int32_t* rp, const int32_t* ci,
                    const void* val, const void* X, int64_t ldx, void* Y, int64_t ldy,
                    hipStream_t s);

// ---- MERBIT (merbit.hip) ----
// one matrix of a fused batch (device-resident table; pointers typed by the
// batch's storage dtype): SpMV uses the first ten fields, SpMM all of them
constexpr int kMaxBatchJobs = 1024;
struct BatchJob {
    const void* ci;
    const void* val;
    const void* flags;
    const void* blk_x;
    const void* blk_y;
    const void* x;
    void* y;
    void* carry;
    void* carry_row;
    int64_t n_rows, nnz, ldx, ldy;
    int32_t kk, pad;
};
// persistent fused MERBIT SpMV over all jobs (begin[j] = first global block of
// job j, begin[n_jobs] = total) followed by a fused carry fix-up
hipError_t merbit_batch_spmv(int dt, int sigma, int nt, const void* jobs, const int64_t* begin, int n_jobs,
                             int64_t total, int grid, hipStream_t s);
// same for SpMM on the 4096-item grid, group width from the batch's k
hipError_t merbit_batch_spmm(int dt, int kk, const void* jobs, const int64_t* begin, int n_jobs, int64_t total,
                             int grid, hipStream_t s);
// merge coordinates at every multiple of `items_per_block` (+ sentinel)
hipError_t merbit_coords(int64_t n_rows, int64_t nnz, const int32_t* rp, int64_t items_per_block, int64_t n_blocks,
                         int32_t* blk_x, int32_t* blk_y, hipStream_t s);
// SpMM over 4096-item blocks, k <= 64 right-hand sides (merbit_mm.hip)
hipError_t merbit_spmm(int dt, int64_t n_rows, int64_t nnz, int64_t n_blocks, const int32_t* ci, const void* val,
                       const uint32_t* flags, const int32_t* blk_x, const int32_t* blk_y, const void* X,
                       int64_t ldx, void* Y, int64_t ldy, int kk, void* carry, int32_t* carry_row, hipStream_t s);
hipError_t merbit_build(int64_t n_rows, int64_t nnz, const int32_t* rp, int sigma,
                        int64_t n_blocks, uint32_t* flags, int64_t flag_words, int32_t* blk_x,
                        int32_t* blk_y, hipStream_t s);
// policy bit 0: non-temporal col/val loads and y stores; bit 1: XCD block remap
hipError_t merbit_spmv(int dt, int sigma, int policy, int64_t n_rows, int64_t n_blocks, const int32_t* ci,
                       const void* val, const uint32_t* flags, const int32_t* blk_x,
                       const int32_t* blk_y, const void* x, void* y, void* carry,
                       int32_t* carry_row, hipStream_t s);

// ---- MERBIT on lane-padded wave tiles with an LDS cache of the xc hottest
// columns (merbit_lp.hip); F32 / F16 / BF16 storage only.  The merge path is
// stored as pc / pv streams of `padded` items (column or -1 at a row end, -2
// after the last item; a cached column is slot | 1 << 31).
// (S, W) in {(8,16), (16,16), (8,8), (16,8), (4,16)} (S = 4 for F32 only)
size_t merbit_lp_lds_bytes(int xc);
// map == NULL: plain columns; pv == NULL: the value stream is left alone
// column map of the lane-padded stream: map[c] = c, except map[hot[i]] =
// i | 1 << 31 for the xc LDS-cached columns
hipError_t merbit_lp_map(int64_t nc, const int32_t* hot, int xc, int32_t* map, hipStream_t s);
hipError_t merbit_lp_build(int dt, int64_t n_rows, const int32_t* rp, const int32_t* ci, const void* val,
                           const int32_t* map, int64_t items, int64_t padded, int32_t* pc, void* pv, hipStream_t s);
// SpMM on the lane-padded tiles for kk <= 16 columns: pcm = plain-column
// stream (merbit_lp_build with map == NULL and pv == NULL), X rows of ldx
// floats (ldx % 4 == 0, ldx >= merbit_lp_mm_chunk(kk) = kk rounded up to 4, X
// 16-byte aligned), the first xr rows of X cached in LDS, carries of
// merbit_lp_mm_chunk(kk) floats per tile
int merbit_lp_mm_chunk(int kk);
bool merbit_lp_mm_fits(int S, int W, int kk);
hipError_t merbit_lp_spmm(int dt, int S, int W, int nt, int64_t n_rows, int64_t n_tiles, const int32_t* pcm,
                          const void* pv, const int32_t* ty, const void* X, int64_t ldx, int kk, int xr, void* Y,
                          int64_t ldy, void* carry, int32_t* carry_row, int grid, hipStream_t s);
hipError_t merbit_lp_spmv(int dt, int S, int W, int nt, int64_t n_rows, int64_t n_tiles, const int32_t* pc,
                          const void* pv, const int32_t* ty, const void* x, const int32_t* hot, int xc, void* y,
                          void* carry, int32_t* carry_row, int grid, hipStream_t s);

}  // namespace k
}  // namespace mr
