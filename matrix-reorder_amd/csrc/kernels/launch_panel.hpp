// launch_panel.hpp -- MERBIT over LDS column panels (merbit_panel.hip): the
// SpMV launcher and the device steps of its plan build.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>

namespace mr {
namespace k {

constexpr int kPanelBits = 15;
constexpr int kPanelCols = 1 << kPanelBits;   // columns per panel: 128 KB of F32 x in LDS

// LDS bytes of a workgroup: one panel of x plus every wave's partial staging
size_t panel_lds_bytes(int items_per_lane, int waves);
// y = A x over the panel plan; part holds one float per partial.  wrow
// (n_chunks + 1 rows) cuts the rows for the partial sums; chunks of at most
// chunk_max partials are added through LDS (chunk_max * 4 bytes <=
// panel_reduce_lds_cap()), else straight from global memory
hipError_t panel_spmv(int dt, int S, int W, int nt, const uint16_t* pcol, const void* pval, const int32_t* tseg,
                      const int32_t* jobs, const int32_t* wg, int n_wg, const void* x, int64_t n_cols, int64_t n_rows,
                      const int32_t* rs, const int32_t* gidx, const int32_t* wrow, int n_chunks, int chunk_max,
                      void* part, void* y, hipStream_t s);
// LDS bytes the partial-sum kernel may use for one chunk
size_t panel_reduce_lds_cap();
// partial-balanced row chunks: wrow[0..n_chunks], *maxc = max chunk (atomicMax
// into a caller-zeroed int)
hipError_t panel_chunks(int64_t n_rows, const int32_t* rs, int n_chunks, int32_t* wrow, int32_t* maxc, hipStream_t s);

// plan build steps (see panel_plan.cpp for the order)
hipError_t panel_seg_flags(int64_t n_rows, const int32_t* rp, const int32_t* ci, int32_t* F, hipStream_t s);
hipError_t panel_seg_info(int64_t nnz, const int32_t* ci, const int32_t* F, const int32_t* sid, int32_t* k0,
                          uint32_t* key, uint32_t* ident, hipStream_t s);
// a[at] = *total (device int64) when total is given, else fixed
hipError_t panel_set_tail(int32_t* a, int64_t at, const int64_t* total, int64_t fixed, hipStream_t s);
hipError_t panel_sorted_len(int64_t n_seg, const uint32_t* order, const int32_t* k0, int32_t* len, hipStream_t s);
hipError_t panel_starts(int64_t n_seg, const uint32_t* skey, const int32_t* X, int32_t* pfirst, int32_t* pstart,
                        hipStream_t s);
hipError_t panel_seg_pos(int64_t n_seg, const uint32_t* order, const uint32_t* skey, const int32_t* X,
                         const int32_t* padd, int32_t* pos, hipStream_t s);
hipError_t panel_scatter(int dt, int64_t n_items, int64_t nnz, int tile, const int32_t* ci, const void* val,
                         const int32_t* F, const int32_t* sid, const int32_t* k0, const int32_t* pos, uint16_t* pcol,
                         void* pval, int32_t* flag, hipStream_t s);
hipError_t panel_tile_base(int64_t n_tiles, int tile, const int32_t* pidx, int32_t* tseg, hipStream_t s);
hipError_t panel_pieces(int64_t n_seg, int tile, const int32_t* pos, const int32_t* k0, int32_t* npieces,
                        hipStream_t s);
hipError_t panel_row_starts(int64_t n_rows, const int32_t* rp, const int32_t* sid, const int32_t* goff, int32_t* rs,
                            hipStream_t s);
hipError_t panel_gidx(int64_t n_seg, int tile, const int32_t* pos, const int32_t* k0, const int32_t* goff,
                      const int32_t* pidx, int32_t* gidx, hipStream_t s);

}  // namespace k
}  // namespace mr
