// launch_panel.hpp -- MERBIT over LDS column panels (merbit_panel.hip): the
// SpMV launcher and the device steps of its plan build.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <climits>
#include <cstdint>

namespace mr {
namespace k {

constexpr int kPanelBits = 15;
constexpr int kPanelCols = 1 << kPanelBits;   // columns per panel: 128 KB of F32 x in LDS

// row-sum stream: tiles of 64 lanes x 8 = kRowSumTile items holding at most
// kRowSumRows rows.  An item is
//   >= 0     a partial: its index in the low ib bits (ib = the stream's index
//            width), and above them c, the number of rows that end after it:
//            0 if the row goes on, else the row ends and c - 1 empty rows follow;
//   -c       (1 <= c <= kRowSumRows) a row without partials in the tile ends (a
//            row too long for a tile, or an empty row no earlier end absorbed),
//            and c - 1 empty rows follow;
//   kRowPad  padding.
// The empty rows get 0 from the tile's y stores.
constexpr int kRowSumTile = 512;
constexpr int kRowSumRows = 1024;
constexpr int32_t kRowPad = INT32_MIN;

// the device arrays of a built panel plan, as the launcher reads them
struct PanelView {
    int S = 8, W = 16, nt = 0;        // items per lane, waves per workgroup, non-temporal streams
    const uint16_t* pcol = nullptr;   // panel-major columns
    const void* pval = nullptr;       // panel-major values
    const int32_t* tseg = nullptr;    // first partial of every tile
    const int32_t* jobs = nullptr;    // (panel, first tile, end tile) per job
    const int32_t* wg = nullptr;      // first job of every workgroup, plus the end
    int n_wg = 0;
    int64_t n_cols = 0, n_rows = 0;
    const int32_t* items = nullptr;   // row-sum stream
    int rs_ib = 31;                   // its index width
    const int32_t* trow = nullptr;    // first row of every row-sum tile, plus n_rows
    int64_t n_rtiles = 0;
    const int32_t* lrow = nullptr;    // rows too long for a row-sum tile
    int n_long = 0;
    const int32_t* rs = nullptr;      // their partials: gidx[rs[r] .. rs[r + 1])
    const int32_t* gidx = nullptr;
};

// LDS bytes of a workgroup: one panel of x plus every wave's partial staging
size_t panel_lds_bytes(int items_per_lane, int waves);
// y = A x over the panel plan; part holds one float per partial
hipError_t panel_spmv(int dt, const PanelView& v, const void* x, void* part, void* y, hipStream_t s);
// the row-sum stream: items[n_items] all padding, then every row r with
// rcnt[r] > 0 at its packed position rpos[r]: its partial indices
// gidx[rs[r] .. rs[r + 1]), the last one carrying rcnt[r] << ib; a row without
// partials or longer than a tile gets the marker -rcnt[r] instead.  Rows with
// rcnt[r] = 0 are empty rows an earlier end closes.
hipError_t panel_rowsum_build(int64_t n_rows, int64_t n_items, int ib, const int32_t* rs, const int32_t* gidx,
                              const int32_t* rpos, const int32_t* rcnt, int32_t* items, hipStream_t s);

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
