/*
 * merbit.c -- oracle restatement of the MERBIT merge-path tile metadata and a
 * sequential emulation of its SpMV.  TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_merge_search      merge_search (include/merbit.cuh:24-53)
 *   oracle_merbit_ref_tiles  Tile_Kernel + TILE ctor (merbit.cuh:55-155), 32-lane
 *                            tiles, u32 lane descriptor for SIGMA < 15:
 *                            flags bits 18..31, delta_y bits 9..17, delta_x 0..8;
 *                            tile_y bit 31 = no lane of the tile ends a row;
 *                            sentinel tile_x[tile_num] = nnz, tile_y = n.
 *   oracle_merbit_spmv       merbit_kernel (merbit.cuh:408-493) walked lane by
 *                            lane in merge-path order: a flag bit set = the item
 *                            ends the current row, clear = consume one nonzero.
 *
 * The concatenation of all lane flag bits is the row-end bitmap of the merge
 * path, bit (row_ptr[r+1] + r) for every row r, whatever the lane width; the
 * HIP library's own wave64 tiling (S items per lane, 256 lanes per block) must
 * reproduce exactly that bitmap and the merge coordinates at its block starts.
 *
 * Pinned by: a brute-force sequential merge of (row ends, nonzeros) compared
 * with merge_search at every diagonal and with the lane flags; the emulated y
 * compared with scipy.sparse (tests/test_oracle_basic.py).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

void oracle_merge_search(const int32_t* rp, int32_t n, int64_t nnz, int64_t diag,
                         int64_t* cx, int64_t* cy)
{
    int64_t y_min = diag - nnz > 0 ? diag - nnz : 0;
    int64_t y_max = diag < n ? diag : n;
    while (y_min < y_max) {
        const int64_t mid = (y_min + y_max) >> 1;
        if (rp[mid + 1] <= diag - mid - 1) y_min = mid + 1;
        else y_max = mid;
    }
    *cx = diag - y_min;
    *cy = y_min < n ? y_min : n;
}

int oracle_merbit_ref_tiles(int32_t n, int64_t nnz, const int32_t* rp, int32_t sigma,
                            uint32_t* tile_x, uint32_t* tile_y, uint32_t* lane_desc)
{
    if (sigma >= 15) return -1;
    const int64_t items = (int64_t)n + nnz;
    const int64_t lane_num = (items + sigma - 1) / sigma;
    const int64_t tile_num = (lane_num + 31) / 32;
    for (int64_t t = 0; t < tile_num; ++t) {
        const int64_t g0 = t * 32;
        const int64_t warp_len = (g0 + 32 < lane_num) ? 32 : lane_num - g0;
        int64_t wx = 0, wy = 0;
        int all_flat = 1;
        for (int64_t l = 0; l < warp_len; ++l) {
            const int64_t gid = g0 + l;
            const int64_t lane_items = ((gid + 1) * sigma < items) ? sigma : items - gid * sigma;
            int64_t x, y;
            oracle_merge_search(rp, n, nnz, gid * sigma, &x, &y);
            if (l == 0) { wx = x; wy = y; }
            const uint32_t dx = (uint32_t)(x - wx), dy = (uint32_t)(y - wy);
            uint32_t bits = 0;
            const int64_t y0 = y;
            for (int64_t j = 0; j < lane_items; ++j) {
                if (x < rp[y + 1]) ++x;
                else { bits |= 1u << j; ++y; }
            }
            lane_desc[gid] = (bits << 18) | (dy << 9) | dx;
            if (y0 != y) all_flat = 0;
        }
        tile_x[t] = (uint32_t)wx;
        tile_y[t] = (uint32_t)wy | (all_flat ? 0x80000000u : 0u);
    }
    tile_x[tile_num] = (uint32_t)nnz;
    tile_y[tile_num] = (uint32_t)n;
    return (int)tile_num;
}

void oracle_merbit_spmv(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                        const double* val, const double* x, int32_t sigma, double* y)
{
    const int64_t items = (int64_t)n + nnz;
    const int64_t lanes = (items + sigma - 1) / sigma;
    for (int32_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t l = 0; l < lanes; ++l) {
        int64_t cx, cy;
        oracle_merge_search(rp, n, nnz, l * sigma, &cx, &cy);
        const int64_t lane_items = ((l + 1) * sigma < items) ? sigma : items - l * sigma;
        double acc = 0.0;
        for (int64_t j = 0; j < lane_items; ++j) {
            if (cx < rp[cy + 1]) { acc += val[cx] * x[ci[cx]]; ++cx; }
            else { y[cy] += acc; acc = 0.0; ++cy; }
        }
        if (cy < n) y[cy] += acc;   /* carry into the row still open */
    }
}
