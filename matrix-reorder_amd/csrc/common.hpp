// common.hpp -- internal helpers of libmr: status/error plumbing, the context
// and matrix handle layouts, and device buffers.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mr.h"
#include "kernels/launch_panel.hpp"

namespace mr {

int set_error(int code, const char* fmt, ...);
void clear_error();

#define MR_HIP_TRY(expr)                                                              \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::mr::set_error(e_ == hipErrorOutOfMemory ? MR_ERR_OUT_OF_MEMORY   \
                                                             : MR_ERR_HIP,            \
                                   "%s failed: %s (%s:%d)", #expr,                    \
                                   hipGetErrorString(e_), __FILE__, __LINE__);        \
    } while (0)

#define MR_TRY(expr)                  \
    do {                              \
        int s_ = (expr);              \
        if (s_ != MR_OK) return s_;   \
    } while (0)

inline size_t dtype_size(int dt)
{
    switch (dt) {
    case MR_F64: return 8;
    case MR_F32: return 4;
    case MR_F16: return 2;
    case MR_BF16: return 2;
    default: return 0;
    }
}
// x / y element size for a matrix stored in dt
inline size_t vec_size(int dt) { return dt == MR_F64 ? 8 : 4; }

// RAII device buffer (hipMalloc / hipFree on the owning device)
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t n)
    {
        release();
        if (n == 0) n = 256;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(MR_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", n,
                             hipGetErrorString(e));
        }
        bytes = n;
        return MR_OK;
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace mr

struct mr_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string arch;
    std::mutex mu;
    // device address of the library's static scratch array on this device
    // (kernels/scratch.hip), looked up when the context is created
    void* scratch = nullptr;
};

// MERBIT plan: merge-path tiles of 256 lanes x sigma items
struct MerbitPlan {
    // defaults = the autotuner's pick on R-MAT s22 after pGOrder (MI355X)
    int sigma = 8;
    int policy = 1;          // bit 0 non-temporal streams, bit 1 XCD block remap
    int64_t items = 0;       // n_rows + nnz
    int64_t n_blocks = 0;
    mr::DevBuf flags;        // u32 words, row-end bitmap (n_blocks * 256 * sigma bits)
    mr::DevBuf blk_x;        // int32[n_blocks + 1]
    mr::DevBuf blk_y;        // int32[n_blocks + 1]
    mr::DevBuf carry;        // accumulator type [n_blocks]
    mr::DevBuf carry_row;    // int32[n_blocks]
    // SpMM grid: 4096 items per block, k-wide carries (k <= 64)
    int64_t mm_blocks = 0;
    mr::DevBuf mm_x, mm_y, mm_carry, mm_carry_row;
    bool mm_built = false;
    // merge coordinates at the other tile widths, for batched launches that
    // pick their own sigma: index 0 / 1 / 2 = sigma 8 / 16 / 32, built on use
    mr::DevBuf cx[3], cy[3];
    int64_t cnb[3] = {-1, -1, -1};
    // lane-padded wave-tile variant with an LDS cache of hot columns (policy
    // bit 2): tiles of 64 x xs merge items, xw waves per persistent workgroup,
    // xc cached columns (xc_req -1 = as many as the LDS holds); pc / pv = the
    // merge path as padded column / value streams, xty = rows ended before
    // every tile
    int xs = 8, xw = 16, xc_req = -1;
    int xc = 0;
    int64_t xtiles = 0;
    mr::DevBuf pc, pv, hot, xty, xcarry, xcarry_row;
    bool xc_built = false;
    // SpMM on the same tiles: plain-column stream and 16-wide carries
    mr::DevBuf pcm, xcarry_mm;
    bool pcm_built = false;
    // column-panel variant (policy bit 3): x in LDS one panel of 32768 columns
    // at a time; tiles of 64 x ps items, pw waves per workgroup; pcol / pval =
    // panel-major streams, ptseg = first partial of every tile, pjobs / pwg =
    // the workgroups' (panel, tile range) jobs, ppart = the partial sums
    // (pparts floats); prsi = the row-sum stream (prtiles tiles of 512 items:
    // each row's partial indices, prs_ib bits wide, the last one carrying the
    // number of rows that end there), ptrow = first row of every row-sum
    // tile; plrow = the pn_long rows too long for a tile, whose partials
    // prs / pgidx list (released when there are none)
    int ps = 8, pw = 16;
    int64_t ptiles = 0, pparts = 0, psegs = 0, prtiles = 0;
    int pn_wg = 0, pn_jobs = 0, pn_long = 0, prs_ib = 31;
    mr::DevBuf pcol, pval, ptseg, pjobs, pwg, prs, pgidx, ppart, prsi, ptrow, plrow;
    bool pn_built = false;
    double pn_build_ms = 0.0;
    int max_lds = 0;
    bool built = false;
    double build_ms = 0.0;
};

struct mr_mat_s {
    mr_ctx ctx = nullptr;
    int64_t n_rows = 0, n_cols = 0, nnz = 0;
    int dtype = MR_F32;
    mr::DevBuf rp;           // int32[n_rows + 1]
    mr::DevBuf ci;           // int32[nnz + pad], zero padded
    mr::DevBuf val;          // dtype[nnz + pad], zero padded
    std::vector<int32_t> perm;   // accumulated relabelling perm[old] = new
    MerbitPlan merbit;
    double avg_row = 0.0;
    mr::DevBuf csr_rows;     // rows grouped by length class (CSR kernels)
    int32_t csr_counts[9] = {0};
    bool csr_binned = false; // false: not built for the current structure
    // the columns with most nonzeros (ties: lower id), as many as an LDS x-cache
    // can hold, and their counts; hot_valid = computed for the current structure
    std::vector<int32_t> hot_order, hot_deg;
    bool hot_valid = false;
    // plan-cache entry of the last cached reorder, and the fingerprint of
    // perm at that point (a later relabel makes the entry stale)
    std::string plan_path;
    uint64_t plan_state = 0;
};

namespace mr {
constexpr int64_t kPad = 256;  // element padding of col/val arrays
int merbit_build(mr_mat m, hipStream_t s);
// merge coordinates of m at tile width 256 * sigma (the plan's own arrays when
// sigma matches, else a cached copy built on first use)
int merbit_coords_sigma(mr_mat m, int sigma, hipStream_t s, const int32_t** bx, const int32_t** by, int64_t* nb);
// the 4096-item SpMM grid of m, built on first use
int merbit_mm_grid(mr_mat m, hipStream_t s);
// metadata of the wave-tile / x-cache variant, built on first use
int merbit_xc_build(mr_mat m, hipStream_t s);
// plain-column stream for SpMM on the lane-padded tiles (builds the tiles too)
int merbit_lp_mm_build(mr_mat m, hipStream_t s);
// rows of X a LP SpMM launch of kk columns caches in LDS
int merbit_lp_mm_rows(mr_mat m, int kk);
// plan of the column-panel variant, built on first use (panel_plan.cpp)
int merbit_panel_build(mr_mat m, hipStream_t s);
// the device arrays of m's built column-panel plan
k::PanelView panel_view(mr_mat m);
}  // namespace mr
