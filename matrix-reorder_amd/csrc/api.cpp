// api.cpp -- C-ABI entry points of libmr: errors, context, device plumbing,
// matrix handles, SpMV/SpMM dispatch and MERBIT metadata.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "common.hpp"
#include "kernels/launch.hpp"
#include "kernels/launch_panel.hpp"
#include "kernels/launch_perm.hpp"
#include "kernels/launch_scratch.hpp"

namespace mr {

static thread_local std::string g_error;

int set_error(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

static int use_device(mr_ctx c)
{
    MR_HIP_TRY(hipSetDevice(c->device));
    return MR_OK;
}

static hipStream_t stream_of(mr_ctx c, mr_stream s)
{
    return s ? static_cast<hipStream_t>(s) : c->stream;
}

int merbit_build(mr_mat m, hipStream_t s)
{
    MerbitPlan& p = m->merbit;
    p.built = false;
    const int64_t T = 256LL * p.sigma;
    p.items = m->n_rows + m->nnz;
    if (p.items > INT32_MAX)
        return set_error(MR_ERR_UNSUPPORTED, "merge path of %lld items exceeds int32",
                         (long long)p.items);
    p.n_blocks = (p.items + T - 1) / T;
    // bitmap padded to whole 8192-item blocks: every tile width (sigma 8..32,
    // the 4096-item SpMM grid) reads inside it
    const int64_t words = (p.items + 8191) / 8192 * (8192 / 32);
    p.mm_built = false;
    p.xc_built = false;
    for (int i = 0; i < 3; ++i) p.cnb[i] = -1;
    MR_TRY(p.flags.alloc(sizeof(uint32_t) * (size_t)(words > 0 ? words : 1)));
    MR_TRY(p.blk_x.alloc(sizeof(int32_t) * (size_t)(p.n_blocks + 1)));
    MR_TRY(p.blk_y.alloc(sizeof(int32_t) * (size_t)(p.n_blocks + 1)));
    MR_TRY(p.carry.alloc(vec_size(m->dtype) * (size_t)(p.n_blocks > 0 ? p.n_blocks : 1)));
    MR_TRY(p.carry_row.alloc(sizeof(int32_t) * (size_t)(p.n_blocks > 0 ? p.n_blocks : 1)));
    hipEvent_t e0, e1;
    MR_HIP_TRY(hipEventCreate(&e0));
    MR_HIP_TRY(hipEventCreate(&e1));
    MR_HIP_TRY(hipEventRecord(e0, s));
    MR_HIP_TRY(k::merbit_build(m->n_rows, m->nnz, m->rp.as<int32_t>(), p.sigma, p.n_blocks,
                               p.flags.as<uint32_t>(), words, p.blk_x.as<int32_t>(),
                               p.blk_y.as<int32_t>(), s));
    MR_HIP_TRY(hipEventRecord(e1, s));
    MR_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    p.build_ms = ms;
    p.built = true;
    return MR_OK;
}

int merbit_coords_sigma(mr_mat m, int sigma, hipStream_t s, const int32_t** bx, const int32_t** by, int64_t* nb)
{
    MerbitPlan& p = m->merbit;
    if (!p.built) MR_TRY(merbit_build(m, s));
    if (sigma == p.sigma) {
        *bx = p.blk_x.as<int32_t>();
        *by = p.blk_y.as<int32_t>();
        *nb = p.n_blocks;
        return MR_OK;
    }
    const int i = sigma == 8 ? 0 : sigma == 16 ? 1 : sigma == 32 ? 2 : -1;
    if (i < 0) return set_error(MR_ERR_INVALID_ARG, "sigma must be 8, 16 or 32 (got %d)", sigma);
    if (p.cnb[i] < 0) {
        const int64_t T = 256LL * sigma;
        const int64_t n = (p.items + T - 1) / T;
        MR_TRY(p.cx[i].alloc(sizeof(int32_t) * (size_t)(n + 1)));
        MR_TRY(p.cy[i].alloc(sizeof(int32_t) * (size_t)(n + 1)));
        MR_HIP_TRY(k::merbit_coords(m->n_rows, m->nnz, m->rp.as<int32_t>(), T, n, p.cx[i].as<int32_t>(),
                                    p.cy[i].as<int32_t>(), s));
        p.cnb[i] = n;
    }
    *bx = p.cx[i].as<int32_t>();
    *by = p.cy[i].as<int32_t>();
    *nb = p.cnb[i];
    return MR_OK;
}

int merbit_mm_grid(mr_mat m, hipStream_t s)
{
    MerbitPlan& p = m->merbit;
    if (!p.built) MR_TRY(merbit_build(m, s));
    if (p.mm_built) return MR_OK;
    p.mm_blocks = (p.items + 4095) / 4096;
    MR_TRY(p.mm_x.alloc(sizeof(int32_t) * (size_t)(p.mm_blocks + 1)));
    MR_TRY(p.mm_y.alloc(sizeof(int32_t) * (size_t)(p.mm_blocks + 1)));
    MR_TRY(p.mm_carry.alloc(vec_size(m->dtype) * 64 * (size_t)std::max<int64_t>(p.mm_blocks, 1)));
    MR_TRY(p.mm_carry_row.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(p.mm_blocks, 1)));
    MR_HIP_TRY(k::merbit_coords(m->n_rows, m->nnz, m->rp.as<int32_t>(), 4096, p.mm_blocks, p.mm_x.as<int32_t>(),
                                p.mm_y.as<int32_t>(), s));
    p.mm_built = true;
    return MR_OK;
}

int merbit_xc_build(mr_mat m, hipStream_t s)
{
    MerbitPlan& p = m->merbit;
    if (m->dtype == MR_F64) return set_error(MR_ERR_UNSUPPORTED, "the x-cache MERBIT variant needs F32/F16/BF16 values");
    if (!p.built) MR_TRY(merbit_build(m, s));
    if (p.xc_built) return MR_OK;
    int max_lds = 0;
    MR_HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->ctx->device));
    // as many hot columns as the workgroup's LDS holds
    const int64_t cap_lds = max_lds > 512 ? ((int64_t)max_lds / 4 - 64) / 64 * 64 : 0;
    int64_t cap = cap_lds;
    if (p.xc_req >= 0) cap = std::min<int64_t>(cap, p.xc_req);
    cap = std::max<int64_t>(0, std::min<int64_t>(cap, m->n_cols));
    const int64_t nc = std::max<int64_t>(m->n_cols, 1);
    // column degrees -> the columns with most nonzeros (ties: lower id), once
    // per structure for the largest cache; smaller caches take a prefix
    if (!m->hot_valid) {
        m->hot_order.clear();
        m->hot_deg.clear();
        const int64_t keep = std::min<int64_t>(cap_lds, m->n_cols);
        if (m->nnz > 0 && keep > 0) {
            std::vector<int32_t> deg((size_t)nc, 0);
            DevBuf cnt;
            MR_TRY(cnt.alloc(sizeof(int32_t) * (size_t)nc));
            MR_HIP_TRY(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));
            MR_HIP_TRY(k::transpose_count(m->nnz, m->ci.as<int32_t>(), cnt.as<int32_t>(), s));
            MR_HIP_TRY(hipMemcpyAsync(deg.data(), cnt.p, sizeof(int32_t) * (size_t)nc, hipMemcpyDeviceToHost, s));
            MR_HIP_TRY(hipStreamSynchronize(s));
            std::vector<int32_t> idx((size_t)nc);
            for (int64_t i = 0; i < nc; ++i) idx[(size_t)i] = (int32_t)i;
            auto hotter = [&](int32_t a, int32_t b) {
                return deg[(size_t)a] != deg[(size_t)b] ? deg[(size_t)a] > deg[(size_t)b] : a < b;
            };
            std::partial_sort(idx.begin(), idx.begin() + keep, idx.end(), hotter);
            m->hot_order.assign(idx.begin(), idx.begin() + keep);
            m->hot_deg.resize((size_t)keep);
            for (int64_t i = 0; i < keep; ++i) m->hot_deg[(size_t)i] = deg[(size_t)idx[(size_t)i]];
        }
        m->hot_valid = true;
    }
    int64_t xc = std::min<int64_t>(cap, (int64_t)m->hot_order.size());
    while (xc > 0 && m->hot_deg[(size_t)xc - 1] == 0) --xc;   // never cache an unused column
    const int64_t T = 64LL * p.xs;
    p.xtiles = (p.items + T - 1) / T;
    const int64_t padded = std::max<int64_t>(p.xtiles * T, T);
    DevBuf dmap, tx;
    MR_TRY(dmap.alloc(sizeof(int32_t) * (size_t)nc));
    MR_TRY(tx.alloc(sizeof(int32_t) * (size_t)(p.xtiles + 1)));
    MR_TRY(p.hot.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(xc, 1)));
    MR_TRY(p.pc.alloc(sizeof(int32_t) * (size_t)padded));
    MR_TRY(p.pv.alloc(dtype_size(m->dtype) * (size_t)padded));
    MR_TRY(p.xty.alloc(sizeof(int32_t) * (size_t)(p.xtiles + 1)));
    MR_TRY(p.xcarry.alloc(sizeof(float) * (size_t)std::max<int64_t>(p.xtiles, 1)));
    MR_TRY(p.xcarry_row.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(p.xtiles, 1)));
    if (xc > 0)
        MR_HIP_TRY(hipMemcpyAsync(p.hot.p, m->hot_order.data(), sizeof(int32_t) * (size_t)xc, hipMemcpyHostToDevice,
                                  s));
    MR_HIP_TRY(k::merbit_lp_map(nc, p.hot.as<int32_t>(), (int)xc, dmap.as<int32_t>(), s));
    MR_HIP_TRY(k::merbit_lp_build(m->dtype, m->n_rows, m->rp.as<int32_t>(), m->ci.as<int32_t>(), m->val.p,
                                  dmap.as<int32_t>(), p.items, padded, p.pc.as<int32_t>(), p.pv.p, s));
    MR_HIP_TRY(k::merbit_coords(m->n_rows, m->nnz, m->rp.as<int32_t>(), T, p.xtiles, tx.as<int32_t>(),
                                p.xty.as<int32_t>(), s));
    MR_HIP_TRY(hipStreamSynchronize(s));   // dmap and tx are freed on return
    p.xc = (int)xc;
    p.max_lds = max_lds;
    p.xc_built = true;
    p.pcm_built = false;
    return MR_OK;
}

int merbit_lp_mm_build(mr_mat m, hipStream_t s)
{
    MerbitPlan& p = m->merbit;
    MR_TRY(merbit_xc_build(m, s));
    if (p.pcm_built) return MR_OK;
    const int64_t T = 64LL * p.xs;
    const int64_t padded = std::max<int64_t>(p.xtiles * T, T);
    MR_TRY(p.pcm.alloc(sizeof(int32_t) * (size_t)padded));
    MR_TRY(p.xcarry_mm.alloc(sizeof(float) * 16 * (size_t)std::max<int64_t>(p.xtiles, 1)));
    MR_HIP_TRY(k::merbit_lp_build(m->dtype, m->n_rows, m->rp.as<int32_t>(), m->ci.as<int32_t>(), m->val.p, nullptr,
                                  p.items, padded, p.pcm.as<int32_t>(), nullptr, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    p.pcm_built = true;
    return MR_OK;
}

int merbit_lp_mm_rows(mr_mat m, int kk)
{
    const int K = k::merbit_lp_mm_chunk(kk);
    const int64_t rows = (int64_t)m->merbit.max_lds / 4 / K;
    return (int)std::max<int64_t>(0, std::min<int64_t>(rows, m->n_cols));
}

static int check_mat(mr_mat m)
{
    if (!m || !m->ctx) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    return use_device(m->ctx);
}

// the CSR kernels' row list, grouped by length class (once per structure)
static int csr_bin_build(mr_mat m, hipStream_t st)
{
    DevBuf counts, offs, cursor;
    MR_TRY(counts.alloc(sizeof(int32_t) * k::kCsrClasses));
    MR_TRY(offs.alloc(sizeof(int32_t) * k::kCsrClasses));
    MR_TRY(cursor.alloc(sizeof(int32_t) * k::kCsrClasses));
    MR_TRY(m->csr_rows.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m->n_rows, 1)));
    MR_HIP_TRY(hipMemsetAsync(counts.p, 0, counts.bytes, st));
    MR_HIP_TRY(hipMemsetAsync(cursor.p, 0, cursor.bytes, st));
    MR_HIP_TRY(k::csr_bin_count(m->n_rows, m->rp.as<int32_t>(), counts.as<int32_t>(), st));
    int32_t h[k::kCsrClasses], o[k::kCsrClasses];
    MR_HIP_TRY(hipMemcpyAsync(h, counts.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MR_HIP_TRY(hipStreamSynchronize(st));
    int32_t run = 0;
    for (int c = 0; c < k::kCsrClasses; ++c) {
        o[c] = run;
        run += h[c];
    }
    MR_HIP_TRY(hipMemcpyAsync(offs.p, o, sizeof(o), hipMemcpyHostToDevice, st));
    MR_HIP_TRY(k::csr_bin_fill(m->n_rows, m->rp.as<int32_t>(), offs.as<int32_t>(), cursor.as<int32_t>(),
                               m->csr_rows.as<int32_t>(), st));
    MR_HIP_TRY(hipStreamSynchronize(st));   // o[] and the counters are released below
    std::memcpy(m->csr_counts, h, sizeof(h));
    m->csr_binned = true;
    return MR_OK;
}

}  // namespace mr

using namespace mr;

extern "C" {

const char* mr_last_error(void) { return g_error.c_str(); }
const char* mr_version(void) { return "matrix-reorder-mi355x 0.1 (gfx950)"; }

int mr_device_count(int* count)
{
    if (!count) return set_error(MR_ERR_INVALID_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return MR_OK;
}

int mr_ctx_create(int device, mr_ctx* out)
{
    if (!out) return set_error(MR_ERR_INVALID_ARG, "null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(MR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n)
        return set_error(MR_ERR_INVALID_ARG, "device %d out of range [0,%d)", device, n);
    MR_HIP_TRY(hipSetDevice(device));
    auto c = std::make_unique<mr_ctx_s>();
    c->device = device;
    MR_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    MR_HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    MR_HIP_TRY(k::scratch_init(device, &c->scratch));
    *out = c.release();
    return MR_OK;
}

int mr_ctx_destroy(mr_ctx ctx)
{
    if (!ctx) return MR_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) {
        hipStreamSynchronize(ctx->stream);
        hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return MR_OK;
}

int mr_ctx_stream(mr_ctx ctx, mr_stream* out)
{
    if (!ctx || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    *out = ctx->stream;
    return MR_OK;
}

int mr_ctx_mem_info(mr_ctx ctx, size_t* free_bytes, size_t* total_bytes)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    MR_TRY(use_device(ctx));
    size_t f = 0, t = 0;
    MR_HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return MR_OK;
}

int mr_ctx_arch(mr_ctx ctx, char* name, int cap, int* cu_count)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    if (name && cap > 0) {
        std::strncpy(name, ctx->arch.c_str(), (size_t)cap - 1);
        name[cap - 1] = 0;
    }
    if (cu_count) *cu_count = ctx->cu_count;
    return MR_OK;
}

int mr_malloc(mr_ctx ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipMalloc(dptr, bytes ? bytes : 256));
    return MR_OK;
}

int mr_free(mr_ctx ctx, void* dptr)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    if (!dptr) return MR_OK;
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipFree(dptr));
    return MR_OK;
}

int mr_memcpy(mr_ctx ctx, void* dst, const void* src, size_t bytes, int kind)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    if (bytes == 0) return MR_OK;
    if (!dst || !src) return set_error(MR_ERR_INVALID_ARG, "null pointer");
    MR_TRY(use_device(ctx));
    hipMemcpyKind k = kind == MR_H2D ? hipMemcpyHostToDevice
                    : kind == MR_D2H ? hipMemcpyDeviceToHost
                                     : hipMemcpyDeviceToDevice;
    MR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    MR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MR_OK;
}

int mr_memset(mr_ctx ctx, void* dptr, int value, size_t bytes)
{
    if (!ctx || !dptr) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipMemsetAsync(dptr, value, bytes, ctx->stream));
    MR_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MR_OK;
}

int mr_bench_copy(mr_ctx ctx, void* dst, const void* src, size_t bytes, mr_stream s)
{
    if (!ctx || !dst || !src) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (bytes % 16) return set_error(MR_ERR_INVALID_ARG, "bytes must be a multiple of 16");
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(k::copy_bytes(dst, src, (int64_t)bytes, ctx->cu_count, stream_of(ctx, s)));
    return MR_OK;
}

int mr_stream_create(mr_ctx ctx, mr_stream* out)
{
    if (!ctx || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(use_device(ctx));
    hipStream_t s;
    MR_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return MR_OK;
}

int mr_stream_destroy(mr_ctx ctx, mr_stream s)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    if (!s) return MR_OK;
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(s)));
    return MR_OK;
}

int mr_stream_sync(mr_ctx ctx, mr_stream s)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipStreamSynchronize(stream_of(ctx, s)));
    return MR_OK;
}

int mr_event_create(mr_ctx ctx, mr_event* out)
{
    if (!ctx || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(use_device(ctx));
    hipEvent_t e;
    MR_HIP_TRY(hipEventCreate(&e));
    *out = e;
    return MR_OK;
}

int mr_event_destroy(mr_ctx ctx, mr_event e)
{
    if (!ctx) return set_error(MR_ERR_INVALID_ARG, "null context");
    if (!e) return MR_OK;
    MR_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(e)));
    return MR_OK;
}

int mr_event_record(mr_ctx ctx, mr_event e, mr_stream s)
{
    if (!ctx || !e) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(use_device(ctx));
    MR_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(e), stream_of(ctx, s)));
    return MR_OK;
}

int mr_event_elapsed_ms(mr_ctx ctx, mr_event start, mr_event stop, float* ms)
{
    if (!ctx || !start || !stop || !ms) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    MR_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return MR_OK;
}

// ---------------------------------------------------------------------------
int mr_csr_create(mr_ctx ctx, int64_t n_rows, int64_t n_cols, int64_t nnz, const int32_t* row_ptr,
                  const int32_t* col_idx, const void* val, int val_dtype, int store_dtype,
                  mr_mat* out)
{
    if (!ctx || !out || !row_ptr) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (n_rows < 0 || n_cols < 0 || nnz < 0 || n_rows > INT32_MAX - 1 || nnz > INT32_MAX - kPad)
        return set_error(MR_ERR_INVALID_ARG, "sizes out of range (n_rows=%lld nnz=%lld)",
                         (long long)n_rows, (long long)nnz);
    if (nnz > 0 && !col_idx) return set_error(MR_ERR_INVALID_ARG, "null col_idx");
    if (dtype_size(store_dtype) == 0 || (val && dtype_size(val_dtype) == 0))
        return set_error(MR_ERR_INVALID_ARG, "unknown dtype");
    // host-side validation: monotone row pointer and column range
    if (row_ptr[0] != 0 || row_ptr[n_rows] != nnz)
        return set_error(MR_ERR_MALFORMED_CSR, "row_ptr[0]=%d row_ptr[n]=%d nnz=%lld", row_ptr[0],
                         row_ptr[n_rows], (long long)nnz);
    for (int64_t i = 0; i < n_rows; ++i)
        if (row_ptr[i + 1] < row_ptr[i])
            return set_error(MR_ERR_MALFORMED_CSR, "row_ptr decreases at row %lld", (long long)i);
    for (int64_t q = 0; q < nnz; ++q)
        if (col_idx[q] < 0 || col_idx[q] >= n_cols)
            return set_error(MR_ERR_MALFORMED_CSR, "col_idx[%lld]=%d outside [0,%lld)", (long long)q,
                             col_idx[q], (long long)n_cols);
    MR_TRY(use_device(ctx));
    auto m = std::make_unique<mr_mat_s>();
    m->ctx = ctx;
    m->n_rows = n_rows;
    m->n_cols = n_cols;
    m->nnz = nnz;
    m->dtype = store_dtype;
    m->avg_row = n_rows > 0 ? (double)nnz / (double)n_rows : 0.0;
    hipStream_t s = ctx->stream;
    MR_TRY(m->rp.alloc(sizeof(int32_t) * (size_t)(n_rows + 1)));
    MR_TRY(m->ci.alloc(sizeof(int32_t) * (size_t)(nnz + kPad)));
    MR_TRY(m->val.alloc(dtype_size(store_dtype) * (size_t)(nnz + kPad)));
    MR_HIP_TRY(hipMemsetAsync(m->ci.p, 0, m->ci.bytes, s));
    MR_HIP_TRY(hipMemsetAsync(m->val.p, 0, m->val.bytes, s));
    MR_HIP_TRY(hipMemcpyAsync(m->rp.p, row_ptr, sizeof(int32_t) * (size_t)(n_rows + 1),
                              hipMemcpyHostToDevice, s));
    if (nnz > 0)
        MR_HIP_TRY(hipMemcpyAsync(m->ci.p, col_idx, sizeof(int32_t) * (size_t)nnz,
                                  hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        if (!val) {
            MR_HIP_TRY(k::fill_ones(m->val.p, store_dtype, nnz, s));
        } else if (val_dtype == store_dtype) {
            MR_HIP_TRY(hipMemcpyAsync(m->val.p, val, dtype_size(store_dtype) * (size_t)nnz,
                                      hipMemcpyHostToDevice, s));
        } else {
            DevBuf tmp;
            MR_TRY(tmp.alloc(dtype_size(val_dtype) * (size_t)nnz));
            MR_HIP_TRY(hipMemcpyAsync(tmp.p, val, dtype_size(val_dtype) * (size_t)nnz,
                                      hipMemcpyHostToDevice, s));
            MR_HIP_TRY(k::convert_values(tmp.p, val_dtype, m->val.p, store_dtype, nnz, s));
            MR_HIP_TRY(hipStreamSynchronize(s));
        }
    }
    m->perm.resize((size_t)n_rows);
    for (int64_t i = 0; i < n_rows; ++i) m->perm[(size_t)i] = (int32_t)i;
    MR_HIP_TRY(hipStreamSynchronize(s));
    MR_TRY(merbit_build(m.get(), s));
    // large FP32-accumulate matrices default to the column-panel variant with
    // non-temporal streams (the autotuner's pick at R-MAT s22; its plan builds
    // in about the time of the lane-padded one's); small ones (panel loads
    // would dominate) to workgroup tiles
    m->merbit.policy = (store_dtype != MR_F64 && nnz >= (int64_t(1) << 22)) ? 9 : 1;
    *out = m.release();
    return MR_OK;
}

int mr_mat_destroy(mr_mat m)
{
    if (!m) return MR_OK;
    if (m->ctx) {
        hipSetDevice(m->ctx->device);
        hipStreamSynchronize(m->ctx->stream);
    }
    delete m;
    return MR_OK;
}

int mr_mat_info(mr_mat m, int64_t* n_rows, int64_t* n_cols, int64_t* nnz, int* dtype)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (n_rows) *n_rows = m->n_rows;
    if (n_cols) *n_cols = m->n_cols;
    if (nnz) *nnz = m->nnz;
    if (dtype) *dtype = m->dtype;
    return MR_OK;
}

int mr_mat_export(mr_mat m, int32_t* row_ptr, int32_t* col_idx, void* val)
{
    MR_TRY(check_mat(m));
    hipStream_t s = m->ctx->stream;
    if (row_ptr)
        MR_HIP_TRY(hipMemcpyAsync(row_ptr, m->rp.p, sizeof(int32_t) * (size_t)(m->n_rows + 1),
                                  hipMemcpyDeviceToHost, s));
    if (col_idx && m->nnz > 0)
        MR_HIP_TRY(hipMemcpyAsync(col_idx, m->ci.p, sizeof(int32_t) * (size_t)m->nnz,
                                  hipMemcpyDeviceToHost, s));
    if (val && m->nnz > 0)
        MR_HIP_TRY(hipMemcpyAsync(val, m->val.p, dtype_size(m->dtype) * (size_t)m->nnz,
                                  hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    return MR_OK;
}

int mr_mat_convert(mr_mat m, int dtype, mr_mat* out)
{
    MR_TRY(check_mat(m));
    if (!out || dtype_size(dtype) == 0) return set_error(MR_ERR_INVALID_ARG, "bad argument");
    hipStream_t s = m->ctx->stream;
    auto c = std::make_unique<mr_mat_s>();
    c->ctx = m->ctx;
    c->n_rows = m->n_rows;
    c->n_cols = m->n_cols;
    c->nnz = m->nnz;
    c->dtype = dtype;
    c->avg_row = m->avg_row;
    c->perm = m->perm;
    c->merbit.sigma = dtype == MR_F64 ? std::min(m->merbit.sigma, 16) : m->merbit.sigma;
    c->merbit.policy = dtype == MR_F64 ? (m->merbit.policy & 3) : m->merbit.policy;
    c->merbit.xs = (dtype != MR_F32 && m->merbit.xs == 4) ? 8 : m->merbit.xs;
    c->merbit.xw = m->merbit.xw;
    c->merbit.xc_req = m->merbit.xc_req;
    c->merbit.ps = m->merbit.ps;
    c->merbit.pw = m->merbit.pw;
    MR_TRY(c->rp.alloc(m->rp.bytes));
    MR_TRY(c->ci.alloc(m->ci.bytes));
    MR_TRY(c->val.alloc(dtype_size(dtype) * (size_t)(m->nnz + kPad)));
    MR_HIP_TRY(hipMemcpyAsync(c->rp.p, m->rp.p, m->rp.bytes, hipMemcpyDeviceToDevice, s));
    MR_HIP_TRY(hipMemcpyAsync(c->ci.p, m->ci.p, m->ci.bytes, hipMemcpyDeviceToDevice, s));
    MR_HIP_TRY(hipMemsetAsync(c->val.p, 0, c->val.bytes, s));
    MR_HIP_TRY(k::convert_values(m->val.p, m->dtype, c->val.p, dtype, m->nnz, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    MR_TRY(merbit_build(c.get(), s));
    *out = c.release();
    return MR_OK;
}

int mr_mat_perm(mr_mat m, int32_t* perm)
{
    if (!m || !perm) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (m->n_rows > 0) std::memcpy(perm, m->perm.data(), sizeof(int32_t) * (size_t)m->n_rows);
    return MR_OK;
}

int mr_check_perm(const int32_t* perm, int64_t n)
{
    if (!perm || n <= 0) return set_error(MR_ERR_INVALID_ARG, "invalid input: perm=%p n=%lld",
                                          (const void*)perm, (long long)n);
    std::vector<uint8_t> seen((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t v = perm[i];
        if (v < 0 || v >= n)
            return set_error(MR_ERR_NOT_PERMUTATION, "perm[%lld]=%d outside [0,%lld)", (long long)i, v,
                             (long long)n);
        if (seen[(size_t)v])
            return set_error(MR_ERR_NOT_PERMUTATION, "value %d repeated (at %lld)", v, (long long)i);
        seen[(size_t)v] = 1;
    }
    return MR_OK;
}

int mr_permute_vec(mr_mat m, int dir, const void* in, void* out, mr_stream s)
{
    MR_TRY(check_mat(m));
    if (!in || !out || (dir != 0 && dir != 1) || m->n_rows != m->n_cols)
        return set_error(MR_ERR_INVALID_ARG, "bad argument");
    if (m->n_rows == 0) return MR_OK;
    hipStream_t st = stream_of(m->ctx, s);
    DevBuf dperm;
    MR_TRY(dperm.alloc(sizeof(int32_t) * (size_t)m->n_rows));
    MR_HIP_TRY(hipMemcpyAsync(dperm.p, m->perm.data(), sizeof(int32_t) * (size_t)m->n_rows,
                              hipMemcpyHostToDevice, st));
    MR_HIP_TRY(k::permute_vec(dperm.as<int32_t>(), m->n_rows, dir, in, out, (int)vec_size(m->dtype), st));
    MR_HIP_TRY(hipStreamSynchronize(st));
    return MR_OK;
}

// ---------------------------------------------------------------------------
int mr_spmv(mr_mat m, int kernel, const void* x, void* y, mr_stream s)
{
    MR_TRY(check_mat(m));
    if ((!x || !y) && m->n_rows > 0) return set_error(MR_ERR_INVALID_ARG, "null x or y");
    hipStream_t st = stream_of(m->ctx, s);
    switch (kernel) {
    case MR_KERNEL_MERBIT: {
        MerbitPlan& p = m->merbit;
        if (!p.built) MR_TRY(merbit_build(m, st));
        if ((p.policy & 8) && m->dtype != MR_F64) {
            MR_TRY(merbit_panel_build(m, st));
            MR_HIP_TRY(k::panel_spmv(m->dtype, panel_view(m), x, p.ppart.p, y, st));
            return MR_OK;
        }
        if ((p.policy & 4) && m->dtype != MR_F64) {
            MR_TRY(merbit_xc_build(m, st));
            MR_HIP_TRY(k::merbit_lp_spmv(m->dtype, p.xs, p.xw, p.policy & 1, m->n_rows, p.xtiles,
                                         p.pc.as<int32_t>(), p.pv.p, p.xty.as<int32_t>(), x, p.hot.as<int32_t>(),
                                         p.xc, y, p.xcarry.p, p.xcarry_row.as<int32_t>(),
                                         std::max(m->ctx->cu_count, 1), st));
            return MR_OK;
        }
        if (m->dtype == MR_F64 && p.sigma > 16)
            return set_error(MR_ERR_UNSUPPORTED, "F64 MERBIT supports sigma <= 16");
        MR_HIP_TRY(k::merbit_spmv(m->dtype, p.sigma, p.policy, m->n_rows, p.n_blocks, m->ci.as<int32_t>(), m->val.p,
                                  p.flags.as<uint32_t>(), p.blk_x.as<int32_t>(), p.blk_y.as<int32_t>(), x,
                                  y, p.carry.p, p.carry_row.as<int32_t>(), st));
        return MR_OK;
    }
    case MR_KERNEL_CSR_SCALAR:
    case MR_KERNEL_CSR_VECTOR:
        if (!m->csr_binned) MR_TRY(csr_bin_build(m, st));
        MR_HIP_TRY(k::csr_spmv(m->dtype, kernel == MR_KERNEL_CSR_VECTOR, m->csr_rows.as<int32_t>(), m->csr_counts,
                               m->rp.as<int32_t>(), m->ci.as<int32_t>(), m->val.p, x, y, st));
        return MR_OK;
    default:
        return set_error(MR_ERR_INVALID_ARG, "unknown kernel %d", kernel);
    }
}

int mr_merbit_ws_bytes(mr_mat m, size_t* bytes)
{
    if (!m || !bytes) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MerbitPlan& p = m->merbit;
    if ((p.policy & 8) && m->dtype != MR_F64) {
        // one float per partial sum of the column-panel plan
        if (!p.pn_built) {
            MR_TRY(check_mat(m));
            MR_TRY(merbit_panel_build(m, m->ctx->stream));
        }
        *bytes = sizeof(float) * (size_t)std::max<int64_t>(p.pparts, 1);
        return MR_OK;
    }
    size_t nb = (size_t)(p.n_blocks > 0 ? p.n_blocks : 1);
    // the x-cache variant needs one carry per tile of 64 x xs items
    const int64_t xt = (p.items + 64LL * p.xs - 1) / (64LL * p.xs);
    if ((p.policy & 4) && m->dtype != MR_F64) nb = std::max<size_t>(nb, (size_t)std::max<int64_t>(xt, 1));
    *bytes = ((nb * vec_size(m->dtype) + 255) / 256) * 256 + nb * sizeof(int32_t);
    return MR_OK;
}

// MERBIT SpMV with caller-provided carries; build = false refuses to build a
// missing plan (the no-allocation entry point)
static int spmv_ws_impl(mr_mat m, const void* x, void* y, void* ws, hipStream_t st, bool build)
{
    MerbitPlan& p = m->merbit;
    const bool pn = (p.policy & 8) && m->dtype != MR_F64;
    const bool lp = !pn && (p.policy & 4) && m->dtype != MR_F64;
    if (!build && (!p.built || (lp && !p.xc_built) || (pn && !p.pn_built)))
        return set_error(MR_ERR_INVALID_ARG,
                         "MERBIT plan not built: run mr_spmv, mr_spmv_ws or mr_merbit_autotune on the matrix first");
    if (!p.built) MR_TRY(merbit_build(m, st));
    if (pn) {
        MR_TRY(merbit_panel_build(m, st));
        MR_HIP_TRY(k::panel_spmv(m->dtype, panel_view(m), x, ws, y, st));
        return MR_OK;
    }
    if (lp) {
        MR_TRY(merbit_xc_build(m, st));
        const size_t nt = (size_t)std::max<int64_t>(p.xtiles, 1);
        char* base = static_cast<char*>(ws);
        int32_t* crow = reinterpret_cast<int32_t*>(base + ((nt * sizeof(float) + 255) / 256) * 256);
        MR_HIP_TRY(k::merbit_lp_spmv(m->dtype, p.xs, p.xw, p.policy & 1, m->n_rows, p.xtiles, p.pc.as<int32_t>(),
                                     p.pv.p, p.xty.as<int32_t>(), x, p.hot.as<int32_t>(), p.xc, y, base, crow,
                                     std::max(m->ctx->cu_count, 1), st));
        return MR_OK;
    }
    if (m->dtype == MR_F64 && p.sigma > 16)
        return set_error(MR_ERR_UNSUPPORTED, "F64 MERBIT supports sigma <= 16");
    const size_t nb = (size_t)(p.n_blocks > 0 ? p.n_blocks : 1);
    char* base = static_cast<char*>(ws);
    void* carry = base;
    int32_t* crow = reinterpret_cast<int32_t*>(base + ((nb * vec_size(m->dtype) + 255) / 256) * 256);
    MR_HIP_TRY(k::merbit_spmv(m->dtype, p.sigma, p.policy, m->n_rows, p.n_blocks, m->ci.as<int32_t>(), m->val.p,
                              p.flags.as<uint32_t>(), p.blk_x.as<int32_t>(), p.blk_y.as<int32_t>(), x, y,
                              carry, crow, st));
    return MR_OK;
}

int mr_spmv_ws(mr_mat m, const void* x, void* y, void* ws, mr_stream s)
{
    MR_TRY(check_mat(m));
    if (!ws) return set_error(MR_ERR_INVALID_ARG, "null workspace");
    if ((!x || !y) && m->n_rows > 0) return set_error(MR_ERR_INVALID_ARG, "null x or y");
    return spmv_ws_impl(m, x, y, ws, stream_of(m->ctx, s), true);
}

int mr_spmv_noalloc(mr_mat m, const void* x, void* y, mr_stream s)
{
    MR_TRY(check_mat(m));
    if ((!x || !y) && m->n_rows > 0) return set_error(MR_ERR_INVALID_ARG, "null x or y");
    if (!m->ctx->scratch) return set_error(MR_ERR_INVALID_ARG, "context has no static scratch address");
    const MerbitPlan& p = m->merbit;
    if ((p.policy & 8) && m->dtype != MR_F64 && !p.pn_built)
        return set_error(MR_ERR_INVALID_ARG,
                         "MERBIT plan not built: run mr_spmv, mr_spmv_ws or mr_merbit_autotune on the matrix first");
    size_t need = 0;
    MR_TRY(mr_merbit_ws_bytes(m, &need));
    if (need > k::kScratchBytes)
        return set_error(MR_ERR_UNSUPPORTED,
                         "mr_spmv_noalloc: this plan needs %zu bytes of carries, above the %zu-byte static scratch "
                         "buffer; use mr_spmv_ws",
                         need, k::kScratchBytes);
    return spmv_ws_impl(m, x, y, m->ctx->scratch, stream_of(m->ctx, s), false);
}

int mr_spmm(mr_mat m, int kernel, int k, const void* X, int64_t ldx, void* Y, int64_t ldy, mr_stream s)
{
    MR_TRY(check_mat(m));
    if (k <= 0 || ldx < k || ldy < k) return set_error(MR_ERR_INVALID_ARG, "bad k / leading dimension");
    if ((!X || !Y) && m->n_rows > 0) return set_error(MR_ERR_INVALID_ARG, "null X or Y");
    hipStream_t st = stream_of(m->ctx, s);
    if (k == 1 && ldx == 1 && ldy == 1) return mr_spmv(m, kernel, X, Y, s);
    if (kernel != MR_KERNEL_MERBIT) {
        MR_HIP_TRY(k::csr_spmm(m->dtype, m->n_rows, k, m->rp.as<int32_t>(), m->ci.as<int32_t>(), m->val.p, X,
                               ldx, Y, ldy, st));
        return MR_OK;
    }
    if (m->n_rows == 0) return MR_OK;
    MerbitPlan& p = m->merbit;
    if ((p.policy & 12) && m->dtype != MR_F64 && ldx % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
        k::merbit_lp_mm_fits(p.xs, p.xw, std::min(k, 16))) {
        // lane-padded tiles: 16 columns per pass, X rows cached in LDS; the
        // streaming hint of a column-panel policy belongs to the panels
        MR_TRY(merbit_lp_mm_build(m, st));
        const size_t vb = vec_size(m->dtype);
        const int nt = (p.policy & 4) ? (p.policy & 1) : 0;
        for (int j0 = 0; j0 < k; j0 += 16) {
            const int kk = std::min(16, k - j0);
            MR_HIP_TRY(k::merbit_lp_spmm(m->dtype, p.xs, p.xw, nt, m->n_rows, p.xtiles, p.pcm.as<int32_t>(),
                                         p.pv.p, p.xty.as<int32_t>(), static_cast<const char*>(X) + vb * (size_t)j0,
                                         ldx, kk, merbit_lp_mm_rows(m, kk), static_cast<char*>(Y) + vb * (size_t)j0,
                                         ldy, p.xcarry_mm.p, p.xcarry_row.as<int32_t>(),
                                         std::max(m->ctx->cu_count, 1), st));
        }
        return MR_OK;
    }
    MR_TRY(merbit_mm_grid(m, st));
    // column chunks of at most 64 right-hand sides
    const size_t vb = vec_size(m->dtype);
    for (int j0 = 0; j0 < k; j0 += 64) {
        const int kk = std::min(64, k - j0);
        MR_HIP_TRY(k::merbit_spmm(m->dtype, m->n_rows, m->nnz, p.mm_blocks, m->ci.as<int32_t>(), m->val.p,
                                  p.flags.as<uint32_t>(), p.mm_x.as<int32_t>(), p.mm_y.as<int32_t>(),
                                  static_cast<const char*>(X) + vb * (size_t)j0, ldx,
                                  static_cast<char*>(Y) + vb * (size_t)j0, ldy, kk, p.mm_carry.p,
                                  p.mm_carry_row.as<int32_t>(), st));
    }
    return MR_OK;
}

int mr_merbit_configure(mr_mat m, int sigma)
{
    MR_TRY(check_mat(m));
    if (sigma != 8 && sigma != 16 && sigma != 32)
        return set_error(MR_ERR_INVALID_ARG, "sigma must be 8, 16 or 32 (got %d)", sigma);
    if (m->dtype == MR_F64 && sigma > 16)
        return set_error(MR_ERR_UNSUPPORTED, "F64 MERBIT supports sigma <= 16");
    m->merbit.sigma = sigma;
    return merbit_build(m, m->ctx->stream);
}

int mr_merbit_options(mr_mat m, int policy)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (policy < 0 || policy > 15) return set_error(MR_ERR_INVALID_ARG, "policy must be in [0, 15]");
    if ((policy & 4) && (policy & 8))
        return set_error(MR_ERR_INVALID_ARG, "policy bits 2 (x-cache tiles) and 3 (column panels) exclude each other");
    if ((policy & 12) && m->dtype == MR_F64)
        return set_error(MR_ERR_UNSUPPORTED, "the x-cache and column-panel MERBIT variants need F32/F16/BF16 values");
    m->merbit.policy = policy;
    return MR_OK;
}

int mr_merbit_get_options(mr_mat m, int* policy)
{
    if (!m || !policy) return set_error(MR_ERR_INVALID_ARG, "null argument");
    *policy = m->merbit.policy;
    return MR_OK;
}

int mr_merbit_autotune(mr_mat m, int reps, int* best_sigma, int* best_policy, float* best_us)
{
    MR_TRY(check_mat(m));
    if (m->n_rows == 0) return MR_OK;
    if (reps < 1) reps = 5;
    hipStream_t s = m->ctx->stream;
    const size_t vb = vec_size(m->dtype);
    DevBuf x, y;
    MR_TRY(x.alloc(vb * (size_t)std::max<int64_t>(m->n_cols, 1)));
    MR_TRY(y.alloc(vb * (size_t)m->n_rows));
    MR_HIP_TRY(hipMemsetAsync(x.p, 0, x.bytes, s));
    hipEvent_t e0, e1;
    MR_HIP_TRY(hipEventCreate(&e0));
    MR_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int bs = m->merbit.sigma, bp = m->merbit.policy, bxs = m->merbit.xs, bxw = m->merbit.xw,
        bxc = m->merbit.xc_req, bps = m->merbit.ps, bpw = m->merbit.pw;
    auto time_it = [&](float* us) -> int {
        MR_TRY(mr_spmv(m, MR_KERNEL_MERBIT, x.p, y.p, s));
        MR_HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) MR_TRY(mr_spmv(m, MR_KERNEL_MERBIT, x.p, y.p, s));
        MR_HIP_TRY(hipEventRecord(e1, s));
        MR_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        MR_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *us = ms * 1000.0f / (float)reps;
        return MR_OK;
    };
    // workgroup tiles: every sigma with the four load policies
    const int sigmas[3] = {8, 16, 32};
    int rc = MR_OK;
    for (int si = 0; si < 3 && rc == MR_OK; ++si) {
        const int sg = sigmas[si];
        if (m->dtype == MR_F64 && sg > 16) continue;
        m->merbit.sigma = sg;
        rc = merbit_build(m, s);
        for (int pol = 0; pol < 4 && rc == MR_OK; ++pol) {
            m->merbit.policy = pol;
            float us = 0;
            rc = time_it(&us);
            if (rc == MR_OK && us < best) { best = us; bs = sg; bp = pol; }
        }
    }
    if (rc == MR_OK && bs != m->merbit.sigma) {
        m->merbit.sigma = bs;
        rc = merbit_build(m, s);
    }
    // wave tiles with the full LDS x-cache: (items per lane, waves) x streaming
    // hint.  Smaller caches (more resident workgroups per CU) were slower on
    // every matrix measured, so they are not searched.
    // the best wave-tile shape is kept even when another variant wins: the
    // SpMM of the column-panel variant runs on the wave tiles
    int lxs = bxs, lxw = bxw;
    if (m->dtype != MR_F64) {
        const int shapes[5][2] = {{8, 16}, {16, 16}, {8, 8}, {16, 8}, {4, 16}};
        const int xc0 = m->merbit.xc_req;
        float lbest = 1e30f;
        for (int q = 0; q < 5 && rc == MR_OK; ++q) {
            if (shapes[q][0] == 4 && m->dtype != MR_F32) continue;
            m->merbit.xs = shapes[q][0];
            m->merbit.xw = shapes[q][1];
            m->merbit.xc_req = -1;
            m->merbit.xc_built = false;
            for (int nt = 0; nt < 2 && rc == MR_OK; ++nt) {
                m->merbit.policy = 4 | nt;
                float us = 0;
                rc = time_it(&us);
                if (rc == MR_OK && us < lbest) {
                    lbest = us; lxs = shapes[q][0]; lxw = shapes[q][1];
                }
                if (rc == MR_OK && us < best) {
                    best = us; bp = 4 | nt; bxs = shapes[q][0]; bxw = shapes[q][1]; bxc = -1;
                }
            }
        }
        if (!(bp & 4)) m->merbit.xc_req = xc0;
        // column panels: (items per lane, waves) x streaming hint; the plan
        // depends on the items per lane only
        const int pshapes[3][2] = {{8, 16}, {8, 8}, {16, 8}};
        for (int q = 0; q < 3 && rc == MR_OK; ++q) {
            if (m->merbit.ps != pshapes[q][0]) m->merbit.pn_built = false;
            m->merbit.ps = pshapes[q][0];
            m->merbit.pw = pshapes[q][1];
            for (int nt = 0; nt < 2 && rc == MR_OK; ++nt) {
                m->merbit.policy = 8 | nt;
                float us = 0;
                rc = time_it(&us);
                if (rc == MR_OK && us < best) {
                    best = us; bp = 8 | nt; bps = pshapes[q][0]; bpw = pshapes[q][1];
                }
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != MR_OK) return rc;
    m->merbit.policy = bp;
    if (bp & 8) {
        // panels won: wave tiles keep their own best shape for SpMM
        bxs = lxs;
        bxw = lxw;
        if (m->merbit.ps != bps) m->merbit.pn_built = false;
        m->merbit.ps = bps;
        m->merbit.pw = bpw;
    }
    if (m->merbit.xs != bxs || m->merbit.xw != bxw || ((bp & 4) && m->merbit.xc_req != bxc)) {
        m->merbit.xs = bxs;
        m->merbit.xw = bxw;
        if (bp & 4) m->merbit.xc_req = bxc;
        m->merbit.xc_built = false;
    }
    if (bp & 4) MR_TRY(merbit_xc_build(m, s));
    if (bp & 8) MR_TRY(merbit_panel_build(m, s));
    if (best_sigma) *best_sigma = bs;
    if (best_policy) *best_policy = bp;
    if (best_us) *best_us = best;
    return MR_OK;
}

int mr_merbit_xc_configure(mr_mat m, int items_per_lane, int waves, int cached_cols)
{
    MR_TRY(check_mat(m));
    if (m->dtype == MR_F64) return set_error(MR_ERR_UNSUPPORTED, "the x-cache MERBIT variant needs F32/F16/BF16 values");
    const int S = items_per_lane, W = waves;
    if (!((S == 8 && W == 16) || (S == 16 && W == 16) || (S == 8 && W == 8) || (S == 16 && W == 8) ||
          (S == 4 && W == 16)))
        return set_error(MR_ERR_INVALID_ARG,
                         "x-cache tiles: (items_per_lane, waves) must be (8,16), (16,16), (8,8), (16,8) or (4,16)");
    if (S == 4 && m->dtype != MR_F32)
        return set_error(MR_ERR_UNSUPPORTED, "x-cache tiles of 4 items per lane need F32 values");
    if (cached_cols < -1) return set_error(MR_ERR_INVALID_ARG, "cached_cols must be >= -1");
    m->merbit.xs = S;
    m->merbit.xw = W;
    m->merbit.xc_req = cached_cols;
    m->merbit.xc_built = false;
    return merbit_xc_build(m, m->ctx->stream);
}

int mr_merbit_panel_configure(mr_mat m, int items_per_lane, int waves)
{
    MR_TRY(check_mat(m));
    if (m->dtype == MR_F64)
        return set_error(MR_ERR_UNSUPPORTED, "the column-panel MERBIT variant needs F32/F16/BF16 values");
    const int S = items_per_lane, W = waves;
    if (!((S == 8 && (W == 8 || W == 16)) || (S == 16 && W == 8)))
        return set_error(MR_ERR_INVALID_ARG,
                         "column panels: (items_per_lane, waves) must be (8, 16), (8, 8) or (16, 8)");
    if (m->merbit.ps != S) m->merbit.pn_built = false;
    m->merbit.ps = S;
    m->merbit.pw = W;
    return merbit_panel_build(m, m->ctx->stream);
}

int mr_merbit_panel_info(mr_mat m, int* items_per_lane, int* waves, int64_t* tiles, int64_t* partials,
                         int64_t* segments, int* workgroups, double* build_ms)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    const MerbitPlan& p = m->merbit;
    if (items_per_lane) *items_per_lane = p.ps;
    if (waves) *waves = p.pw;
    if (tiles) *tiles = p.pn_built ? p.ptiles : -1;
    if (partials) *partials = p.pn_built ? p.pparts : -1;
    if (segments) *segments = p.pn_built ? p.psegs : -1;
    if (workgroups) *workgroups = p.pn_built ? p.pn_wg : -1;
    if (build_ms) *build_ms = p.pn_built ? p.pn_build_ms : 0.0;
    return MR_OK;
}

int mr_merbit_stored_bytes(mr_mat m, int64_t* bytes)
{
    if (!m || !bytes) return set_error(MR_ERR_INVALID_ARG, "null argument");
    const MerbitPlan& p = m->merbit;
    const int64_t dsz = (int64_t)dtype_size(m->dtype);
    const int64_t vec = m->dtype == MR_F64 ? 8 : 4;   // x / y elements
    const int64_t xy = (m->n_cols + m->n_rows) * vec;
    *bytes = -1;
    if (p.policy & 8) {
        if (!p.pn_built) return MR_OK;
        const int64_t tile = 64LL * p.ps;
        // item streams and tile bases; partials written, then gathered;
        // row-sum stream and tile rows
        *bytes = p.ptiles * tile * (2 + dsz) + p.ptiles * 4 + 2 * p.pparts * 4 +
                 p.prtiles * k::kRowSumTile * 4 + (p.prtiles + 1) * 4 + xy;
    } else if (p.policy & 4) {
        if (!p.xc_built) return MR_OK;
        // padded (column, value) streams, rows ended before every tile, the
        // cached column ids, carries written and read back
        *bytes = p.xtiles * 64LL * p.xs * (4 + dsz) + p.xtiles * 4 + (int64_t)p.xc * 4 + 2 * p.xtiles * (vec + 4) + xy;
    } else {
        if (!p.built) return MR_OK;
        // columns and values, the row-end bitmap, block coordinates, carries
        *bytes = m->nnz * (4 + dsz) + p.n_blocks * 256LL * p.sigma / 8 + (p.n_blocks + 1) * 8 +
                 2 * p.n_blocks * (vec + 4) + xy;
    }
    return MR_OK;
}

int mr_merbit_xc_info(mr_mat m, int* items_per_lane, int* waves, int* cached_cols, int64_t* tiles)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    const MerbitPlan& p = m->merbit;
    if (items_per_lane) *items_per_lane = p.xs;
    if (waves) *waves = p.xw;
    if (cached_cols) *cached_cols = p.xc_built ? p.xc : -1;
    if (tiles) *tiles = p.xc_built ? p.xtiles : 0;
    return MR_OK;
}

int mr_merbit_info(mr_mat m, int* sigma, int64_t* items, int64_t* n_blocks, double* build_ms)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (sigma) *sigma = m->merbit.sigma;
    if (items) *items = m->merbit.items;
    if (n_blocks) *n_blocks = m->merbit.n_blocks;
    if (build_ms) *build_ms = m->merbit.build_ms;
    return MR_OK;
}

int mr_merbit_export(mr_mat m, uint32_t* flags, int32_t* blk_x, int32_t* blk_y)
{
    MR_TRY(check_mat(m));
    MerbitPlan& p = m->merbit;
    if (!p.built) MR_TRY(merbit_build(m, m->ctx->stream));
    hipStream_t s = m->ctx->stream;
    const int64_t words = p.n_blocks * 256LL * p.sigma / 32;
    if (flags && words > 0)
        MR_HIP_TRY(hipMemcpyAsync(flags, p.flags.p, sizeof(uint32_t) * (size_t)words, hipMemcpyDeviceToHost, s));
    if (blk_x)
        MR_HIP_TRY(hipMemcpyAsync(blk_x, p.blk_x.p, sizeof(int32_t) * (size_t)(p.n_blocks + 1),
                                  hipMemcpyDeviceToHost, s));
    if (blk_y)
        MR_HIP_TRY(hipMemcpyAsync(blk_y, p.blk_y.p, sizeof(int32_t) * (size_t)(p.n_blocks + 1),
                                  hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    return MR_OK;
}

}  // extern "C"
