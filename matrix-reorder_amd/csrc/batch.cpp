// batch.cpp -- fused batches: many MERBIT matrices multiplied in one persistent
// launch plus one carry fix-up, an autotuner over the launch shape, and hipGraph
// capture of the fused launch.
//
// The reference runs one merbit_kernel launch per matrix and iteration
// (class.cu:1421-1432); a batch concatenates the MERBIT tiles of all jobs into
// one global tile range (begin[j] = first tile of job j) so that small and
// large matrices share one grid and the machine stays full.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "kernels/launch.hpp"

using namespace mr;

struct mr_batch_s {
    mr_ctx ctx = nullptr;
    int dtype = MR_F32;
    int k = 1;
    int n = 0;
    std::vector<mr_mat> mats;
    int sigma = 8, nt = 1, grid = 0;
    int64_t total = 0;
    std::vector<int64_t> begin;                 // n + 1 tile offsets (current sigma)
    std::vector<const int32_t*> bx, by;         // merge coordinates per job
    std::vector<const void*> x;
    std::vector<void*> y;
    std::vector<int64_t> ldx, ldy;
    bool io_set = false;
    void* host = nullptr;                       // pinned staging of jobs + begin
    size_t table_bytes = 0, begin_off = 0;
    DevBuf d_table, carry, carry_row;
    hipEvent_t copied = nullptr;                // last staging copy completed
    bool copy_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~mr_batch_s()
    {
        drop_graph();
        if (copied) hipEventDestroy(copied);
        if (host) hipHostFree(host);
    }
    void drop_graph()
    {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    int group() const   // carry width per tile
    {
        if (k == 1) return 1;
        int g = 2;
        while (g < k) g <<= 1;
        return g;
    }
};

namespace {

hipStream_t stream_of(mr_batch b, mr_stream s) { return s ? static_cast<hipStream_t>(s) : b->ctx->stream; }

// tile coordinates of every job for the current sigma / grid kind
int plan_tiles(mr_batch b, hipStream_t st)
{
    b->begin.assign((size_t)b->n + 1, 0);
    b->bx.assign((size_t)b->n, nullptr);
    b->by.assign((size_t)b->n, nullptr);
    for (int j = 0; j < b->n; ++j) {
        mr_mat m = b->mats[(size_t)j];
        int64_t nb = 0;
        if (m->n_rows > 0) {
            if (b->k == 1) {
                MR_TRY(merbit_coords_sigma(m, b->sigma, st, &b->bx[(size_t)j], &b->by[(size_t)j], &nb));
            } else {
                MR_TRY(merbit_mm_grid(m, st));
                b->bx[(size_t)j] = m->merbit.mm_x.as<int32_t>();
                b->by[(size_t)j] = m->merbit.mm_y.as<int32_t>();
                nb = m->merbit.mm_blocks;
            }
        }
        b->begin[(size_t)j + 1] = b->begin[(size_t)j] + nb;
    }
    b->total = b->begin[(size_t)b->n];
    return MR_OK;
}

// write the job table into the pinned staging buffer and copy it on st
int upload_table(mr_batch b, hipStream_t st)
{
    if (b->copy_pending) {
        MR_HIP_TRY(hipEventSynchronize(b->copied));
        b->copy_pending = false;
    }
    auto* jobs = static_cast<k::BatchJob*>(b->host);
    const size_t vb = vec_size(b->dtype);
    const int G = b->group();
    for (int j = 0; j < b->n; ++j) {
        mr_mat m = b->mats[(size_t)j];
        k::BatchJob& J = jobs[j];
        std::memset(&J, 0, sizeof(J));
        J.ci = m->ci.p;
        J.val = m->val.p;
        J.flags = m->merbit.flags.p;
        J.blk_x = b->bx[(size_t)j];
        J.blk_y = b->by[(size_t)j];
        J.x = b->io_set ? b->x[(size_t)j] : nullptr;
        J.y = b->io_set ? b->y[(size_t)j] : nullptr;
        J.carry = static_cast<char*>(b->carry.p) + vb * (size_t)G * (size_t)b->begin[(size_t)j];
        J.carry_row = b->carry_row.as<int32_t>() + b->begin[(size_t)j];
        J.n_rows = m->n_rows;
        J.nnz = m->nnz;
        J.ldx = b->io_set ? b->ldx[(size_t)j] : b->k;
        J.ldy = b->io_set ? b->ldy[(size_t)j] : b->k;
        J.kk = b->k;
    }
    std::memcpy(static_cast<char*>(b->host) + b->begin_off, b->begin.data(), sizeof(int64_t) * b->begin.size());
    MR_HIP_TRY(hipMemcpyAsync(b->d_table.p, b->host, b->table_bytes, hipMemcpyHostToDevice, st));
    MR_HIP_TRY(hipEventRecord(b->copied, st));
    b->copy_pending = true;
    return MR_OK;
}

int launch(mr_batch b, hipStream_t st)
{
    if (b->total == 0) return MR_OK;
    if (!b->io_set) return set_error(MR_ERR_INVALID_ARG, "batch: mr_batch_set_io was not called");
    const int64_t grid = b->grid > 0 ? std::min<int64_t>(b->grid, b->total) : b->total;
    const void* jobs = b->d_table.p;
    const int64_t* begin = reinterpret_cast<const int64_t*>(static_cast<const char*>(b->d_table.p) + b->begin_off);
    if (b->k == 1)
        MR_HIP_TRY(k::merbit_batch_spmv(b->dtype, b->sigma, b->nt, jobs, begin, b->n, b->total, (int)grid, st));
    else
        MR_HIP_TRY(k::merbit_batch_spmm(b->dtype, b->k, jobs, begin, b->n, b->total, (int)grid, st));
    return MR_OK;
}

int apply_config(mr_batch b, int sigma, int nt, int grid)
{
    hipStream_t st = b->ctx->stream;
    b->drop_graph();
    b->sigma = sigma;
    b->nt = nt;
    b->grid = grid;
    MR_TRY(plan_tiles(b, st));
    MR_TRY(upload_table(b, st));
    MR_HIP_TRY(hipStreamSynchronize(st));
    return MR_OK;
}

}  // namespace

extern "C" {

int mr_batch_create(mr_ctx ctx, int n_jobs, const mr_mat* mats, int k, mr_batch* out)
{
    if (!ctx || !out || (n_jobs > 0 && !mats)) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (n_jobs <= 0 || n_jobs > k::kMaxBatchJobs)
        return set_error(MR_ERR_INVALID_ARG, "batch: n_jobs must be in [1, %d] (got %d)", k::kMaxBatchJobs, n_jobs);
    if (k < 1 || k > 64) return set_error(MR_ERR_UNSUPPORTED, "batch: k must be in [1, 64] (got %d)", k);
    MR_HIP_TRY(hipSetDevice(ctx->device));
    auto b = std::make_unique<mr_batch_s>();
    b->ctx = ctx;
    b->k = k;
    b->n = n_jobs;
    b->mats.assign(mats, mats + n_jobs);
    for (int j = 0; j < n_jobs; ++j) {
        mr_mat m = mats[j];
        if (!m) return set_error(MR_ERR_INVALID_ARG, "batch: job %d has a null matrix", j);
        if (m->ctx != ctx) return set_error(MR_ERR_INVALID_ARG, "batch: job %d belongs to another context", j);
        if (j == 0) b->dtype = m->dtype;
        if (m->dtype != b->dtype)
            return set_error(MR_ERR_UNSUPPORTED, "batch: job %d has dtype %d, job 0 has %d", j, m->dtype, b->dtype);
    }
    b->sigma = 8;
    // carry arenas sized for the finest tile grid (sigma 8 for SpMV)
    int64_t max_tiles = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const int64_t items = mats[j]->n_rows + mats[j]->nnz;
        max_tiles += (k == 1) ? (items + 2047) / 2048 : (items + 4095) / 4096;
    }
    const size_t vb = vec_size(b->dtype);
    MR_TRY(b->carry.alloc(vb * (size_t)b->group() * (size_t)std::max<int64_t>(max_tiles, 1)));
    MR_TRY(b->carry_row.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(max_tiles, 1)));
    b->begin_off = ((sizeof(k::BatchJob) * (size_t)n_jobs + 255) / 256) * 256;
    b->table_bytes = b->begin_off + sizeof(int64_t) * ((size_t)n_jobs + 1);
    MR_HIP_TRY(hipHostMalloc(&b->host, b->table_bytes, hipHostMallocDefault));
    MR_TRY(b->d_table.alloc(b->table_bytes));
    MR_HIP_TRY(hipEventCreateWithFlags(&b->copied, hipEventDisableTiming));
    MR_TRY(apply_config(b.get(), b->sigma, 1, 0));
    *out = b.release();
    return MR_OK;
}

int mr_batch_destroy(mr_batch b)
{
    if (!b) return MR_OK;
    hipSetDevice(b->ctx->device);
    if (b->copy_pending) hipEventSynchronize(b->copied);
    delete b;
    return MR_OK;
}

int mr_batch_set_io(mr_batch b, const void* const* x, const int64_t* ldx, void* const* y, const int64_t* ldy,
                    mr_stream s)
{
    if (!b || !x || !y) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    b->x.assign(x, x + b->n);
    b->y.assign(y, y + b->n);
    b->ldx.assign((size_t)b->n, b->k);
    b->ldy.assign((size_t)b->n, b->k);
    for (int j = 0; j < b->n; ++j) {
        if (ldx) b->ldx[(size_t)j] = ldx[j];
        if (ldy) b->ldy[(size_t)j] = ldy[j];
        if (b->ldx[(size_t)j] < b->k || b->ldy[(size_t)j] < b->k)
            return set_error(MR_ERR_INVALID_ARG, "batch: job %d leading dimension below k", j);
        if (b->mats[(size_t)j]->n_rows > 0 && (!x[j] || !y[j]))
            return set_error(MR_ERR_INVALID_ARG, "batch: job %d has a null x or y", j);
    }
    b->io_set = true;
    return upload_table(b, stream_of(b, s));
}

int mr_batch_configure(mr_batch b, int sigma, int nt, int grid)
{
    if (!b) return set_error(MR_ERR_INVALID_ARG, "null batch");
    if (sigma != 8 && sigma != 16 && sigma != 32)
        return set_error(MR_ERR_INVALID_ARG, "sigma must be 8, 16 or 32 (got %d)", sigma);
    if (b->dtype == MR_F64 && sigma > 16) return set_error(MR_ERR_UNSUPPORTED, "F64 MERBIT supports sigma <= 16");
    if (grid < 0) return set_error(MR_ERR_INVALID_ARG, "grid must be >= 0");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    return apply_config(b, sigma, nt ? 1 : 0, grid);
}

int mr_batch_autotune(mr_batch b, int reps, int* sigma, int* nt, int* grid, float* best_us)
{
    if (!b) return set_error(MR_ERR_INVALID_ARG, "null batch");
    if (!b->io_set) return set_error(MR_ERR_INVALID_ARG, "batch: mr_batch_set_io must precede autotuning");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    if (reps < 1) reps = 5;
    hipStream_t st = b->ctx->stream;
    const int cu = std::max(b->ctx->cu_count, 1);
    std::vector<int> sigmas = {8, 16, 32};
    if (b->k > 1 || b->dtype == MR_F64) sigmas = {8, 16};
    if (b->k > 1) sigmas = {b->sigma};   // the SpMM grid does not depend on sigma
    const std::vector<int> nts = b->k == 1 ? std::vector<int>{0, 1} : std::vector<int>{0};
    const int grids[4] = {0, 2 * cu, 4 * cu, 8 * cu};
    hipEvent_t e0, e1;
    MR_HIP_TRY(hipEventCreate(&e0));
    MR_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int bs = b->sigma, bn = b->nt, bg = b->grid;
    int rc = MR_OK;
    for (int sg : sigmas) {
        for (int t : nts) {
            for (int g : grids) {
                rc = apply_config(b, sg, t, g);
                if (rc == MR_OK) rc = launch(b, st);   // warm-up
                if (rc == MR_OK && hipEventRecord(e0, st) != hipSuccess) rc = MR_ERR_HIP;
                for (int r = 0; r < reps && rc == MR_OK; ++r) rc = launch(b, st);
                if (rc == MR_OK && hipEventRecord(e1, st) != hipSuccess) rc = MR_ERR_HIP;
                float ms = 0;
                if (rc == MR_OK && hipEventSynchronize(e1) != hipSuccess) rc = MR_ERR_HIP;
                if (rc == MR_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = MR_ERR_HIP;
                if (rc != MR_OK) break;
                const float us = ms * 1000.0f / (float)reps;
                if (us < best) { best = us; bs = sg; bn = t; bg = g; }
            }
            if (rc != MR_OK) break;
        }
        if (rc != MR_OK) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != MR_OK) return rc == MR_ERR_HIP ? set_error(MR_ERR_HIP, "batch autotune: HIP event failure") : rc;
    MR_TRY(apply_config(b, bs, bn, bg));
    if (sigma) *sigma = bs;
    if (nt) *nt = bn;
    if (grid) *grid = bg;
    if (best_us) *best_us = best;
    return MR_OK;
}

int mr_batch_run(mr_batch b, mr_stream s)
{
    if (!b) return set_error(MR_ERR_INVALID_ARG, "null batch");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    return launch(b, stream_of(b, s));
}

int mr_batch_graph_capture(mr_batch b, mr_stream s)
{
    if (!b || !s) return set_error(MR_ERR_INVALID_ARG, "batch graph capture needs a batch and a non-NULL stream");
    if (!b->io_set) return set_error(MR_ERR_INVALID_ARG, "batch: mr_batch_set_io must precede capture");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    hipStream_t st = static_cast<hipStream_t>(s);
    b->drop_graph();
    // the staging copy must not become part of the graph
    MR_HIP_TRY(hipStreamSynchronize(st));
    if (b->copy_pending) {
        MR_HIP_TRY(hipEventSynchronize(b->copied));
        b->copy_pending = false;
    }
    MR_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = launch(b, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != MR_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(MR_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    b->graph = g;
    MR_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return MR_OK;
}

int mr_batch_graph_launch(mr_batch b, mr_stream s)
{
    if (!b) return set_error(MR_ERR_INVALID_ARG, "null batch");
    if (!b->exec) return set_error(MR_ERR_INVALID_ARG, "batch: no captured graph");
    MR_HIP_TRY(hipSetDevice(b->ctx->device));
    MR_HIP_TRY(hipGraphLaunch(b->exec, stream_of(b, s)));
    return MR_OK;
}

int mr_batch_info(mr_batch b, int64_t* tiles, int64_t* nnz, double* bytes, double* flops)
{
    if (!b) return set_error(MR_ERR_INVALID_ARG, "null batch");
    int64_t z = 0;
    double by = 0.0;
    const double vb = (double)vec_size(b->dtype);
    for (mr_mat m : b->mats) {
        z += m->nnz;
        by += (double)m->nnz * (double)(dtype_size(m->dtype) + 4) + (double)(m->n_rows + 1) * 4.0 +
              (double)b->k * vb * (double)(m->n_cols + m->n_rows);
    }
    if (tiles) *tiles = b->total;
    if (nnz) *nnz = z;
    if (bytes) *bytes = by;
    if (flops) *flops = 2.0 * (double)z * (double)b->k;
    return MR_OK;
}

}  // extern "C"
