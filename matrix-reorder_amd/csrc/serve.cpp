// serve.cpp -- the serving loop: reordered matrices resident in HBM, SpMV /
// SpMM requests batched continuously into fused MERBIT launches on several
// streams, admission by memory, background uploads that run pGOrder, and a
// Poisson load generator that reports latency percentiles.
//
// The reference has no serving path (its main.cu runs one matrix per process);
// this file is built on the batched kernels of the library.
//
// Threads: the caller's threads submit and wait; one dispatcher thread forms
// batches and retires them; one upload thread generates, admits and reorders
// new matrices on the context's stream.  All server state is guarded by one
// mutex; HIP calls are made with it released.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "kernels/launch.hpp"
#include "kernels/launch_panel.hpp"
#include "kernels/launch_serve.hpp"
#include "parallel.hpp"

using namespace mr;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kClasses = 7;   // K = 1 (SpMV), then carry widths G = 2, 4, ..., 64
// requests that stream at least this many bytes (matrix, X and Y) are heavy:
// they queue apart, and light requests are dispatched first
constexpr double kHeavyBytes = 32.0 * 1024 * 1024;
// effective rates of heavy single-vector and multi-vector jobs before any
// heavy batch has completed (bytes per us); then learned from batch times
constexpr double kHeavyRate0[2] = {4.0e6, 1.0e6};

struct SMat {
    mr_mat m = nullptr;
    bool owned = false;       // uploaded (and freed) by the server
    bool live = true;
    int64_t bytes = 0;
    int inflight = 0;         // queued or running requests
    uint64_t last_use = 0;    // dispatch stamp, for least-recently-used eviction
};

struct Req {
    int mat = 0, k = 1;
    uint64_t seed = 0;
    double t_arr = 0.0, t_done = 0.0;   // microseconds since the server epoch
    int status = -1;                    // -1 while queued or running
    uint64_t hash = 0;
    double flops = 0.0, bytes = 0.0;
    bool heavy = false;                 // bytes >= kHeavyBytes
    std::vector<float> y;               // verify level 2
};

// one fused job: a matrix and the requests coalesced onto it (columns side by
// side: request i owns columns [off_i, off_i + k_i) of X and Y, ld = K)
struct Job {
    int mat = 0;
    mr_mat m = nullptr;
    int K = 0;
    bool lp = false;                  // runs through the lane-padded kernels
    bool pn = false;                  // runs through the column-panel kernels
    int64_t ldx = 0;
    std::vector<int64_t> reqs;
    std::vector<int> off;             // first column of each request
    std::vector<int> ks;              // its k and seed (copied under the lock)
    std::vector<uint64_t> seeds;
    size_t x_off = 0, y_off = 0, c_off = 0, r_off = 0;
    int64_t tiles = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf arena;                 // [tables | hashes | job data]
    void* pinned = nullptr;       // [tables staging | hash readback]
    bool busy = false;
    bool heavy = false;           // the batch holds heavy requests
    double t_launch = 0.0;        // us since the server epoch
    double est_us = 0.0;          // predicted service time (heavy batches)
    int kinds = 0;                // bit 0: a single-vector job, bit 1: a multi-vector job
    std::vector<Job> jobs;
    std::vector<int64_t> reqs;    // request-table order
};

struct UploadTask {
    int scale = 18, ef = 16;
    uint64_t seed = 0;
    int ticket = 0;
};

struct UploadRes {
    int status = -1;
    int mat = -1;
    double ms = 0.0;
};

int carry_group(int K)
{
    if (K == 1) return 1;
    int g = 2;
    while (g < K) g <<= 1;
    return g;
}

int class_of(int K)
{
    if (K == 1) return 0;
    int c = 1, g = 2;
    while (g < K) { g <<= 1; ++c; }
    return c;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// single-vector jobs on matrices tuned to the column-panel variant run its two
// kernels (one launch each, partial sums in the slot's arena)
bool use_panel(const mr_mat m, int K)
{
    return K == 1 && (m->merbit.policy & 8) && m->dtype != MR_F64 && m->merbit.pn_built;
}

// other jobs on matrices with a lane-padded or column-panel plan run through
// the lane-padded kernels (one launch each) instead of the fused batch
bool use_lp(const mr_mat m, int K)
{
    return (m->merbit.policy & 12) && m->dtype != MR_F64 && m->merbit.xc_built && m->merbit.pcm_built &&
           (K == 1 || k::merbit_lp_mm_fits(m->merbit.xs, m->merbit.xw, std::min(K, 16)));
}

int64_t lp_ldx(int K) { return K == 1 ? 1 : (K + 3) / 4 * 4; }

}  // namespace

struct mr_server_s {
    mr_ctx ctx = nullptr;
    mr_serve_config cfg{};
    int dtype = -1;
    Clock::time_point epoch;
    std::mutex mu;
    std::condition_variable cv;        // dispatcher: work arrived / stop
    std::condition_variable done_cv;   // waiters: a request or upload finished
    std::condition_variable up_cv;     // upload thread: task arrived / stop
    std::vector<SMat> mats;
    std::deque<Req> reqs;              // index = request id
    std::deque<int64_t> queue;         // light requests
    std::deque<int64_t> hqueue;        // heavy requests
    std::vector<Slot> slots;
    int heavy_busy = 0;                // slots running a heavy batch
    // bytes per us of heavy single-vector [0] and multi-vector [1] jobs
    double heavy_rate[2] = {kHeavyRate0[0], kHeavyRate0[1]};
    std::deque<UploadTask> up_tasks;
    std::vector<UploadRes> up_res;
    std::thread dispatcher, uploader;
    bool stop = false;
    bool dead = false;                 // a HIP failure stopped the dispatcher
    std::string dead_msg;
    uint64_t stamp = 0;
    int64_t batches = 0, batch_jobs = 0, batch_reqs = 0, evictions = 0;
    int64_t uploaded_bytes = 0;
    std::string plan_dir;                 // plan cache of the uploads ("" = off)
    int64_t plan_hits = 0, plan_misses = 0;
    size_t table_bytes = 0;            // table region of an arena / staging buffer
    size_t hash_off = 0;               // hash region inside the table region

    double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - epoch).count(); }
};

namespace {

// device bytes of a job with K columns on matrix m
size_t job_bytes(const mr_mat m, int K, int sigma, int64_t* tiles)
{
    if (use_panel(m, K)) {   // x, y and one float per partial sum
        if (tiles) *tiles = 0;
        return align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_cols, 1)) +
               align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_rows, 1)) +
               align256(sizeof(float) * (size_t)std::max<int64_t>(m->merbit.pparts, 1)) + align256(sizeof(int32_t));
    }
    if (use_lp(m, K)) {
        const int64_t t = m->merbit.xtiles;
        if (tiles) *tiles = t;
        return align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_cols, 1) * (size_t)lp_ldx(K)) +
               align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_rows, 1) * (size_t)K) +
               align256(sizeof(float) * (size_t)(K == 1 ? 1 : 16) * (size_t)std::max<int64_t>(t, 1)) +
               align256(sizeof(int32_t) * (size_t)std::max<int64_t>(t, 1));
    }
    const int64_t items = m->n_rows + m->nnz;
    const int64_t t = (K == 1) ? (items + 256LL * sigma - 1) / (256LL * sigma) : (items + 4095) / 4096;
    if (tiles) *tiles = t;
    const size_t G = (size_t)carry_group(K);
    return align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_cols, 1) * (size_t)K) +
           align256(sizeof(float) * (size_t)std::max<int64_t>(m->n_rows, 1) * (size_t)K) +
           align256(sizeof(float) * G * (size_t)std::max<int64_t>(t, 1)) +
           align256(sizeof(int32_t) * (size_t)std::max<int64_t>(t, 1));
}

double req_bytes(const mr_mat m, int k)
{
    return (double)m->nnz * (double)(dtype_size(m->dtype) + 4) + (double)(m->n_rows + 1) * 4.0 +
           (double)k * 4.0 * (double)(m->n_cols + m->n_rows);
}

// final status of a request (lock held); wakes its waiters
void finish_req(mr_server s, int64_t r, int status, double t)
{
    Req& R = s->reqs[(size_t)r];
    R.status = status;
    R.t_done = t;
    SMat& M = s->mats[(size_t)R.mat];
    if (M.inflight > 0) --M.inflight;
    s->done_cv.notify_all();
}

// pop requests from the light or the heavy queue into jobs that fit the
// slot's arena (lock held).  With a deadline (timeout_ms > 0):
//  * a request queued past its deadline is shed (MR_ERR_TIMEOUT);
//  * a heavy batch is predicted to take the bytes each job streams over the
//    learned rate of its kind (one vector or several); a heavy request that
//    would finish late even in a batch of its own is shed, and the batch stops
//    growing before a request would make any of its members finish late (the
//    rest wait for the next batch).
void form_batch(mr_server s, Slot& sl, bool heavy)
{
    std::deque<int64_t>& q = heavy ? s->hqueue : s->queue;
    const double now = s->now_us();
    const double tmo = s->cfg.timeout_ms * 1000.0;
    const size_t cap = sl.arena.bytes - s->table_bytes;
    size_t used = 0;
    int n_reqs = 0;
    double est = 0.0;                         // predicted service time so far (us)
    int kinds = 0;
    double first_deadline = 1e300;            // earliest deadline in the batch
    std::unordered_map<int, size_t> open;   // matrix -> job still below 64 columns
    sl.jobs.clear();
    sl.reqs.clear();
    while (!q.empty() && n_reqs < s->cfg.max_jobs) {
        const int64_t r = q.front();
        Req& R = s->reqs[(size_t)r];
        if (tmo > 0 && now - R.t_arr > tmo) {
            q.pop_front();
            finish_req(s, r, MR_ERR_TIMEOUT, now);
            continue;
        }
        if (R.status >= 0) {   // cancelled while queued
            q.pop_front();
            continue;
        }
        const mr_mat m = s->mats[(size_t)R.mat].m;
        size_t j = SIZE_MAX, need;
        if (s->cfg.coalesce) {
            auto it = open.find(R.mat);
            if (it != open.end() && sl.jobs[it->second].K + R.k <= 64) j = it->second;
        }
        // time the request adds: the whole request on a new job, only its
        // vectors when it joins a job on the same matrix (which then runs as a
        // multi-vector job)
        const int kind = (j == SIZE_MAX && R.k == 1) ? 0 : 1;
        const double add = (j == SIZE_MAX ? R.bytes : 4.0 * R.k * (double)(m->n_cols + m->n_rows)) /
                           s->heavy_rate[kind];
        if (heavy && tmo > 0) {
            const double done_at = now + est + add;
            const double dl = R.t_arr + tmo;
            if (done_at > dl || done_at > first_deadline) {
                if (sl.jobs.empty()) {   // late even on its own
                    q.pop_front();
                    finish_req(s, r, MR_ERR_TIMEOUT, now);
                    continue;
                }
                break;
            }
        }
        if (j != SIZE_MAX) {
            const int K = sl.jobs[j].K;
            const size_t grown = job_bytes(m, K + R.k, s->cfg.sigma, nullptr);
            const size_t before = job_bytes(m, K, s->cfg.sigma, nullptr);
            need = grown > before ? grown - before : 0;
        } else {
            need = job_bytes(m, R.k, s->cfg.sigma, nullptr);
        }
        if (used + need > cap) {
            if (sl.jobs.empty()) {   // can never fit: refuse it
                q.pop_front();
                finish_req(s, r, MR_ERR_ADMISSION, now);
                continue;
            }
            break;
        }
        q.pop_front();
        used += need;
        est += add;
        kinds |= 1 << kind;
        if (tmo > 0) first_deadline = std::min(first_deadline, R.t_arr + tmo);
        if (j == SIZE_MAX) {
            Job J;
            J.mat = R.mat;
            J.m = m;
            sl.jobs.push_back(J);
            j = sl.jobs.size() - 1;
            open[R.mat] = j;
        }
        Job& J = sl.jobs[j];
        J.off.push_back(J.K);
        J.reqs.push_back(r);
        J.ks.push_back(R.k);
        J.seeds.push_back(R.seed);
        J.K += R.k;
        if (J.K >= 64) open.erase(R.mat);
        s->mats[(size_t)R.mat].last_use = ++s->stamp;
        ++n_reqs;
    }
    sl.heavy = heavy;
    sl.est_us = est;
    sl.kinds = kinds;
}

// write the tables, copy them, and launch the batch on the slot's stream
// (lock released; the slot is marked busy by the caller)
int launch_slot(mr_server s, Slot& sl, hipStream_t st)
{
    const int sigma = s->cfg.sigma;
    const size_t vb = sizeof(float);
    // job data layout
    size_t cur = s->table_bytes;
    for (Job& J : sl.jobs) {
        const mr_mat m = J.m;
        J.pn = use_panel(m, J.K);
        J.lp = !J.pn && use_lp(m, J.K);
        J.ldx = J.lp ? lp_ldx(J.K) : J.K;
        const int G = J.lp ? (J.K == 1 ? 1 : 16) : carry_group(J.K);
        job_bytes(m, J.K, sigma, &J.tiles);
        J.x_off = cur;
        cur += align256(vb * (size_t)std::max<int64_t>(m->n_cols, 1) * (size_t)J.ldx);
        J.y_off = cur;
        cur += align256(vb * (size_t)std::max<int64_t>(m->n_rows, 1) * (size_t)J.K);
        J.c_off = cur;   // carries, or the partial sums of a column-panel job
        cur += J.pn ? align256(vb * (size_t)std::max<int64_t>(m->merbit.pparts, 1))
                    : align256(vb * (size_t)G * (size_t)std::max<int64_t>(J.tiles, 1));
        J.r_off = cur;
        cur += align256(sizeof(int32_t) * (size_t)std::max<int64_t>(J.tiles, 1));
    }
    char* dev = static_cast<char*>(sl.arena.p);
    char* host = static_cast<char*>(sl.pinned);
    size_t off = 0;
    // per-class job tables
    struct ClassTab { size_t jobs_off = 0, begin_off = 0; int n = 0; int64_t total = 0; };
    ClassTab ct[kClasses];
    std::vector<std::vector<size_t>> members(kClasses);
    for (size_t j = 0; j < sl.jobs.size(); ++j)
        if (!sl.jobs[j].lp && !sl.jobs[j].pn) members[(size_t)class_of(sl.jobs[j].K)].push_back(j);
    for (int c = 0; c < kClasses; ++c) {
        if (members[(size_t)c].empty()) continue;
        ClassTab& T = ct[c];
        T.n = (int)members[(size_t)c].size();
        T.jobs_off = off;
        off = align256(off + sizeof(k::BatchJob) * (size_t)T.n);
        T.begin_off = off;
        off = align256(off + sizeof(int64_t) * ((size_t)T.n + 1));
        auto* tab = reinterpret_cast<k::BatchJob*>(host + T.jobs_off);
        auto* beg = reinterpret_cast<int64_t*>(host + T.begin_off);
        beg[0] = 0;
        for (int i = 0; i < T.n; ++i) {
            const Job& J = sl.jobs[members[(size_t)c][(size_t)i]];
            const mr_mat m = J.m;
            k::BatchJob& B = tab[i];
            std::memset(&B, 0, sizeof(B));
            const int32_t* bx = nullptr;
            const int32_t* by = nullptr;
            int64_t nb = 0;
            if (m->n_rows > 0) {
                if (J.K == 1) {
                    MR_TRY(merbit_coords_sigma(m, sigma, st, &bx, &by, &nb));
                } else {
                    MR_TRY(merbit_mm_grid(m, st));
                    bx = m->merbit.mm_x.as<int32_t>();
                    by = m->merbit.mm_y.as<int32_t>();
                    nb = m->merbit.mm_blocks;
                }
            }
            B.ci = m->ci.p;
            B.val = m->val.p;
            B.flags = m->merbit.flags.p;
            B.blk_x = bx;
            B.blk_y = by;
            B.x = dev + J.x_off;
            B.y = dev + J.y_off;
            B.carry = dev + J.c_off;
            B.carry_row = dev + J.r_off;
            B.n_rows = m->n_rows;
            B.nnz = m->nnz;
            B.ldx = J.K;
            B.ldy = J.K;
            B.kk = J.K;
            beg[i + 1] = beg[i] + nb;
        }
        T.total = beg[T.n];
    }
    // request table (right-hand sides and hashes)
    const int nr = (int)sl.reqs.size();
    const size_t req_off = off;
    off = align256(off + sizeof(k::ServeReq) * (size_t)nr);
    const size_t rb_off = off;
    off = align256(off + sizeof(int64_t) * ((size_t)nr + 1));
    const size_t hb_off = off;
    off = align256(off + sizeof(int64_t) * ((size_t)nr + 1));
    auto* rq = reinterpret_cast<k::ServeReq*>(host + req_off);
    auto* rb = reinterpret_cast<int64_t*>(host + rb_off);
    auto* hb = reinterpret_cast<int64_t*>(host + hb_off);
    rb[0] = 0;
    hb[0] = 0;
    int i = 0;
    for (const Job& J : sl.jobs) {
        for (size_t q = 0; q < J.reqs.size(); ++q, ++i) {
            k::ServeReq& E = rq[i];
            E.x = dev + J.x_off + vb * (size_t)J.off[q];
            E.y = dev + J.y_off + vb * (size_t)J.off[q];
            E.ldx = J.ldx;
            E.ldy = J.K;
            E.n_cols = J.m->n_cols;
            E.n_rows = J.m->n_rows;
            E.k = J.ks[q];
            E.pad = 0;
            E.seed = J.seeds[q];
            rb[i + 1] = rb[i] + J.m->n_cols * (int64_t)J.ks[q];
            hb[i + 1] = hb[i] + J.m->n_rows * (int64_t)J.ks[q];
        }
    }
    MR_HIP_TRY(hipMemcpyAsync(dev, host, off, hipMemcpyHostToDevice, st));
    MR_HIP_TRY(k::serve_rhs(reinterpret_cast<const k::ServeReq*>(dev + req_off),
                            reinterpret_cast<const int64_t*>(dev + rb_off), nr, rb[nr], st));
    const int grid_cfg = s->cfg.grid;
    for (int c = 0; c < kClasses; ++c) {
        const ClassTab& T = ct[c];
        if (T.n == 0 || T.total == 0) continue;
        const int64_t grid = grid_cfg > 0 ? std::min<int64_t>(grid_cfg, T.total) : T.total;
        const void* tab = dev + T.jobs_off;
        const int64_t* beg = reinterpret_cast<const int64_t*>(dev + T.begin_off);
        if (c == 0)
            MR_HIP_TRY(k::merbit_batch_spmv(s->dtype, sigma, 1, tab, beg, T.n, T.total, (int)grid, st));
        else
            MR_HIP_TRY(k::merbit_batch_spmm(s->dtype, 1 << c, tab, beg, T.n, T.total, (int)grid, st));
    }
    // large matrices: one launch per job -- the column-panel kernels for a
    // single vector, else the lane-padded ones (SpMM in 16-column passes)
    for (const Job& J : sl.jobs) {
        if (J.pn) {
            MR_HIP_TRY(k::panel_spmv(J.m->dtype, panel_view(J.m), dev + J.x_off, dev + J.c_off, dev + J.y_off, st));
            continue;
        }
        if (!J.lp) continue;
        const mr_mat m = J.m;
        const MerbitPlan& p = m->merbit;
        // persistent launches of one workgroup per CU would hold every CU for
        // the whole pass (up to ~0.8 ms at R-MAT s22, k = 16): with other slots
        // an eighth of the CUs is left to the light batches running beside them
        const int cus = std::max(s->ctx->cu_count, 1);
        const int grid = s->slots.size() > 1 ? std::max(1, cus - cus / 8) : cus;
        if (J.K == 1) {
            MR_HIP_TRY(k::merbit_lp_spmv(m->dtype, p.xs, p.xw, p.policy & 1, m->n_rows, p.xtiles, p.pc.as<int32_t>(),
                                         p.pv.p, p.xty.as<int32_t>(), dev + J.x_off, p.hot.as<int32_t>(), p.xc,
                                         dev + J.y_off, dev + J.c_off, reinterpret_cast<int32_t*>(dev + J.r_off), grid,
                                         st));
        } else {
            // the streaming hint of a column-panel policy belongs to the panels
            const int nt = (p.policy & 4) ? (p.policy & 1) : 0;
            for (int j0 = 0; j0 < J.K; j0 += 16) {
                const int kk = std::min(16, J.K - j0);
                MR_HIP_TRY(k::merbit_lp_spmm(m->dtype, p.xs, p.xw, nt, m->n_rows, p.xtiles,
                                             p.pcm.as<int32_t>(), p.pv.p, p.xty.as<int32_t>(),
                                             dev + J.x_off + vb * (size_t)j0, J.ldx, kk, merbit_lp_mm_rows(m, kk),
                                             dev + J.y_off + vb * (size_t)j0, J.K, dev + J.c_off,
                                             reinterpret_cast<int32_t*>(dev + J.r_off), grid, st));
            }
        }
    }
    if (s->cfg.verify) {
        auto* hd = reinterpret_cast<uint64_t*>(dev + s->hash_off);
        MR_HIP_TRY(k::serve_hash(reinterpret_cast<const k::ServeReq*>(dev + req_off),
                                 reinterpret_cast<const int64_t*>(dev + hb_off), nr, hb[nr], hd, st));
        MR_HIP_TRY(hipMemcpyAsync(host + s->table_bytes, hd, sizeof(uint64_t) * (size_t)nr, hipMemcpyDeviceToHost,
                                  st));
    }
    MR_HIP_TRY(hipEventRecord(sl.done, st));
    return MR_OK;
}

// completed slot: stamp every request (lock held; releases it for verify 2)
void retire_slot(mr_server s, Slot& sl, std::unique_lock<std::mutex>& lk)
{
    const double t = s->now_us();
    const auto* hashes = reinterpret_cast<const uint64_t*>(static_cast<char*>(sl.pinned) + s->table_bytes);
    if (s->cfg.verify >= 2) {
        // copy every result out while the slot is still reserved
        std::vector<std::pair<int64_t, std::vector<float>>> outs;
        lk.unlock();
        for (const Job& J : sl.jobs) {
            // the job's whole n_rows x K block in one contiguous copy, split
            // into its requests' columns on the host (a strided 2-D copy of
            // millions of short rows per request is far slower)
            const size_t nr = (size_t)J.m->n_rows, K = (size_t)J.K;
            std::vector<float> blk(nr * K);
            if (!blk.empty())
                hipMemcpy(blk.data(), static_cast<char*>(sl.arena.p) + J.y_off, sizeof(float) * nr * K,
                          hipMemcpyDeviceToHost);
            for (size_t q = 0; q < J.reqs.size(); ++q) {
                const int64_t r = J.reqs[q];
                const size_t k = (size_t)J.ks[q], off = (size_t)J.off[q];
                std::vector<float> y(nr * k);
                for (size_t i = 0; i < nr; ++i)
                    for (size_t c = 0; c < k; ++c) y[i * k + c] = blk[i * K + off + c];
                outs.emplace_back(r, std::move(y));
            }
        }
        lk.lock();
        for (auto& o : outs) s->reqs[(size_t)o.first].y = std::move(o.second);
    }
    int i = 0;
    for (const Job& J : sl.jobs) {
        for (int64_t r : J.reqs) {
            if (s->cfg.verify) s->reqs[(size_t)r].hash = hashes[i];
            ++i;
            finish_req(s, r, MR_OK, t);
        }
    }
    ++s->batches;
    s->batch_jobs += (int64_t)sl.jobs.size();
    s->batch_reqs += i;
    if (sl.heavy) {
        // launch -> retire of a heavy batch, sharing the device with the other
        // slots as it did, against its prediction: the rates of the kinds it
        // ran move a fifth of the way (in log space) toward what it measured
        const double dt = t - sl.t_launch;
        if (dt > 0.0 && sl.est_us > 0.0) {
            const double f = std::pow(sl.est_us / dt, 0.2);
            for (int kd = 0; kd < 2; ++kd)
                if (sl.kinds & (1 << kd)) s->heavy_rate[kd] = std::min(1.0e8, std::max(1.0e4, s->heavy_rate[kd] * f));
        }
        --s->heavy_busy;
    }
    sl.busy = false;
    sl.heavy = false;
    sl.jobs.clear();
    sl.reqs.clear();
    s->done_cv.notify_all();
}

void fail_all(mr_server s, const std::string& msg)
{
    s->dead = true;
    s->dead_msg = msg;
    const double t = s->now_us();
    for (Slot& sl : s->slots) {
        if (!sl.busy) continue;
        for (const Job& J : sl.jobs)
            for (int64_t r : J.reqs) finish_req(s, r, MR_ERR_HIP, t);
        sl.busy = false;
        sl.heavy = false;
        sl.jobs.clear();
    }
    s->heavy_busy = 0;
    for (std::deque<int64_t>* q : {&s->queue, &s->hqueue})
        while (!q->empty()) {
            finish_req(s, q->front(), MR_ERR_HIP, t);
            q->pop_front();
        }
    s->done_cv.notify_all();
}

void dispatch_loop(mr_server s)
{
    hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool busy = false;
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                retire_slot(s, sl, lk);
            } else if (q == hipErrorNotReady) {
                busy = true;
            } else {
                fail_all(s, std::string("batch failed: ") + hipGetErrorString(q));
                busy = false;
                break;
            }
        }
        if (!s->dead) {
            for (Slot& sl : s->slots) {
                if (sl.busy) continue;
                // light requests first; heavy batches leave a slot to them
                bool heavy;
                if (!s->queue.empty()) heavy = false;
                else if (!s->hqueue.empty() &&
                         s->heavy_busy < std::max(1, std::min(s->cfg.heavy_slots, (int)s->slots.size() - 1)))
                    heavy = true;
                else continue;
                form_batch(s, sl, heavy);
                if (sl.jobs.empty()) continue;
                for (const Job& J : sl.jobs)
                    for (int64_t r : J.reqs) sl.reqs.push_back(r);
                sl.busy = true;
                sl.t_launch = s->now_us();
                if (sl.heavy) ++s->heavy_busy;
                busy = true;
                lk.unlock();
                const int rc = launch_slot(s, sl, sl.stream);
                lk.lock();
                if (rc != MR_OK) {
                    fail_all(s, std::string("batch launch failed: ") + mr_last_error());
                    break;
                }
            }
        }
        const bool queued = !s->queue.empty() || !s->hqueue.empty();
        if (s->stop && !queued && !busy) break;
        if (!busy && (!queued || s->dead)) {
            if (s->stop) break;
            s->cv.wait(lk);
        } else {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

// generate, admit, reorder and register uploaded R-MAT matrices
void upload_loop(mr_server s)
{
    hipSetDevice(s->ctx->device);
    host_thread_cap() = s->cfg.upload_threads;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->up_cv.wait(lk, [s] { return s->stop || !s->up_tasks.empty(); });
        if (s->up_tasks.empty()) break;
        const UploadTask task = s->up_tasks.front();
        s->up_tasks.pop_front();
        const std::string plan_dir = s->plan_dir;
        lk.unlock();
        const auto t0 = Clock::now();
        UploadRes res;
        int rc = MR_OK;
        const int64_t n = 1LL << task.scale;
        const int64_t m_edges = (int64_t)task.ef << task.scale;
        std::vector<int32_t> src((size_t)m_edges), dst((size_t)m_edges), rp((size_t)n + 1), ci((size_t)m_edges);
        int64_t nnz = 0;
        rc = mr_gen_rmat_edges(task.scale, task.ef, 0.57, 0.19, 0.19, task.seed, 1, src.data(), dst.data());
        if (rc == MR_OK) rc = mr_coo_to_csr(n, m_edges, src.data(), dst.data(), 1, rp.data(), ci.data(), &nnz);
        std::vector<double> val;
        if (rc == MR_OK) {
            val.resize((size_t)std::max<int64_t>(nnz, 1));
            rc = mr_gen_values(task.seed ^ 0x5EEDull, 0, -1.0, 1.0, nnz, val.data());
        }
        src.clear(); src.shrink_to_fit();
        dst.clear(); dst.shrink_to_fit();
        // resident bytes, and the peak while pGOrder runs (transpose, working
        // lists, relabel scratch)
        const int64_t dsz = (int64_t)dtype_size(s->dtype);
        const int64_t bytes = (n + 1) * 4 + (nnz + kPad) * (4 + dsz) + (n + nnz) / 8 + 4 * (n + nnz) / 256 + (1 << 20);
        const int64_t peak = bytes + 4 * (n + 1 + nnz) * 3 + nnz * (4 + dsz) + 16 * n;
        std::vector<mr_mat> victims;
        if (rc == MR_OK) {
            lk.lock();
            while (true) {
                size_t free_b = 0, total_b = 0;
                hipMemGetInfo(&free_b, &total_b);
                const bool fits_dev = (int64_t)free_b - s->cfg.hbm_reserve >= peak;
                const bool fits_budget = s->cfg.hbm_budget <= 0 || s->uploaded_bytes + bytes <= s->cfg.hbm_budget;
                if (fits_dev && fits_budget) break;
                // evict the least recently used uploaded matrix with no requests
                int v = -1;
                for (size_t i = 0; i < s->mats.size(); ++i) {
                    const SMat& M = s->mats[i];
                    if (M.owned && M.live && M.inflight == 0 &&
                        (v < 0 || M.last_use < s->mats[(size_t)v].last_use))
                        v = (int)i;
                }
                if (v < 0) {
                    rc = MR_ERR_ADMISSION;
                    break;
                }
                SMat& M = s->mats[(size_t)v];
                M.live = false;
                s->uploaded_bytes -= M.bytes;
                ++s->evictions;
                mr_mat dead = M.m;
                M.m = nullptr;
                lk.unlock();
                mr_mat_destroy(dead);
                lk.lock();
            }
            lk.unlock();
        }
        mr_mat mat = nullptr;
        if (rc == MR_OK)
            rc = mr_csr_create(s->ctx, n, n, nnz, rp.data(), ci.data(), val.data(), MR_F64, s->dtype, &mat);
        int plan_hit = 0;
        if (rc == MR_OK) {
            if (plan_dir.empty()) rc = mr_reorder_pgorder(mat, 0, 5, nullptr, nullptr);
            else rc = mr_reorder_pgorder_cached(mat, 0, 5, plan_dir.c_str(), nullptr, nullptr, &plan_hit);
        }
        if (rc == MR_OK) {
            const int32_t* bx;
            const int32_t* by;
            int64_t nb;
            rc = merbit_coords_sigma(mat, s->cfg.sigma, s->ctx->stream, &bx, &by, &nb);
            if (rc == MR_OK) rc = merbit_mm_grid(mat, s->ctx->stream);
            if (rc == MR_OK && (mat->merbit.policy & 4)) rc = merbit_lp_mm_build(mat, s->ctx->stream);
            if (rc == MR_OK && hipStreamSynchronize(s->ctx->stream) != hipSuccess) rc = MR_ERR_HIP;
        }
        if (rc != MR_OK && mat) {
            mr_mat_destroy(mat);
            mat = nullptr;
        }
        res.ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        lk.lock();
        if (rc == MR_OK) {
            SMat M;
            M.m = mat;
            M.owned = true;
            M.bytes = bytes;
            M.last_use = ++s->stamp;
            s->mats.push_back(M);
            s->uploaded_bytes += bytes;
            res.mat = (int)s->mats.size() - 1;
        }
        if (rc == MR_OK && !plan_dir.empty()) ++(plan_hit == 1 ? s->plan_hits : s->plan_misses);
        res.status = rc;
        s->up_res[(size_t)task.ticket] = res;
        s->done_cv.notify_all();
    }
}

int submit_locked(mr_server s, int mat, int k, uint64_t seed, double t_arr, int64_t* req)
{
    if (s->dead) return set_error(MR_ERR_HIP, "server stopped: %s", s->dead_msg.c_str());
    if (s->stop) return set_error(MR_ERR_CANCELLED, "server is shutting down");
    if (mat < 0 || mat >= (int)s->mats.size() || !s->mats[(size_t)mat].live)
        return set_error(MR_ERR_INVALID_ARG, "serve: matrix %d is not resident", mat);
    if (k < 1 || k > 64) return set_error(MR_ERR_UNSUPPORTED, "serve: k must be in [1, 64] (got %d)", k);
    Req R;
    R.mat = mat;
    R.k = k;
    R.seed = seed;
    R.t_arr = t_arr;
    const mr_mat m = s->mats[(size_t)mat].m;
    R.flops = 2.0 * (double)m->nnz * (double)k;
    R.bytes = req_bytes(m, k);
    R.heavy = R.bytes >= kHeavyBytes;
    const bool heavy = R.heavy;
    s->reqs.push_back(std::move(R));
    const int64_t id = (int64_t)s->reqs.size() - 1;
    (heavy ? s->hqueue : s->queue).push_back(id);
    ++s->mats[(size_t)mat].inflight;
    if (req) *req = id;
    s->cv.notify_one();
    return MR_OK;
}

void release_slots(mr_server s)
{
    for (Slot& sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.stream) hipStreamDestroy(sl.stream);
        if (sl.pinned) hipHostFree(sl.pinned);
        sl.done = nullptr;
        sl.stream = nullptr;
        sl.pinned = nullptr;
        sl.arena.release();
    }
}

uint64_t mix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit(uint64_t seed, uint64_t i) { return (double)(mix(seed, i) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

int mr_serve_config_default(mr_serve_config* c)
{
    if (!c) return set_error(MR_ERR_INVALID_ARG, "null config");
    std::memset(c, 0, sizeof(*c));
    c->slots = 4;
    c->max_jobs = 256;
    c->coalesce = 1;
    c->verify = 0;
    c->sigma = 8;
    c->grid = 0;
    c->upload_threads = 0;
    c->heavy_slots = 1;
    c->slot_bytes = 1LL << 30;
    c->hbm_budget = 0;
    c->hbm_reserve = 4LL << 30;
    c->timeout_ms = 0.0;
    return MR_OK;
}

int mr_serve_create(mr_ctx ctx, const mr_serve_config* cfg, mr_server* out)
{
    if (!ctx || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    mr_serve_config c;
    if (cfg) c = *cfg;
    else mr_serve_config_default(&c);
    if (c.slots < 1 || c.slots > 16) return set_error(MR_ERR_INVALID_ARG, "serve: slots must be in [1, 16]");
    if (c.max_jobs < 1 || c.max_jobs > k::kMaxServeReqs || c.max_jobs > k::kMaxBatchJobs)
        return set_error(MR_ERR_INVALID_ARG, "serve: max_jobs must be in [1, %d]", k::kMaxServeReqs);
    if (c.sigma != 8 && c.sigma != 16 && c.sigma != 32)
        return set_error(MR_ERR_INVALID_ARG, "serve: sigma must be 8, 16 or 32");
    if (c.verify < 0 || c.verify > 2) return set_error(MR_ERR_INVALID_ARG, "serve: verify must be 0, 1 or 2");
    if (c.grid < 0 || c.slot_bytes < 0 || c.hbm_budget < 0 || c.hbm_reserve < 0)
        return set_error(MR_ERR_INVALID_ARG, "serve: negative size");
    if (c.upload_threads <= 0) c.upload_threads = std::max(1, host_threads() / 2);
    if (c.heavy_slots < 0) return set_error(MR_ERR_INVALID_ARG, "serve: heavy_slots must be >= 0");
    if (c.heavy_slots == 0) c.heavy_slots = 1;
    MR_HIP_TRY(hipSetDevice(ctx->device));
    auto s = std::make_unique<mr_server_s>();
    s->ctx = ctx;
    s->cfg = c;
    s->epoch = Clock::now();
    // table region: per class a job table + begin array, the request table
    // and its two begin arrays, then the per-request hashes
    const size_t J = (size_t)c.max_jobs;
    s->table_bytes = kClasses * (align256(sizeof(k::BatchJob) * J) + align256(sizeof(int64_t) * (J + 1))) +
                     align256(sizeof(k::ServeReq) * J) + 2 * align256(sizeof(int64_t) * (J + 1));
    s->hash_off = s->table_bytes;
    s->table_bytes += align256(sizeof(uint64_t) * J);
    const size_t min_arena = s->table_bytes + (16u << 20);
    if ((size_t)c.slot_bytes < min_arena)
        return set_error(MR_ERR_INVALID_ARG, "serve: slot_bytes must be at least %zu", min_arena);
    s->slots.resize((size_t)c.slots);
    auto make_slot = [&](Slot& sl) -> int {
        MR_HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        MR_HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        MR_TRY(sl.arena.alloc((size_t)c.slot_bytes));
        MR_HIP_TRY(hipHostMalloc(&sl.pinned, s->table_bytes + sizeof(uint64_t) * J, hipHostMallocDefault));
        return MR_OK;
    };
    for (Slot& sl : s->slots) {
        const int rc = make_slot(sl);
        if (rc != MR_OK) {
            release_slots(s.get());
            return rc;
        }
    }
    s->dispatcher = std::thread(dispatch_loop, s.get());
    s->uploader = std::thread(upload_loop, s.get());
    *out = s.release();
    return MR_OK;
}

int mr_serve_destroy(mr_server s)
{
    if (!s) return MR_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->up_tasks.clear();
    }
    s->cv.notify_all();
    s->up_cv.notify_all();
    if (s->uploader.joinable()) s->uploader.join();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    hipSetDevice(s->ctx->device);
    release_slots(s);
    for (SMat& M : s->mats)
        if (M.owned && M.m) mr_mat_destroy(M.m);
    delete s;
    return MR_OK;
}

int mr_serve_add(mr_server s, mr_mat m, int* id)
{
    if (!s || !m) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (m->ctx != s->ctx) return set_error(MR_ERR_INVALID_ARG, "serve: matrix belongs to another context");
    if (m->dtype == MR_F64) return set_error(MR_ERR_UNSUPPORTED, "serve: F64 matrices are not served (FP32 x / y)");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->dtype >= 0 && m->dtype != s->dtype)
            return set_error(MR_ERR_UNSUPPORTED, "serve: all matrices share one storage dtype (%d)", s->dtype);
    }
    MR_HIP_TRY(hipSetDevice(s->ctx->device));
    // build the tile metadata now so the dispatcher never has to
    const int32_t* bx;
    const int32_t* by;
    int64_t nb;
    if (m->n_rows > 0) {
        MR_TRY(merbit_coords_sigma(m, s->cfg.sigma, s->ctx->stream, &bx, &by, &nb));
        MR_TRY(merbit_mm_grid(m, s->ctx->stream));
        // a column-panel plan serves single vectors; its SpMM runs on the
        // lane-padded tiles, as mr_spmm does
        if ((m->merbit.policy & 8) && m->dtype != MR_F64) MR_TRY(merbit_panel_build(m, s->ctx->stream));
        if (m->merbit.policy & 12) MR_TRY(merbit_lp_mm_build(m, s->ctx->stream));
        MR_HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    }
    std::lock_guard<std::mutex> g(s->mu);
    // checked again: another add may have fixed the dtype while the metadata
    // was built without the lock (the fused kernels read every matrix as it)
    if (s->dtype >= 0 && m->dtype != s->dtype)
        return set_error(MR_ERR_UNSUPPORTED, "serve: all matrices share one storage dtype (%d)", s->dtype);
    s->dtype = m->dtype;
    SMat M;
    M.m = m;
    M.last_use = ++s->stamp;
    s->mats.push_back(M);
    if (id) *id = (int)s->mats.size() - 1;
    return MR_OK;
}

int mr_serve_upload_rmat(mr_server s, int scale, int edge_factor, uint64_t seed, int* ticket)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    if (scale < 1 || scale > 26 || edge_factor < 1 || ((int64_t)edge_factor << scale) > INT32_MAX)
        return set_error(MR_ERR_INVALID_ARG, "serve: upload size out of range");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->dtype < 0) return set_error(MR_ERR_INVALID_ARG, "serve: add a matrix first (it fixes the dtype)");
    if (s->stop) return set_error(MR_ERR_CANCELLED, "server is shutting down");
    UploadTask t;
    t.scale = scale;
    t.ef = edge_factor;
    t.seed = seed;
    t.ticket = (int)s->up_res.size();
    s->up_res.emplace_back();
    s->up_tasks.push_back(t);
    if (ticket) *ticket = t.ticket;
    s->up_cv.notify_one();
    return MR_OK;
}

int mr_serve_upload_wait(mr_server s, int ticket, int* status, int* mat_id, double* ms)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket < 0 || ticket >= (int)s->up_res.size()) return set_error(MR_ERR_INVALID_ARG, "unknown upload ticket");
    s->done_cv.wait(lk, [&] { return s->up_res[(size_t)ticket].status >= 0 || s->stop; });
    const UploadRes& r = s->up_res[(size_t)ticket];
    if (status) *status = r.status < 0 ? MR_ERR_CANCELLED : r.status;
    if (mat_id) *mat_id = r.mat;
    if (ms) *ms = r.ms;
    return MR_OK;
}

int mr_serve_set_plan_cache(mr_server s, const char* dir)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    s->plan_dir = dir ? dir : "";
    return MR_OK;
}

int mr_serve_plan_cache_stats(mr_server s, int64_t* hits, int64_t* misses)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    if (hits) *hits = s->plan_hits;
    if (misses) *misses = s->plan_misses;
    return MR_OK;
}

int mr_serve_submit(mr_server s, int mat_id, int k, uint64_t seed, int64_t* req)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    return submit_locked(s, mat_id, k, seed, s->now_us(), req);
}

int mr_serve_cancel(mr_server s, int64_t req)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    if (req < 0 || req >= (int64_t)s->reqs.size()) return set_error(MR_ERR_INVALID_ARG, "unknown request");
    std::deque<int64_t>& q = s->reqs[(size_t)req].heavy ? s->hqueue : s->queue;
    auto it = std::find(q.begin(), q.end(), req);
    if (it == q.end()) return set_error(MR_ERR_INVALID_ARG, "request %lld is not queued", (long long)req);
    q.erase(it);
    finish_req(s, req, MR_ERR_CANCELLED, s->now_us());
    s->done_cv.notify_all();
    return MR_OK;
}

int mr_serve_wait(mr_server s, int64_t req, int* status, double* latency_us, uint64_t* hash)
{
    if (!s) return set_error(MR_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (req < 0 || req >= (int64_t)s->reqs.size()) return set_error(MR_ERR_INVALID_ARG, "unknown request");
    s->done_cv.wait(lk, [&] { return s->reqs[(size_t)req].status >= 0; });
    const Req& R = s->reqs[(size_t)req];
    if (status) *status = R.status;
    if (latency_us) *latency_us = R.t_done - R.t_arr;
    if (hash) *hash = R.hash;
    return MR_OK;
}

int mr_serve_result(mr_server s, int64_t req, float* y)
{
    if (!s || !y) return set_error(MR_ERR_INVALID_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->cfg.verify < 2) return set_error(MR_ERR_INVALID_ARG, "serve: results are kept only with verify = 2");
    if (req < 0 || req >= (int64_t)s->reqs.size()) return set_error(MR_ERR_INVALID_ARG, "unknown request");
    s->done_cv.wait(lk, [&] { return s->reqs[(size_t)req].status >= 0; });
    const Req& R = s->reqs[(size_t)req];
    if (R.status != MR_OK) return set_error(R.status, "request %lld failed", (long long)req);
    std::memcpy(y, R.y.data(), sizeof(float) * R.y.size());
    return MR_OK;
}

int mr_serve_run_load(mr_server s, const mr_serve_load* load, mr_serve_report* rep)
{
    if (!s || !load || !rep) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (!(load->rate > 0) || load->n_requests < 1 || load->n_k < 1 || load->n_k > 4)
        return set_error(MR_ERR_INVALID_ARG, "serve: bad load (rate > 0, n_requests >= 1, 1 <= n_k <= 4)");
    for (int i = 0; i < load->n_k; ++i)
        if (load->k_choices[i] < 1 || load->k_choices[i] > 64)
            return set_error(MR_ERR_UNSUPPORTED, "serve: k must be in [1, 64]");
    std::memset(rep, 0, sizeof(*rep));
    const int64_t N = load->n_requests;
    std::vector<double> arr((size_t)N);
    double t = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        t += -std::log1p(-unit(load->seed, (uint64_t)(3 * i))) / load->rate * 1e6;
        arr[(size_t)i] = t;
    }
    std::vector<int64_t> ids;
    ids.reserve((size_t)N);
    std::vector<int> tickets;
    int64_t b0, j0, r0;
    double t0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->mats.empty()) return set_error(MR_ERR_INVALID_ARG, "serve: no resident matrix");
        b0 = s->batches;
        j0 = s->batch_jobs;
        r0 = s->batch_reqs;
        t0 = s->now_us() + 1000.0;
    }
    const int64_t ev0 = s->evictions;
    for (int64_t i = 0; i < N; ++i) {
        const double target = t0 + arr[(size_t)i];
        for (;;) {
            const double now = s->now_us();
            if (now >= target) break;
            if (target - now > 300.0)
                std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(target - now - 200.0)));
        }
        if (load->upload_every > 0 && (i + 1) % load->upload_every == 0) {
            int tk = -1;
            if (mr_serve_upload_rmat(s, load->upload_scale, load->upload_edge_factor, mix(load->seed, (uint64_t)i),
                                     &tk) == MR_OK)
                tickets.push_back(tk);
            continue;
        }
        const int k = load->k_choices[(size_t)(mix(load->seed, (uint64_t)(3 * i + 1)) % (uint64_t)load->n_k)];
        std::lock_guard<std::mutex> g(s->mu);
        // a matrix chosen uniformly among the resident ones
        std::vector<int> live;
        live.reserve(s->mats.size());
        for (size_t m = 0; m < s->mats.size(); ++m)
            if (s->mats[m].live) live.push_back((int)m);
        if (live.empty()) continue;
        const int mat = live[(size_t)(mix(load->seed, (uint64_t)(3 * i + 2)) % (uint64_t)live.size())];
        int64_t id = -1;
        if (submit_locked(s, mat, k, mix(load->seed ^ 0xABCDull, (uint64_t)i), target, &id) == MR_OK)
            ids.push_back(id);
    }
    std::unique_lock<std::mutex> lk(s->mu);
    for (int64_t id : ids) s->done_cv.wait(lk, [&] { return s->reqs[(size_t)id].status >= 0; });
    double t_end = t0;
    std::vector<double> lat, lat_light, lat_heavy;
    lat.reserve(ids.size());
    double flops = 0.0, bytes = 0.0;
    for (int64_t id : ids) {
        const Req& R = s->reqs[(size_t)id];
        if (R.status == MR_OK) {
            lat.push_back(R.t_done - R.t_arr);
            (R.heavy ? lat_heavy : lat_light).push_back(R.t_done - R.t_arr);
            if (R.heavy) ++rep->completed_heavy;
            flops += R.flops;
            bytes += R.bytes;
            t_end = std::max(t_end, R.t_done);
            ++rep->completed;
        } else if (R.status == MR_ERR_TIMEOUT) {
            ++rep->timed_out;
            if (R.heavy) ++rep->timed_out_heavy;
        } else if (R.status == MR_ERR_CANCELLED) {
            ++rep->cancelled;
        } else {
            ++rep->rejected;
        }
    }
    rep->batches = (double)(s->batches - b0);
    const double nb = std::max(rep->batches, 1.0);
    rep->mean_jobs_per_batch = (double)(s->batch_jobs - j0) / nb;
    rep->mean_requests_per_batch = (double)(s->batch_reqs - r0) / nb;
    lk.unlock();
    double up_ms = 0.0;
    for (int tk : tickets) {
        int st = 0;
        double ms = 0.0;
        mr_serve_upload_wait(s, tk, &st, nullptr, &ms);
        if (st == MR_OK) {
            ++rep->uploads_admitted;
            up_ms += ms;
        } else {
            ++rep->uploads_rejected;
        }
    }
    rep->upload_ms_mean = rep->uploads_admitted ? up_ms / (double)rep->uploads_admitted : 0.0;
    rep->evictions = s->evictions - ev0;
    std::sort(lat.begin(), lat.end());
    std::sort(lat_light.begin(), lat_light.end());
    std::sort(lat_heavy.begin(), lat_heavy.end());
    auto pct_of = [](const std::vector<double>& v, double q) {
        if (v.empty()) return 0.0;
        size_t i = (size_t)std::ceil(q * (double)v.size());
        if (i > 0) --i;
        return v[std::min(i, v.size() - 1)];
    };
    auto pct = [&](double q) { return pct_of(lat, q); };
    rep->p99_light_us = pct_of(lat_light, 0.99);
    rep->p99_heavy_us = pct_of(lat_heavy, 0.99);
    rep->p50_us = pct(0.50);
    rep->p90_us = pct(0.90);
    rep->p99_us = pct(0.99);
    rep->max_us = lat.empty() ? 0.0 : lat.back();
    double sum = 0.0;
    for (double v : lat) sum += v;
    rep->mean_us = lat.empty() ? 0.0 : sum / (double)lat.size();
    rep->wall_s = (t_end - t0) * 1e-6;
    rep->offered_rate = load->rate;
    rep->achieved_rate = rep->wall_s > 0 ? (double)rep->completed / rep->wall_s : 0.0;
    rep->gflops = rep->wall_s > 0 ? flops / rep->wall_s * 1e-9 : 0.0;
    rep->gbytes_s = rep->wall_s > 0 ? bytes / rep->wall_s * 1e-9 : 0.0;
    return MR_OK;
}

int mr_serve_rhs(uint64_t seed, int64_t n_cols, int k, float* out)
{
    if (!out || n_cols < 0 || k < 1) return set_error(MR_ERR_INVALID_ARG, "bad argument");
    const int64_t total = n_cols * (int64_t)k;
    for (int64_t j = 0; j < total; ++j) out[j] = k::serve_rhs_value(seed, (uint64_t)j);
    return MR_OK;
}

int mr_serve_hash(const float* y, int64_t n_rows, int k, uint64_t* hash)
{
    if ((!y && n_rows > 0) || !hash || n_rows < 0 || k < 1) return set_error(MR_ERR_INVALID_ARG, "bad argument");
    uint64_t h = 0;
    const int64_t total = n_rows * (int64_t)k;
    for (int64_t j = 0; j < total; ++j) {
        uint32_t bits;
        std::memcpy(&bits, &y[j], 4);
        h += k::serve_hash_term((uint64_t)j, bits);
    }
    *hash = h;
    return MR_OK;
}

}  // extern "C"
