// reorder.cpp -- reordering entry points of libmr: the GPU relabel
// (P A P^T), pGOrder (GPU greedy placement + per-batch CPU Gorder), the
// sequential CPU Gorder and Reverse Cuthill-McKee baselines.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <condition_variable>
#include <mutex>
#include <thread>

#include "bucket_heap.hpp"
#include "common.hpp"
#include "kernels/launch.hpp"
#include "kernels/launch_degree.hpp"
#include "kernels/launch_perm.hpp"
#include "kernels/launch_pg.hpp"
#include "kernels/launch_radix.hpp"
#include "kernels/launch_transpose.hpp"
#include "parallel.hpp"
#include "reorder.hpp"

namespace mr {

namespace {

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0)
{
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

// segmented row sort of a CSR-shaped structure on the device:
// new row p = old row inv[p], entries mapped through perm and sorted.
int sort_rows(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, const void* val, int val_bytes,
              const int32_t* perm, const int32_t* inv, const int32_t* new_rp, int32_t* new_ci, void* new_val,
              hipStream_t s)
{
    DevBuf d_len, d_counts, d_offs, d_cursor, d_lists;
    MR_TRY(d_len.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1)));
    MR_TRY(d_counts.alloc(sizeof(int32_t) * k::kPermClasses));
    MR_TRY(d_offs.alloc(sizeof(int32_t) * k::kPermClasses));
    MR_TRY(d_cursor.alloc(sizeof(int32_t) * k::kPermClasses));
    MR_TRY(d_lists.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1)));
    MR_HIP_TRY(k::perm_row_lengths(n, rp, inv, d_len.as<int32_t>(), s));
    MR_HIP_TRY(hipMemsetAsync(d_counts.p, 0, d_counts.bytes, s));
    MR_HIP_TRY(hipMemsetAsync(d_cursor.p, 0, d_cursor.bytes, s));
    MR_HIP_TRY(k::perm_class_count(n, d_len.as<int32_t>(), d_counts.as<int32_t>(), s));
    int32_t counts[k::kPermClasses], offs[k::kPermClasses];
    MR_HIP_TRY(hipMemcpyAsync(counts, d_counts.p, sizeof(counts), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    int32_t run = 0;
    for (int c = 0; c < k::kPermClasses; ++c) { offs[c] = run; run += counts[c]; }
    MR_HIP_TRY(hipMemcpyAsync(d_offs.p, offs, sizeof(offs), hipMemcpyHostToDevice, s));
    MR_HIP_TRY(k::perm_class_fill(n, d_len.as<int32_t>(), d_offs.as<int32_t>(), d_cursor.as<int32_t>(),
                                  d_lists.as<int32_t>(), s));
    // the longest class: lengths of its rows only, their offsets, one sort
    const int32_t nbig = counts[k::kPermClasses - 1];
    DevBuf d_boff, d_lws;
    std::vector<int64_t> boff;   // kept until the end: its upload is asynchronous
    int64_t big_total = 0;
    const int32_t* big_rows = d_lists.as<int32_t>() + offs[k::kPermClasses - 1];
    if (nbig > 0) {
        DevBuf d_blen;
        MR_TRY(d_blen.alloc(sizeof(int32_t) * (size_t)nbig));
        MR_HIP_TRY(k::perm_long_lengths(nbig, big_rows, d_len.as<int32_t>(), d_blen.as<int32_t>(), s));
        std::vector<int32_t> blen((size_t)nbig);
        MR_HIP_TRY(hipMemcpyAsync(blen.data(), d_blen.p, sizeof(int32_t) * (size_t)nbig, hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipStreamSynchronize(s));
        boff.resize((size_t)nbig);
        for (int32_t i = 0; i < nbig; ++i) {
            boff[(size_t)i] = big_total;
            big_total += blen[(size_t)i];
        }
        MR_TRY(d_boff.alloc(sizeof(int64_t) * (size_t)nbig));
        MR_HIP_TRY(hipMemcpyAsync(d_boff.p, boff.data(), sizeof(int64_t) * (size_t)nbig, hipMemcpyHostToDevice, s));
        MR_TRY(d_lws.alloc(k::perm_long_ws_bytes(big_total, nbig, n)));
    }
    for (int c = 0; c < k::kPermClasses - 1; ++c)
        MR_HIP_TRY(k::perm_sort_class(c, counts[c], d_lists.as<int32_t>() + offs[c], rp, ci, val, val_bytes, perm,
                                      inv, new_rp, new_ci, new_val, s));
    if (nbig > 0)
        MR_HIP_TRY(k::perm_sort_long(nbig, big_rows, rp, ci, val, val_bytes, perm, inv, new_rp, new_ci, new_val,
                                     d_boff.as<int64_t>(), big_total, n, d_lws.p, d_lws.bytes, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    (void)nnz;
    return MR_OK;
}

}  // namespace

int gpu_transpose(mr_mat m, DevBuf& tp, DevBuf& ti, hipStream_t s)
{
    const int64_t n = m->n_rows, nc = m->n_cols, nnz = m->nnz;
    MR_TRY(tp.alloc(sizeof(int32_t) * (size_t)(nc + 1)));
    MR_TRY(ti.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    DevBuf ws;
    MR_TRY(ws.alloc(k::transpose_ws_bytes(nnz, nc)));
    MR_HIP_TRY(k::transpose_sorted(n, nc, nnz, m->rp.as<int32_t>(), m->ci.as<int32_t>(), tp.as<int32_t>(),
                                   ti.as<int32_t>(), ws.p, ws.bytes, s));
    return MR_OK;
}

int relabel(mr_mat m, const std::vector<int32_t>& perm, double* ms_out)
{
    const auto t0 = clk::now();
    const int64_t n = m->n_rows, nnz = m->nnz;
    hipStream_t s = m->ctx->stream;
    if (n == 0) return MR_OK;
    DevBuf d_perm, d_inv, d_len, d_newrp, d_newci, d_newval, ws;
    MR_TRY(d_perm.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(d_inv.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(d_len.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(d_newrp.alloc(sizeof(int32_t) * (size_t)(n + 1)));
    MR_TRY(d_newci.alloc(sizeof(int32_t) * (size_t)(nnz + kPad)));
    MR_TRY(d_newval.alloc(dtype_size(m->dtype) * (size_t)(nnz + kPad)));
    MR_TRY(ws.alloc(k::scan_workspace_bytes(n)));
    MR_HIP_TRY(hipMemcpyAsync(d_perm.p, perm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, s));
    {
        DevBuf d_bad;
        MR_TRY(d_bad.alloc(sizeof(int32_t)));
        MR_HIP_TRY(k::invert_perm(n, d_perm.as<int32_t>(), d_inv.as<int32_t>(), d_bad.as<int32_t>(), s));
        int32_t bad = 0;
        MR_HIP_TRY(hipMemcpyAsync(&bad, d_bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipStreamSynchronize(s));
        if (bad) return set_error(MR_ERR_INVALID_ARG, "relabel: the order is not a permutation of [0, %lld)",
                                  (long long)n);
    }
    // only the padding past nnz needs a value: every entry below it is written
    MR_HIP_TRY(hipMemsetAsync(d_newci.as<int32_t>() + nnz, 0, sizeof(int32_t) * (size_t)kPad, s));
    MR_HIP_TRY(hipMemsetAsync(static_cast<char*>(d_newval.p) + dtype_size(m->dtype) * (size_t)nnz, 0,
                              dtype_size(m->dtype) * (size_t)kPad, s));
    MR_HIP_TRY(k::perm_row_lengths(n, m->rp.as<int32_t>(), d_inv.as<int32_t>(), d_len.as<int32_t>(), s));
    MR_HIP_TRY(k::exclusive_scan_i32(d_len.as<int32_t>(), n, d_newrp.as<int32_t>(), nullptr, ws.p, s));
    const int32_t tail = (int32_t)nnz;
    MR_HIP_TRY(hipMemcpyAsync(d_newrp.as<int32_t>() + n, &tail, sizeof(int32_t), hipMemcpyHostToDevice, s));
    MR_TRY(sort_rows(n, nnz, m->rp.as<int32_t>(), m->ci.as<int32_t>(), m->val.p, (int)dtype_size(m->dtype),
                     d_perm.as<int32_t>(), d_inv.as<int32_t>(), d_newrp.as<int32_t>(), d_newci.as<int32_t>(),
                     d_newval.p, s));
    m->rp = std::move(d_newrp);
    m->ci = std::move(d_newci);
    m->val = std::move(d_newval);
    for (auto& p : m->perm) p = perm[(size_t)p];
    m->csr_binned = false;   // row lengths moved: the CSR row bins are rebuilt on next use
    m->hot_valid = false;   // column ids moved: the hot columns are recounted on next use
    m->merbit.pn_built = false;   // and the column panels are rebuilt
    if (ms_out) *ms_out = ms_since(t0);
    return merbit_build(m, s);
}

// ---------------------------------------------------------------------------
// pGOrder batch schedule (src/class.cu:2995-3016)
int32_t pg_schedule(int32_t n, int32_t W, std::vector<int32_t>& Ps, std::vector<int32_t>& Pl,
                    std::vector<int32_t>& Os, std::vector<int32_t>& Ol)
{
    const int base = 5;
    const int32_t kk = (int32_t)std::log2((double)W / std::pow(2.0, base));
    const int32_t B = (n + W - 1) / W + kk;
    Ps.assign((size_t)B, 0); Pl.assign((size_t)B, 0); Os.assign((size_t)B, 0); Ol.assign((size_t)B, 0);
    for (int32_t b = 0; b < B; ++b) {
        if (b == 0) { Ps[0] = 0; Pl[0] = 32; Os[0] = 0; Ol[0] = 32; continue; }
        const double p1 = std::pow(2.0, b + base - 1), p2 = std::pow(2.0, b + base);
        int32_t start = (p1 < W) ? (int32_t)p1 : (b - kk) * W;
        int32_t leng = (p1 < W) ? (int32_t)p1 : W;
        Ps[(size_t)b] = start;
        Pl[(size_t)b] = (start + leng < n) ? leng : n - start;
        start = (p1 < W) ? 0 : (b - kk) * W;
        leng = (p2 < W) ? (int32_t)p2 : W;
        Os[(size_t)b] = start;
        Ol[(size_t)b] = (start + leng < n) ? leng : n - start;
    }
    return B;
}

// ---------------------------------------------------------------------------
// order_on_cpu (src/class.cu:2220-2314) for one batch.  In-batch siblings of a
// vertex are reached through y -> z lists restricted to the batch, built in
// ascending z order, which issues exactly the reference's sequence of heap
// updates (it walks all of rows(y) in ascending order and skips vertices of
// other batches).
struct PgCpuShared {
    const int32_t *rp, *ci, *tp, *ti;
    const int32_t* local_idx;  // position of a vertex inside its batch slice
    int32_t window, degree;
};

struct PgCpuWorker {
    BucketHeap heap;
    std::vector<int32_t> yslot;      // per vertex, -1 when not a y of this batch
    std::vector<int32_t> ys, ycnt, ystart, zlist, sorted, out;
    std::vector<uint8_t> sts;
};

void order_batch(const PgCpuShared& g, PgCpuWorker& w, uint32_t* slice, int32_t len)
{
    if (len <= 0) return;
    const int32_t* rp = g.rp;
    const int32_t* ci = g.ci;
    const int32_t* tp = g.tp;
    const int32_t L = g.degree;
    auto collen = [&](int32_t v) { return tp[v + 1] - tp[v]; };
    auto rowlen = [&](int32_t v) { return rp[v + 1] - rp[v]; };
    // start: first vertex with strictly maximal out-list length
    int32_t start = 0, deg = collen((int32_t)slice[0]);
    for (int32_t i = 0; i < len; ++i)
        if (collen((int32_t)slice[i]) > deg) { deg = collen((int32_t)slice[i]); start = i; }
    // y -> in-batch z lists, z ascending
    w.sorted.assign(slice, slice + len);
    std::sort(w.sorted.begin(), w.sorted.end());
    w.ys.clear();
    w.ycnt.clear();
    for (int32_t z : w.sorted)
        for (int32_t a = rp[z]; a < rp[z + 1]; ++a) {
            const int32_t y = ci[a];
            if (collen(y) >= L) continue;
            int32_t& s = w.yslot[(size_t)y];
            if (s < 0) { s = (int32_t)w.ys.size(); w.ys.push_back(y); w.ycnt.push_back(0); }
            w.ycnt[(size_t)s]++;
        }
    w.ystart.assign(w.ys.size() + 1, 0);
    for (size_t i = 0; i < w.ys.size(); ++i) w.ystart[i + 1] = w.ystart[i] + w.ycnt[i];
    w.zlist.resize((size_t)w.ystart.back());
    std::fill(w.ycnt.begin(), w.ycnt.end(), 0);
    for (int32_t z : w.sorted) {
        const int32_t lz = g.local_idx[z];
        for (int32_t a = rp[z]; a < rp[z + 1]; ++a) {
            const int32_t y = ci[a];
            if (collen(y) >= L) continue;
            const int32_t s = w.yslot[(size_t)y];
            w.zlist[(size_t)(w.ystart[(size_t)s] + w.ycnt[(size_t)s]++)] = lz;
        }
    }
    w.heap.reset(len);
    for (int32_t i = 0; i < len; ++i) w.heap.push(i);
    w.heap.modify(start, 1);
    w.sts.assign((size_t)len, 1);
    w.out.resize((size_t)len);
    auto apply = [&](int32_t x, int32_t d) {
        if (rowlen(x) >= L) return;
        for (int32_t a = rp[x]; a < rp[x + 1]; ++a) {
            const int32_t y = ci[a];
            if (collen(y) >= L) continue;
            const int32_t s = w.yslot[(size_t)y];
            if (s < 0) continue;
            for (int32_t q = w.ystart[(size_t)s]; q < w.ystart[(size_t)s + 1]; ++q) {
                const int32_t z = w.zlist[(size_t)q];
                if (w.sts[(size_t)z]) w.heap.modify(z, d);
            }
        }
    };
    for (int32_t id = 0; id < len; ++id) {
        const int32_t li = w.heap.pop();
        w.sts[(size_t)li] = 0;
        const int32_t node = (int32_t)slice[li];
        w.out[(size_t)id] = node;
        apply(node, 1);
        if (id >= g.window) apply(w.out[(size_t)(id - g.window)], -1);
    }
    for (int32_t i = 0; i < len; ++i) slice[i] = (uint32_t)w.out[(size_t)i];
    for (int32_t y : w.ys) w.yslot[(size_t)y] = -1;
}

}  // namespace mr

using namespace mr;

extern "C" {

int mr_get_width(int64_t n, int64_t nnz, int* width)
{
    if (!width || n <= 0 || nnz < 0) return set_error(MR_ERR_INVALID_ARG, "bad arguments");
    // GRAPH::get_width (src/class.cu:733-750): float arithmetic, integer nnz/n
    const int32_t nrow = (int32_t)n, nz = (int32_t)nnz;
    const float q = (float)nrow / (float)(nz / nrow);
    const float bound = std::round(std::log2(std::sqrt(q)));
    *width = bound < 10.0f ? 1024 : (bound < 12.0f ? 16384 : 262144);
    return MR_OK;
}

int mr_permute(mr_mat m, const int32_t* perm)
{
    if (!m || !perm) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "symmetric relabel needs a square matrix");
    if (m->n_rows == 0) return MR_OK;
    MR_TRY(mr_check_perm(perm, m->n_rows));
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    std::vector<int32_t> p(perm, perm + m->n_rows);
    return relabel(m, p, nullptr);
}

// The Gorder-family placement shared by gorder6 (pGOrder) and gorder5: the
// GPU loop over the batch schedule (Ps, Pl) with windows (Os, Ol), then
// order_on_cpu on the host for every slice (Cs, Cl).
struct GorderPlan {
    std::vector<int32_t> Ps, Pl, Os, Ol;   // GPU batches and their windows
    std::vector<int32_t> Cs, Cl;           // slices of the CPU phase
    bool start_columns = false;            // start scores: column counts (else row lengths)
    int32_t hub_cap = 10000;               // window rows / columns at or above it are skipped
    int32_t cpu_window = 5;
};

static int gorder_family_order(mr_mat m, const GorderPlan& P, uint32_t* gpu_order, uint32_t* final_order,
                               mr_reorder_stats* stats)
{
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const auto t_all = clk::now();
    const int64_t nnz = m->nnz;
    const int32_t n = (int32_t)m->n_rows;
    const int32_t B = (int32_t)P.Ps.size();
    hipStream_t s = m->ctx->stream;

    // ---- GPU loop ----
    const auto t_gpu = clk::now();
    DevBuf tp, ti;
    MR_TRY(gpu_transpose(m, tp, ti, s));
    DevBuf score, T0, T1, sure, tie, hist, chunk, placement, state, wprefix, hcol, items;
    MR_TRY(score.alloc(sizeof(uint32_t) * (size_t)n));
    MR_TRY(T0.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(T1.alloc(sizeof(int32_t) * (size_t)n));
    // batches with |P| or a window above 1024 take the wide selection
    const int32_t max_pl = B > 0 ? *std::max_element(P.Pl.begin(), P.Pl.end()) : 0;
    const int32_t max_ol = B > 0 ? *std::max_element(P.Ol.begin(), P.Ol.end()) : 0;
    const bool any_wide = max_pl > k::kPgNarrow || max_ol > k::kPgNarrow;
    MR_TRY(sure.alloc(sizeof(int32_t) * (size_t)std::max<int32_t>(4096, max_pl)));
    MR_TRY(tie.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(hist.alloc(sizeof(int32_t) * (size_t)k::pg_hist_bins()));
    MR_TRY(chunk.alloc(sizeof(int32_t) * (size_t)((n + 1023) / 1024 + 1)));
    MR_TRY(placement.alloc(sizeof(uint32_t) * (size_t)n));
    MR_TRY(state.alloc(sizeof(k::PgState)));
    MR_TRY(wprefix.alloc(sizeof(int32_t) * (size_t)(std::max<int32_t>(1024, max_ol) + 1)));
    DevBuf wka, wva, wkb, wvb, wrws, wcoff, wlens, wsws, whist2;
    k::PgWide wide{};
    if (any_wide) {
        const int32_t cap = std::max<int32_t>(1, max_pl);
        MR_TRY(wka.alloc(sizeof(uint32_t) * (size_t)cap));
        MR_TRY(wva.alloc(sizeof(uint32_t) * (size_t)cap));
        MR_TRY(wkb.alloc(sizeof(uint32_t) * (size_t)cap));
        MR_TRY(wvb.alloc(sizeof(uint32_t) * (size_t)cap));
        MR_TRY(wrws.alloc(k::radix_sort_ws_bytes(cap)));
        MR_TRY(wcoff.alloc(sizeof(int32_t) * (size_t)((n + 1023) / 1024 + 1)));
        MR_TRY(wlens.alloc(sizeof(int32_t) * (size_t)(max_ol + 1)));
        MR_TRY(wsws.alloc(k::scan_workspace_bytes((int64_t)max_ol + 1)));
        MR_TRY(whist2.alloc(sizeof(int32_t) * 2048));
        MR_HIP_TRY(hipMemsetAsync(whist2.p, 0, whist2.bytes, s));
        wide.cap = cap;
        wide.ka = wka.as<uint32_t>();
        wide.va = wva.as<uint32_t>();
        wide.kb = wkb.as<uint32_t>();
        wide.vb = wvb.as<uint32_t>();
        wide.rws = wrws.p;
        wide.rws_bytes = wrws.bytes;
        wide.coff = wcoff.as<int32_t>();
        wide.lens = wlens.as<int32_t>();
        wide.sws = wsws.p;
        wide.hist2 = whist2.as<int32_t>();
    }
    MR_TRY(hcol.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(items.alloc(sizeof(uint64_t) * (size_t)k::pg_item_capacity(n, nnz)));
    MR_HIP_TRY(hipMemsetAsync(hcol.p, 0, hcol.bytes, s));
    MR_HIP_TRY(hipMemsetAsync(hist.p, 0, hist.bytes, s));
    MR_HIP_TRY(hipMemsetAsync(state.p, 0, state.bytes, s));
    MR_HIP_TRY(hipMemsetAsync(placement.p, 0xff, placement.bytes, s));
    k::PgState* st = state.as<k::PgState>();
    // working copy of the column lists: placed vertices are dropped from it
    // every `compact_every` batches (the host phase needs the full lists)
    DevBuf ti_live, live;
    MR_TRY(ti_live.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    MR_TRY(live.alloc(sizeof(int32_t) * (size_t)n));
    if (nnz > 0)
        MR_HIP_TRY(hipMemcpyAsync(ti_live.p, ti.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToDevice, s));
    MR_HIP_TRY(k::pg_live_init(n, tp.as<int32_t>(), live.as<int32_t>(), s));
    const int32_t compact_every = std::max<int32_t>(32, B / 32);
    MR_HIP_TRY(k::pg_init(n, P.start_columns ? tp.as<int32_t>() : m->rp.as<int32_t>(), score.as<uint32_t>(),
                          T0.as<int32_t>(), st, chunk.as<int32_t>(), hist.as<int32_t>(),
                          B > 0 && (P.Pl[0] > k::kPgNarrow || P.Ol[0] > k::kPgNarrow), s));
    int32_t* Tl[2] = {T0.as<int32_t>(), T1.as<int32_t>()};

    // ---- host copies of both adjacency lists for the CPU phase ----
    // Pinned, so each is one DMA transfer queued on the stream ahead of the
    // loop (the matrix does not change).  Every placement event the refining
    // threads wait for comes later in the same stream, so no other wait is
    // needed before they read the lists.
    struct PinnedLists {
        int32_t* p = nullptr;
        ~PinnedLists() { if (p) hipHostFree(p); }
    } lists;
    const size_t nnz_a = (size_t)std::max<int64_t>(nnz, 1);
    MR_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&lists.p),
                             sizeof(int32_t) * (2 * ((size_t)n + 1) + 2 * nnz_a), hipHostMallocDefault));
    int32_t* const hrp = lists.p;
    int32_t* const htp = hrp + ((size_t)n + 1);
    int32_t* const hci = htp + ((size_t)n + 1);
    int32_t* const hti = hci + nnz_a;
    MR_HIP_TRY(hipMemcpyAsync(hrp, m->rp.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(htp, tp.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    if (nnz > 0) {
        MR_HIP_TRY(hipMemcpyAsync(hci, m->ci.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipMemcpyAsync(hti, ti.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s));
    }

    // ---- CPU phase overlapped with the GPU loop ----
    // A slice is final once the GPU has placed every batch that writes into
    // it; the loop copies the placement to pinned host memory every kChunk
    // batches and records an event, and host threads refine each slice
    // (order_on_cpu) as soon as the placement covering it has landed.
    struct Pinned {
        uint32_t* p = nullptr;
        ~Pinned() { if (p) hipHostFree(p); }
    } ghost;
    MR_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&ghost.p), sizeof(uint32_t) * (size_t)n, hipHostMallocDefault));
    std::vector<uint32_t> order((size_t)n);
    const int32_t C = (int32_t)P.Cs.size();
    std::vector<int32_t> ready((size_t)C, 0);   // batch after which slice c is final
    for (int32_t c = 0, b2 = 0; c < C; ++c) {
        const int64_t last = (int64_t)P.Cs[(size_t)c] + P.Cl[(size_t)c] - 1;
        while (b2 + 1 < B && (int64_t)P.Ps[(size_t)b2] + P.Pl[(size_t)b2] - 1 < last) ++b2;
        ready[(size_t)c] = b2;
    }
    std::vector<int32_t> local_idx((size_t)n, -1);
    PgCpuShared g{hrp, hci, htp, hti, local_idx.data(), P.cpu_window,
                  (int32_t)(std::sqrt((double)n) / 4)};
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::pair<int32_t, hipEvent_t>> published;   // (last batch copied, event)
    bool abort_all = false;
    int32_t done_batch = -1;   // the placement up to this batch is in ghost
    std::atomic<int32_t> next{0};
    std::atomic<int64_t> cpu_busy_us{0};
    const int device = m->ctx->device;
    // One thread waits for the copy events in order and publishes how far the
    // placement has landed; the refining threads make no HIP call (thousands
    // of event waits from every thread contend in the runtime with the thread
    // that launches the loop).
    auto watcher = [&] {
        hipSetDevice(device);
        size_t next_ev = 0;
        for (;;) {
            std::pair<int32_t, hipEvent_t> pe;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return abort_all || next_ev < published.size(); });
                if (abort_all) return;
                pe = published[next_ev++];
            }
            const bool ok = hipEventSynchronize(pe.second) == hipSuccess;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (ok) done_batch = pe.first;
                else abort_all = true;
            }
            cv.notify_all();
            if (!ok || pe.first >= B - 1) return;
        }
    };
    auto worker = [&] {
        PgCpuWorker w;
        w.yslot.assign((size_t)n, -1);
        for (int32_t c = next++; c < C; c = next++) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return abort_all || done_batch >= ready[(size_t)c]; });
                if (abort_all) return;
            }
            const auto t0 = clk::now();
            const int32_t cs = P.Cs[(size_t)c], cl = P.Cl[(size_t)c];
            for (int32_t i = 0; i < cl; ++i) {
                order[(size_t)(cs + i)] = ghost.p[cs + i];
                local_idx[(size_t)ghost.p[cs + i]] = i;
            }
            order_batch(g, w, order.data() + cs, cl);
            cpu_busy_us += (int64_t)(ms_since(t0) * 1000.0);
        }
    };
    // one core stays with the thread that launches the GPU loop
    const int nth = C > 0 ? std::max(1, std::min<int>(host_threads() - 1, C)) : 0;
    std::vector<std::thread> pool;
    if (nth > 0) pool.emplace_back(watcher);
    for (int t = 0; t < nth; ++t) pool.emplace_back(worker);
    auto finish = [&](bool ok) {
        if (!ok) {
            std::lock_guard<std::mutex> lk(mu);
            abort_all = true;
        }
        cv.notify_all();
        for (auto& th : pool) th.join();
        pool.clear();
        for (auto& pe : published) hipEventDestroy(pe.second);
        published.clear();
    };
    constexpr int32_t kChunk = 32;
    int64_t copied = 0;
    auto publish = [&](int32_t b) -> int {
        // batch 0 of the gorder6 schedule keeps |P| = 32 even when n < 32
        const int64_t upto = std::min<int64_t>((int64_t)P.Ps[(size_t)b] + P.Pl[(size_t)b], n);
        if (upto > copied) {
            const hipError_t e = hipMemcpyAsync(ghost.p + copied, placement.as<uint32_t>() + copied,
                                                sizeof(uint32_t) * (size_t)(upto - copied), hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) return set_error(MR_ERR_HIP, "placement copy failed: %s", hipGetErrorString(e));
            copied = upto;
        }
        hipEvent_t ev;
        // blocking sync: waiting host threads sleep instead of spinning beside
        // the thread that launches the loop
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming | hipEventBlockingSync) != hipSuccess ||
            hipEventRecord(ev, s) != hipSuccess)
            return set_error(MR_ERR_HIP, "event for the placement copy failed");
        {
            std::lock_guard<std::mutex> lk(mu);
            published.emplace_back(b, ev);
        }
        cv.notify_all();
        return MR_OK;
    };

    int cur = 0;
    int rc = MR_OK;
    for (int32_t b = 0; b < B && rc == MR_OK; ++b) {
        hipError_t e = hipSuccess;
        if (b > 0 && b % compact_every == 0)
            e = k::pg_compact(n, tp.as<int32_t>(), ti_live.as<int32_t>(), live.as<int32_t>(), score.as<uint32_t>(), s);
        if (e == hipSuccess)
            e = k::pg_batch(n, cur, m->rp.as<int32_t>(), m->ci.as<int32_t>(), tp.as<int32_t>(), ti_live.as<int32_t>(),
                            score.as<uint32_t>(), Tl, st, sure.as<int32_t>(), tie.as<int32_t>(), hist.as<int32_t>(),
                            chunk.as<int32_t>(), placement.as<uint32_t>(), wprefix.as<int32_t>(), hcol.as<int32_t>(),
                            items.as<uint64_t>(), live.as<int32_t>(), P.Ps[(size_t)b], P.Pl[(size_t)b],
                            P.Os[(size_t)b], P.Ol[(size_t)b], P.hub_cap, any_wide ? &wide : nullptr, s);
        if (e != hipSuccess) rc = set_error(MR_ERR_HIP, "pGOrder batch %d failed: %s", b, hipGetErrorString(e));
        cur ^= 1;
        if (rc == MR_OK && (b % kChunk == kChunk - 1 || b == B - 1)) rc = publish(b);
    }
    if (rc == MR_OK && hipStreamSynchronize(s) != hipSuccess) rc = set_error(MR_ERR_HIP, "pGOrder loop failed");
    if (rc == MR_OK && any_wide) {
        k::PgState hst{};
        if (hipMemcpy(&hst, st, sizeof(hst), hipMemcpyDeviceToHost) != hipSuccess)
            rc = set_error(MR_ERR_HIP, "pGOrder state copy failed");
        else if (hst.overflow)
            rc = set_error(MR_ERR_NUMERIC, "pGOrder wide selection listed more vertices than the batch holds");
    }
    // on failure too, every queued copy into the pinned buffers must land
    // before their destructors free them
    if (rc != MR_OK) (void)hipStreamSynchronize(s);
    const double gpu_ms = ms_since(t_gpu);
    if (rc == MR_OK) {
        std::vector<int32_t> chk(ghost.p, ghost.p + n);
        if (mr_check_perm(chk.data(), n) != MR_OK)
            rc = set_error(MR_ERR_NUMERIC, "Gorder GPU loop produced an invalid placement");
    }
    const auto t_tail = clk::now();
    finish(rc == MR_OK);
    if (rc != MR_OK) return rc;
    if (abort_all) return set_error(MR_ERR_HIP, "waiting for the pGOrder placement copy failed");
    const double tail_ms = ms_since(t_tail);
    // positions outside every CPU slice keep the GPU placement
    {
        std::vector<uint8_t> in_slice((size_t)n, 0);
        for (int32_t c = 0; c < C; ++c)
            std::fill(in_slice.begin() + P.Cs[(size_t)c], in_slice.begin() + P.Cs[(size_t)c] + P.Cl[(size_t)c], 1);
        for (int32_t i = 0; i < n; ++i)
            if (!in_slice[(size_t)i]) order[(size_t)i] = ghost.p[i];
    }
    if (gpu_order) std::memcpy(gpu_order, ghost.p, sizeof(uint32_t) * (size_t)n);
    if (final_order) std::memcpy(final_order, order.data(), sizeof(uint32_t) * (size_t)n);
    if (stats) {
        stats->batches = B;
        stats->gpu_ms = gpu_ms;
        // host refinement: summed busy time over the threads that overlapped
        // the GPU loop; the part after the loop is in total_ms
        stats->cpu_ms = (double)cpu_busy_us.load() / 1000.0 / std::max(1, nth);
        stats->total_ms = ms_since(t_all);
    }
    (void)tail_ms;
    return MR_OK;
}

int mr_pgorder_order(mr_mat m, int width, int window, uint32_t* gpu_order, uint32_t* final_order,
                     mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "pGOrder needs a square matrix");
    if (window < 1) return set_error(MR_ERR_INVALID_ARG, "window must be >= 1");
    const int64_t n64 = m->n_rows, nnz = m->nnz;
    if (n64 == 0) return MR_OK;
    int W = width;
    if (W <= 0) MR_TRY(mr_get_width(n64, nnz, &W));
    if (W < 32 || W > (1 << 22) || (W & (W - 1)))
        return set_error(MR_ERR_UNSUPPORTED, "pGOrder GPU loop supports power-of-two widths 32..2^22 (got %d)", W);
    const int32_t n = (int32_t)n64;
    GorderPlan P;
    const int32_t B = pg_schedule(n, W, P.Ps, P.Pl, P.Os, P.Ol);
    // the reference never clips the first window (class.cu:3003); for n < 32 it
    // would read past the placement, here it is clipped to n
    if (P.Ol[0] > n) P.Ol[0] = n;
    for (int32_t b = 0; b < B; ++b)
        if (P.Os[(size_t)b] + P.Ol[(size_t)b] > n) P.Ol[(size_t)b] = n - P.Os[(size_t)b];
    // CPU phase: the windows of every batch from log2(W) (class.cu:3077-3093)
    for (int32_t b = (int32_t)std::log2((double)W); b < B; ++b) {
        P.Cs.push_back(P.Os[(size_t)b]);
        P.Cl.push_back(P.Ol[(size_t)b]);
    }
    P.start_columns = false;
    P.hub_cap = 10000;
    P.cpu_window = window;
    MR_TRY(gorder_family_order(m, P, gpu_order, final_order, stats));
    if (stats) {
        stats->width = W;
        stats->window = window;
    }
    return MR_OK;
}

int mr_pgorder_cpu_phase(mr_mat m, int width, int window, const uint32_t* gpu_order, uint32_t* final_order,
                         int threads, double* ms)
{
    if (!m || !gpu_order || !final_order) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "pGOrder needs a square matrix");
    if (window < 1) return set_error(MR_ERR_INVALID_ARG, "window must be >= 1");
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const int64_t n64 = m->n_rows, nnz = m->nnz;
    if (ms) *ms = 0.0;
    if (n64 == 0) return MR_OK;
    int W = width;
    if (W <= 0) MR_TRY(mr_get_width(n64, nnz, &W));
    if (W < 32 || W > (1 << 22) || (W & (W - 1)))
        return set_error(MR_ERR_UNSUPPORTED, "pGOrder supports power-of-two widths 32..2^22 (got %d)", W);
    const int32_t n = (int32_t)n64;
    {
        // the placement must be a permutation of the vertices (it indexes
        // the host lists below)
        std::vector<uint8_t> seen((size_t)n, 0);
        for (int32_t i = 0; i < n; ++i) {
            const uint32_t v = gpu_order[i];
            if (v >= (uint32_t)n || seen[v])
                return set_error(MR_ERR_NOT_PERMUTATION, "gpu_order is not a permutation of the %d vertices", n);
            seen[v] = 1;
        }
    }
    std::vector<int32_t> Ps, Pl, Os, Ol;
    const int32_t B = pg_schedule(n, W, Ps, Pl, Os, Ol);
    for (int32_t b = 0; b < B; ++b)
        if (Os[(size_t)b] + Ol[(size_t)b] > n) Ol[(size_t)b] = n - Os[(size_t)b];
    // the slices gorder6 refines on the host: the windows of every batch from
    // log2(width) (class.cu:3077-3093), as mr_pgorder_order
    std::vector<int32_t> Cs, Cl;
    for (int32_t b = (int32_t)std::log2((double)W); b < B; ++b) {
        Cs.push_back(Os[(size_t)b]);
        Cl.push_back(Ol[(size_t)b]);
    }
    // host lists: rows, and the column counts order_on_cpu reads
    std::vector<int32_t> hrp((size_t)n + 1), hci((size_t)std::max<int64_t>(nnz, 1)), htp((size_t)n + 1, 0);
    hipStream_t s = m->ctx->stream;
    MR_HIP_TRY(hipMemcpyAsync(hrp.data(), m->rp.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    if (nnz > 0)
        MR_HIP_TRY(hipMemcpyAsync(hci.data(), m->ci.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    for (int64_t a = 0; a < nnz; ++a) ++htp[(size_t)hci[(size_t)a] + 1];
    for (int32_t v = 0; v < n; ++v) htp[(size_t)v + 1] += htp[(size_t)v];
    std::memcpy(final_order, gpu_order, sizeof(uint32_t) * (size_t)n);
    std::vector<int32_t> local_idx((size_t)n, -1);
    // ti is not read by order_batch (only the column counts are)
    PgCpuShared g{hrp.data(), hci.data(), htp.data(), nullptr, local_idx.data(), (int32_t)window,
                  (int32_t)(std::sqrt((double)n) / 4)};
    const int32_t C = (int32_t)Cs.size();
    int nth = threads > 0 ? threads : host_threads();
    nth = std::max(1, std::min<int>(nth, std::max<int32_t>(C, 1)));
    std::atomic<int32_t> next{0};
    const auto t0 = clk::now();
    auto worker = [&] {
        PgCpuWorker w;
        w.yslot.assign((size_t)n, -1);
        for (int32_t c = next++; c < C; c = next++) {
            const int32_t cs = Cs[(size_t)c], cl = Cl[(size_t)c];
            for (int32_t i = 0; i < cl; ++i) local_idx[(size_t)gpu_order[cs + i]] = i;
            order_batch(g, w, final_order + cs, cl);
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nth; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    if (ms) *ms = ms_since(t0);
    return MR_OK;
}

int mr_gorder5_order(mr_mat m, int vsize, int width1, int width2, uint32_t* gpu_order, uint32_t* final_order,
                     mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "gorder5 needs a square matrix");
    const int64_t n64 = m->n_rows, nnz = m->nnz;
    if (n64 == 0) return MR_OK;
    if (vsize <= 0) MR_TRY(mr_get_width(n64, nnz, &vsize));
    if (vsize < 1 || vsize > (1 << 22))
        return set_error(MR_ERR_UNSUPPORTED, "gorder5: batch size must be in [1, 2^22] (got %d)", vsize);
    if (width1 < 1 || (int64_t)width1 * vsize > (int64_t)1 << 24)
        return set_error(MR_ERR_UNSUPPORTED, "gorder5: the GPU window (width1 x batch size) must hold 1..2^24 vertices");
    if (width2 < 1) return set_error(MR_ERR_INVALID_ARG, "gorder5: width2 must be >= 1");
    const int32_t n = (int32_t)n64;
    GorderPlan P;
    // fixed batches and a window of the last width1 batches (class.cu:2800-2813)
    const int32_t B = (n + vsize - 1) / vsize;
    for (int32_t b = 0; b < B; ++b) {
        const int32_t start = b * vsize;
        const int32_t len = (start + vsize < n) ? vsize : n - start;
        const int32_t os = (b < width1) ? 0 : (b - width1 + 1) * vsize;
        int32_t ol = (b < width1) ? (b + 1) * vsize : width1 * vsize;
        if (os + ol > n) ol = n - os;
        P.Ps.push_back(start);
        P.Pl.push_back(len);
        P.Os.push_back(os);
        P.Ol.push_back(ol);
        P.Cs.push_back(start);     // CPU phase on every batch (class.cu:2833-2850)
        P.Cl.push_back(len);
    }
    P.start_columns = true;
    P.hub_cap = INT32_MAX;
    P.cpu_window = width2;
    MR_TRY(gorder_family_order(m, P, gpu_order, final_order, stats));
    if (stats) {
        stats->width = vsize;
        stats->window = width2;
    }
    return MR_OK;
}

int mr_reorder_gorder5(mr_mat m, int vsize, int width1, int width2, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    const int64_t n = m->n_rows;
    if (n == 0) return MR_OK;
    const auto t0 = clk::now();
    std::vector<uint32_t> order((size_t)n);
    mr_reorder_stats st{};
    MR_TRY(mr_gorder5_order(m, vsize, width1, width2, nullptr, order.data(), &st));
    std::vector<int32_t> perm((size_t)n);
    for (int64_t i = 0; i < n; ++i) perm[(size_t)order[(size_t)i]] = (int32_t)i;
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        *stats = st;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

int mr_reorder_pgorder(mr_mat m, int width, int window, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    const int64_t n = m->n_rows;
    if (n == 0) return MR_OK;
    const auto t0 = clk::now();
    std::vector<uint32_t> order((size_t)n);
    mr_reorder_stats st{};
    MR_TRY(mr_pgorder_order(m, width, window, nullptr, order.data(), &st));
    std::vector<int32_t> perm((size_t)n);
    for (int64_t i = 0; i < n; ++i) perm[(size_t)order[(size_t)i]] = (int32_t)i;
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        *stats = st;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

}  // extern "C"

// a sequential host order over the CSR and CSC of the matrix (the CSC is
// built on the device), applied with the device relabel
template <typename F>
static int host_order_relabel(mr_mat m, F&& fn, int32_t width, int32_t window, int32_t* perm_out,
                              mr_reorder_stats* stats)
{
    const int64_t nnz = m->nnz;
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const auto t0 = clk::now();
    const int32_t n = (int32_t)m->n_rows;
    hipStream_t s = m->ctx->stream;
    DevBuf tp, ti;
    MR_TRY(gpu_transpose(m, tp, ti, s));
    std::vector<int32_t> rp((size_t)n + 1), ci((size_t)std::max<int64_t>(nnz, 1)), tpp((size_t)n + 1),
        tii((size_t)std::max<int64_t>(nnz, 1));
    MR_HIP_TRY(hipMemcpyAsync(rp.data(), m->rp.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(tpp.data(), tp.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    if (nnz > 0) {
        MR_HIP_TRY(hipMemcpyAsync(ci.data(), m->ci.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipMemcpyAsync(tii.data(), ti.p, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s));
    }
    MR_HIP_TRY(hipStreamSynchronize(s));
    std::vector<uint32_t> order((size_t)n);
    fn(n, rp.data(), ci.data(), tpp.data(), tii.data(), order.data());
    std::vector<int32_t> perm((size_t)n);
    for (int32_t i = 0; i < n; ++i) perm[(size_t)order[(size_t)i]] = i;
    const double order_ms = ms_since(t0);
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        stats->width = width;
        stats->window = window;
        stats->batches = 1;
        stats->gpu_ms = 0;
        stats->cpu_ms = order_ms;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

extern "C" {

int mr_reorder_gorder_cpu(mr_mat m, int window, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "Gorder needs a square matrix");
    if (window < 1) return set_error(MR_ERR_INVALID_ARG, "window must be >= 1");
    if (m->n_rows == 0) return MR_OK;
    return host_order_relabel(
        m,
        [&](int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti, uint32_t* order) {
            gorder_cpu(n, rp, ci, tp, ti, window, order);
        },
        0, window, perm_out, stats);
}

int mr_reorder_tile_order(mr_mat m, int width1, int width2, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "the tile order needs a square matrix");
    if (width1 < 1 || width2 < 1) return set_error(MR_ERR_INVALID_ARG, "tile order: widths must be >= 1");
    if (m->n_rows == 0) return MR_OK;
    return host_order_relabel(
        m,
        [&](int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti, uint32_t* order) {
            tile_order(n, rp, ci, tp, ti, width1, width2, order);
        },
        width1, width2, perm_out, stats);
}

int mr_reorder_rcm(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "RCM needs a square matrix");
    const int64_t n64 = m->n_rows, nnz = m->nnz;
    if (n64 == 0) return MR_OK;
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const auto t0 = clk::now();
    const int32_t n = (int32_t)n64;
    std::vector<int32_t> rp((size_t)n + 1), ci((size_t)std::max<int64_t>(nnz, 1));
    MR_TRY(mr_mat_export(m, rp.data(), ci.data(), nullptr));
    // the reference passes its COO (row_ind, col_ind); as a CSR of M the pair
    // is (col, row) -- RCM symmetrises, so the entry order is what matters:
    // the reference COO is sorted by (col, row) = rows of M in order.
    std::vector<int32_t> rows((size_t)std::max<int64_t>(nnz, 1)), cols((size_t)std::max<int64_t>(nnz, 1));
    for (int32_t i = 0; i < n; ++i)
        for (int32_t q = rp[(size_t)i]; q < rp[(size_t)i + 1]; ++q) {
            rows[(size_t)q] = ci[(size_t)q];
            cols[(size_t)q] = i;
        }
    std::vector<uint32_t> order((size_t)n);
    rcm_order(n, nnz, rows.data(), cols.data(), order.data());
    std::vector<int32_t> perm((size_t)n);
    for (int32_t i = 0; i < n; ++i) perm[(size_t)order[(size_t)i]] = i;
    const double order_ms = ms_since(t0);
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->cpu_ms = order_ms;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

// GRAPH::relabel_degree (src/class.cu:1496-1549) and GRAPH::dbg
// (src/class.cu:1683-1756): column counts of M on the device, the order
// (key desc, id asc) from one radix sort of composite keys, then the relabel
static int degree_relabel(mr_mat m, int grouped, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "a symmetric relabel needs a square matrix");
    const int64_t n = m->n_rows;
    if (n == 0) return MR_OK;
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const auto t0 = clk::now();
    hipStream_t s = m->ctx->stream;
    DevBuf cnt, dperm, ws;
    const size_t ws_bytes = k::degree_order_ws_bytes(n);
    MR_TRY(cnt.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(dperm.alloc(sizeof(int32_t) * (size_t)n));
    MR_TRY(ws.alloc(ws_bytes));
    MR_HIP_TRY(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));
    MR_HIP_TRY(k::transpose_count(m->nnz, m->ci.as<int32_t>(), cnt.as<int32_t>(), s));
    MR_HIP_TRY(k::degree_order(n, cnt.as<int32_t>(), grouped, dperm.as<int32_t>(), ws.p, ws_bytes, s));
    std::vector<int32_t> perm((size_t)n);
    MR_HIP_TRY(hipMemcpyAsync(perm.data(), dperm.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    const double order_ms = ms_since(t0);
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->gpu_ms = order_ms;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

int mr_reorder_degree(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats)
{
    return degree_relabel(m, 0, perm_out, stats);
}

int mr_reorder_dbg(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats)
{
    return degree_relabel(m, 1, perm_out, stats);
}

int mr_reorder_rabbit(mr_mat m, int reference_scatter, int32_t* perm_out, mr_reorder_stats* stats)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "Rabbit Order needs a square matrix");
    const int64_t n64 = m->n_rows, nnz = m->nnz;
    if (n64 == 0) return MR_OK;
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    const auto t0 = clk::now();
    const int32_t n = (int32_t)n64;
    std::vector<int32_t> rp((size_t)n + 1), ci((size_t)std::max<int64_t>(nnz, 1));
    MR_TRY(mr_mat_export(m, rp.data(), ci.data(), nullptr));
    std::vector<uint32_t> id((size_t)n);
    rabbit_order(n, rp.data(), ci.data(), id.data());
    std::vector<int32_t> perm((size_t)n);
    if (reference_scatter) {
        // class.cu:1622-1627 scatters the returned ids as a position -> vertex
        // sequence, i.e. applies the inverse of Rabbit's vertex -> id map
        for (int32_t i = 0; i < n; ++i) perm[(size_t)id[(size_t)i]] = i;
    } else {
        for (int32_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)id[(size_t)i];
    }
    const double order_ms = ms_since(t0);
    double rl = 0;
    MR_TRY(relabel(m, perm, &rl));
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->cpu_ms = order_ms;
        stats->relabel_ms = rl;
        stats->total_ms = ms_since(t0);
    }
    return MR_OK;
}

}  // extern "C"
