// panel_plan.cpp -- plan of the column-panel MERBIT variant (policy bit 3).
//
// The device builds the panel-major streams in these steps (kernels in
// kernels/merbit_panel.hip):
//   1. segment starts: a row's first nonzero or a change of panel in the row;
//      their prefix sum numbers the segments in row-major order (sid);
//   2. a stable radix sort of the segments by panel gives panel-major order
//      (rows ascending inside a panel);
//   3. the host lays every panel on whole tiles and splits the tiles into the
//      workgroups' jobs (below);
//   4. every nonzero is scattered to its item; items closing a segment or a
//      tile are flagged, and the prefix sum of the flags numbers the partials;
//   5. every row lists its partials in row-major segment order (rs / gidx);
//   6. the row-sum stream: the host packs the rows into 512-item tiles that no
//      row crosses (runs of empty rows folded into the end marker before
//      them), the device writes each row's partial indices and an end marker
//      at its packed position.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels/launch_panel.hpp"
#include "kernels/launch_perm.hpp"
#include "kernels/launch_radix.hpp"

namespace mr {

namespace {

struct JobTable {
    std::vector<int32_t> jobs;   // (panel, first tile, end tile) triples
    std::vector<int32_t> wg;     // first job of every workgroup, plus the end
};

// Greedy split of the panels' tiles into at most `groups` contiguous job
// lists of cost <= budget.  C is the prefix sum of the tile costs (tile t
// costs C[t + 1] - C[t]); loading a panel into LDS costs `load`.  Returns the
// number of workgroups used.
int split_jobs(const std::vector<int32_t>& tile0, const std::vector<int32_t>& ntile, const std::vector<double>& C,
               double budget, double load, JobTable* out)
{
    out->jobs.clear();
    out->wg.assign(1, 0);
    double rem = budget;
    bool open = false;   // the current workgroup has a job
    for (size_t p = 0; p < ntile.size(); ++p) {
        int32_t t = tile0[p];
        const int32_t te = tile0[p] + ntile[p];
        while (t < te) {
            if (open && rem < load + (C[(size_t)t + 1] - C[(size_t)t])) {
                out->wg.push_back((int32_t)(out->jobs.size() / 3));
                rem = budget;
                open = false;
            }
            // the most tiles from t whose cost fits rem - load, at least one
            const double lim = C[(size_t)t] + (rem - load);
            int32_t e = (int32_t)(std::upper_bound(C.begin() + t + 1, C.begin() + te + 1, lim) - C.begin()) - 1;
            if (e <= t) e = t + 1;
            out->jobs.push_back((int32_t)p);
            out->jobs.push_back(t);
            out->jobs.push_back(e);
            rem -= (C[(size_t)e] - C[(size_t)t]) + load;
            open = true;
            t = e;
        }
    }
    if (open) out->wg.push_back((int32_t)(out->jobs.size() / 3));
    return (int)out->wg.size() - 1;
}

JobTable plan_jobs(const std::vector<int32_t>& tile0, const std::vector<int32_t>& ntile, const std::vector<double>& C,
                   int groups, double load)
{
    double total = 0.0, tmax = 1.0;
    for (size_t p = 0; p < ntile.size(); ++p)
        if (ntile[p] > 0) total += (C[(size_t)tile0[p] + ntile[p]] - C[(size_t)tile0[p]]) + load;
    for (size_t t = 0; t + 1 < C.size(); ++t) tmax = std::max(tmax, C[t + 1] - C[t]);
    JobTable best;
    if (total <= 0.0) {
        best.wg.assign(1, 0);
        return best;
    }
    // smallest budget whose greedy split fits the workgroups
    double lo = std::max(tmax + load, total / groups), hi = total + load + tmax;
    JobTable t;
    if (split_jobs(tile0, ntile, C, lo, load, &t) <= groups) return t;
    split_jobs(tile0, ntile, C, hi, load, &best);
    for (int it = 0; it < 40 && hi - lo > 0.5; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (split_jobs(tile0, ntile, C, mid, load, &t) <= groups) {
            hi = mid;
            best = t;
        } else {
            lo = mid;
        }
    }
    return best;
}

}  // namespace

int merbit_panel_build(mr_mat m, hipStream_t s)
{
    MerbitPlan& p = m->merbit;
    if (m->dtype == MR_F64) return set_error(MR_ERR_UNSUPPORTED, "the column-panel MERBIT variant needs F32/F16/BF16 values");
    if (p.pn_built) return MR_OK;
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = m->n_rows, nnz = m->nnz, nc = m->n_cols;
    const int tile = 64 * p.ps;
    const int64_t P = std::max<int64_t>(1, (nc + k::kPanelCols - 1) / k::kPanelCols);
    if (nnz + P * (int64_t)tile >= INT32_MAX)
        return set_error(MR_ERR_UNSUPPORTED, "column panels: %lld nonzeros exceed the int32 item index",
                         (long long)nnz);
    int key_bits = 1;
    while (key_bits < 31 && ((int64_t)1 << key_bits) < P) ++key_bits;

    DevBuf F, sid, total, scan_ws;
    MR_TRY(F.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
    MR_TRY(sid.alloc(sizeof(int32_t) * (size_t)(nnz + 1)));
    MR_TRY(total.alloc(sizeof(int64_t) * 4));
    // 1. segments in row-major order
    MR_HIP_TRY(k::panel_seg_flags(n, m->rp.as<int32_t>(), m->ci.as<int32_t>(), F.as<int32_t>(), s));
    int64_t n_seg = 0;
    if (nnz > 0) {
        MR_TRY(scan_ws.alloc(k::scan_workspace_bytes(nnz)));
        MR_HIP_TRY(k::exclusive_scan_i32(F.as<int32_t>(), nnz, sid.as<int32_t>(), total.as<int64_t>(), scan_ws.p, s));
        MR_HIP_TRY(k::panel_set_tail(sid.as<int32_t>(), nnz, total.as<int64_t>(), 0, s));
        MR_HIP_TRY(hipMemcpyAsync(&n_seg, total.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipStreamSynchronize(s));
    } else {
        MR_HIP_TRY(k::panel_set_tail(sid.as<int32_t>(), 0, nullptr, 0, s));
    }
    const int64_t ns1 = std::max<int64_t>(n_seg, 1);
    DevBuf k0, key, ident, skey, order;
    MR_TRY(k0.alloc(sizeof(int32_t) * (size_t)(n_seg + 1)));
    MR_TRY(key.alloc(sizeof(uint32_t) * (size_t)ns1));
    MR_TRY(ident.alloc(sizeof(uint32_t) * (size_t)ns1));
    MR_TRY(skey.alloc(sizeof(uint32_t) * (size_t)ns1));
    MR_TRY(order.alloc(sizeof(uint32_t) * (size_t)ns1));
    MR_HIP_TRY(k::panel_seg_info(nnz, m->ci.as<int32_t>(), F.as<int32_t>(), sid.as<int32_t>(), k0.as<int32_t>(),
                                 key.as<uint32_t>(), ident.as<uint32_t>(), s));
    MR_HIP_TRY(k::panel_set_tail(k0.as<int32_t>(), n_seg, nullptr, nnz, s));
    // 2. panel-major order of the segments (stable: rows stay ascending)
    if (n_seg > 0) {
        DevBuf sort_ws;
        const size_t wsb = k::radix_sort_ws_bytes(n_seg);
        MR_TRY(sort_ws.alloc(wsb));
        MR_HIP_TRY(k::radix_sort_pairs(key.as<uint32_t>(), ident.as<uint32_t>(), skey.as<uint32_t>(),
                                       order.as<uint32_t>(), n_seg, key_bits, sort_ws.p, wsb, s));
    }
    key.release();
    ident.release();
    DevBuf len, X, pfirst, pstart;
    MR_TRY(len.alloc(sizeof(int32_t) * (size_t)ns1));
    MR_TRY(X.alloc(sizeof(int32_t) * (size_t)(n_seg + 1)));
    MR_TRY(pfirst.alloc(sizeof(int32_t) * (size_t)P));
    MR_TRY(pstart.alloc(sizeof(int32_t) * (size_t)P));
    MR_HIP_TRY(hipMemsetAsync(pfirst.p, 0xff, pfirst.bytes, s));
    if (n_seg > 0) {
        DevBuf ws2;
        MR_TRY(ws2.alloc(k::scan_workspace_bytes(n_seg)));
        MR_HIP_TRY(k::panel_sorted_len(n_seg, order.as<uint32_t>(), k0.as<int32_t>(), len.as<int32_t>(), s));
        MR_HIP_TRY(k::exclusive_scan_i32(len.as<int32_t>(), n_seg, X.as<int32_t>(), nullptr, ws2.p, s));
        MR_HIP_TRY(k::panel_starts(n_seg, skey.as<uint32_t>(), X.as<int32_t>(), pfirst.as<int32_t>(),
                                   pstart.as<int32_t>(), s));
        MR_HIP_TRY(hipStreamSynchronize(s));
    }
    len.release();
    // 3. panels on whole tiles, tiles into jobs
    std::vector<int32_t> h_first((size_t)P), h_start((size_t)P), tile0((size_t)P, 0), ntile((size_t)P, 0),
        padd((size_t)P, 0);
    MR_HIP_TRY(hipMemcpyAsync(h_first.data(), pfirst.p, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(h_start.data(), pstart.p, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    int64_t tiles = 0;
    for (int64_t q = 0; q < P; ++q) {
        if (h_first[(size_t)q] < 0) continue;
        int64_t next = nnz;
        for (int64_t r = q + 1; r < P; ++r)
            if (h_first[(size_t)r] >= 0) {
                next = h_start[(size_t)r];
                break;
            }
        const int64_t items = next - h_start[(size_t)q];
        const int64_t nt = (items + tile - 1) / tile;
        tile0[(size_t)q] = (int32_t)tiles;
        ntile[(size_t)q] = (int32_t)nt;
        padd[(size_t)q] = (int32_t)(tiles * tile - h_start[(size_t)q]);
        tiles += nt;
    }
    const int64_t n_items = tiles * tile;
    DevBuf d_padd, pos;
    MR_TRY(d_padd.alloc(sizeof(int32_t) * (size_t)P));
    MR_TRY(pos.alloc(sizeof(int32_t) * (size_t)ns1));
    MR_HIP_TRY(hipMemcpyAsync(d_padd.p, padd.data(), sizeof(int32_t) * (size_t)P, hipMemcpyHostToDevice, s));
    MR_HIP_TRY(k::panel_seg_pos(n_seg, order.as<uint32_t>(), skey.as<uint32_t>(), X.as<int32_t>(),
                                d_padd.as<int32_t>(), pos.as<int32_t>(), s));
    skey.release();
    order.release();
    X.release();
    // 4. items and the partial numbering
    DevBuf flag, pidx;
    MR_TRY(p.pcol.alloc(sizeof(uint16_t) * (size_t)std::max<int64_t>(n_items, 1)));
    MR_TRY(p.pval.alloc(dtype_size(m->dtype) * (size_t)std::max<int64_t>(n_items, 1)));
    MR_TRY(flag.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(n_items, 1)));
    MR_TRY(pidx.alloc(sizeof(int32_t) * (size_t)(n_items + 1)));
    MR_HIP_TRY(k::panel_scatter(m->dtype, n_items, nnz, tile, m->ci.as<int32_t>(), m->val.p, F.as<int32_t>(),
                                sid.as<int32_t>(), k0.as<int32_t>(), pos.as<int32_t>(), p.pcol.as<uint16_t>(), p.pval.p,
                                flag.as<int32_t>(), s));
    int64_t n_parts = 0;
    if (n_items > 0) {
        DevBuf ws3;
        MR_TRY(ws3.alloc(k::scan_workspace_bytes(n_items)));
        MR_HIP_TRY(k::exclusive_scan_i32(flag.as<int32_t>(), n_items, pidx.as<int32_t>(), total.as<int64_t>(), ws3.p,
                                         s));
        MR_HIP_TRY(k::panel_set_tail(pidx.as<int32_t>(), n_items, total.as<int64_t>(), 0, s));
        MR_HIP_TRY(hipMemcpyAsync(&n_parts, total.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipStreamSynchronize(s));
    } else {
        MR_HIP_TRY(k::panel_set_tail(pidx.as<int32_t>(), 0, nullptr, 0, s));
    }
    flag.release();
    MR_TRY(p.ptseg.alloc(sizeof(int32_t) * (size_t)(tiles + 1)));
    MR_HIP_TRY(k::panel_tile_base(tiles, tile, pidx.as<int32_t>(), p.ptseg.as<int32_t>(), s));
    MR_HIP_TRY(k::panel_set_tail(p.ptseg.as<int32_t>(), tiles, nullptr, n_parts, s));
    // jobs, from per-tile costs.  A linear fit to per-workgroup clock stamps
    // of the panel kernel at R-MAT s22 (FP32) gives 0.168 us per tile of 3 KB,
    // 0.22 ns per partial sum the tile writes and 6.0 us per job (the x panel
    // load of the first job and the pipeline ramp; later panels are
    // prefetched into registers), 1.3 us residual.  In bytes: a partial costs
    // about 4 B, a job about the panel's x bytes; SpMV times over a grid of
    // both settled on 5 B and 1 panel (F32 and BF16 alike).  Partials vary
    // inside a panel: its first rows (the long ones after a locality order)
    // write few, its last rows about one per item.
    constexpr double kPartialBytes = 5.0, kJobLoad = 1.0;
    const double tile_bytes = (double)tile * (2.0 + (double)dtype_size(m->dtype));
    std::vector<int32_t> h_tseg((size_t)tiles + 1);
    MR_HIP_TRY(hipMemcpyAsync(h_tseg.data(), p.ptseg.p, sizeof(int32_t) * h_tseg.size(), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    std::vector<double> C((size_t)tiles + 1, 0.0);
    for (int64_t t = 0; t < tiles; ++t)
        C[(size_t)t + 1] = C[(size_t)t] + 1.0 + kPartialBytes * (double)(h_tseg[(size_t)t + 1] - h_tseg[(size_t)t]) / tile_bytes;
    const double load = kJobLoad * (double)(sizeof(float) * k::kPanelCols) / tile_bytes;
    JobTable jt = plan_jobs(tile0, ntile, C, std::max(1, m->ctx->cu_count), load);
    // the kernel maps its blocks onto the job lists by XCD: a multiple of 8
    // lists, the added ones empty
    if (jt.wg.size() > 1)
        while ((jt.wg.size() - 1) % 8 != 0) jt.wg.push_back(jt.wg.back());
    // 5. every row's partials, in row-major segment order
    DevBuf npieces, goff;
    MR_TRY(npieces.alloc(sizeof(int32_t) * (size_t)ns1));
    MR_TRY(goff.alloc(sizeof(int32_t) * (size_t)(n_seg + 1)));
    int64_t n_pieces = 0;
    if (n_seg > 0) {
        DevBuf ws4;
        MR_TRY(ws4.alloc(k::scan_workspace_bytes(n_seg)));
        MR_HIP_TRY(k::panel_pieces(n_seg, tile, pos.as<int32_t>(), k0.as<int32_t>(), npieces.as<int32_t>(), s));
        MR_HIP_TRY(k::exclusive_scan_i32(npieces.as<int32_t>(), n_seg, goff.as<int32_t>(), total.as<int64_t>() + 1,
                                         ws4.p, s));
        MR_HIP_TRY(k::panel_set_tail(goff.as<int32_t>(), n_seg, total.as<int64_t>() + 1, 0, s));
        MR_HIP_TRY(hipMemcpyAsync(&n_pieces, total.as<int64_t>() + 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MR_HIP_TRY(hipStreamSynchronize(s));
    } else {
        MR_HIP_TRY(k::panel_set_tail(goff.as<int32_t>(), 0, nullptr, 0, s));
    }
    if (n_pieces != n_parts)
        return set_error(MR_ERR_HIP, "column panels: %lld row pieces but %lld partials", (long long)n_pieces,
                         (long long)n_parts);
    MR_TRY(p.prs.alloc(sizeof(int32_t) * (size_t)(n + 1)));
    MR_TRY(p.pgidx.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(n_parts, 1)));
    MR_TRY(p.ppart.alloc(sizeof(float) * (size_t)std::max<int64_t>(n_parts, 1)));
    MR_HIP_TRY(k::panel_row_starts(n, m->rp.as<int32_t>(), sid.as<int32_t>(), goff.as<int32_t>(), p.prs.as<int32_t>(),
                                   s));
    MR_HIP_TRY(k::panel_gidx(n_seg, tile, pos.as<int32_t>(), k0.as<int32_t>(), goff.as<int32_t>(), pidx.as<int32_t>(),
                             p.pgidx.as<int32_t>(), s));
    // 6. the row-sum stream.  Rows are packed in order into tiles of 512 items
    //    and at most 1024 rows: a row takes its partials (the last one carries
    //    the end) and moves to the next tile when they or the row do not fit;
    //    an empty row after an end in the same tile adds one to that end's
    //    count instead (up to what the bits above the partial index hold); a
    //    row with more partials than a tile holds, or an empty row no end
    //    absorbs, takes a marker, and the long rows are listed.
    std::vector<int32_t> hrs((size_t)n + 1), rpos((size_t)std::max<int64_t>(n, 1)),
        rcnt((size_t)std::max<int64_t>(n, 1), 0), trow, lrow;
    MR_HIP_TRY(hipMemcpyAsync(hrs.data(), p.prs.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    constexpr int64_t RT = k::kRowSumTile, RR = k::kRowSumRows;
    int ib = 1;   // partial index width
    while (((int64_t)1 << ib) < n_parts) ++ib;
    if (ib > 30)
        return set_error(MR_ERR_UNSUPPORTED, "column panels: %lld partials exceed the row-sum stream's index",
                         (long long)n_parts);
    const int64_t cap_part = std::min<int64_t>(RR, ((int64_t)1 << (31 - ib)) - 1);   // rows one partial can end
    int64_t cur = 0, rt = 0, rows_in = 0;
    int64_t last_end = -1;      // row that owns the tile's last end (-1: none)
    int64_t last_cap = 0;       // how many rows that end may close
    if (n > 0) trow.push_back(0);
    for (int64_t r = 0; r < n; ++r) {
        const int64_t len = (int64_t)hrs[(size_t)r + 1] - hrs[(size_t)r];
        if (len == 0 && last_end >= 0 && rows_in < RR && rcnt[(size_t)last_end] < last_cap) {
            ++rcnt[(size_t)last_end];
            ++rows_in;
            continue;
        }
        int64_t L = len;
        bool part_end = len > 0;
        if (len == 0 || len > RT) {
            if (len > RT) lrow.push_back((int32_t)r);
            L = 1;
            part_end = false;
        }
        if (cur + L > RT || rows_in >= RR) {
            ++rt;
            cur = 0;
            rows_in = 0;
            trow.push_back((int32_t)r);
        }
        rpos[(size_t)r] = (int32_t)(rt * RT + cur);
        rcnt[(size_t)r] = 1;
        last_end = r;
        last_cap = part_end ? cap_part : RR;
        cur += L;
        ++rows_in;
        if (rt * RT + cur >= INT32_MAX)
            return set_error(MR_ERR_UNSUPPORTED, "column panels: the row-sum stream exceeds the int32 item index");
    }
    const int64_t n_rtiles = n > 0 ? rt + 1 : 0;
    trow.push_back((int32_t)n);
    {
        DevBuf d_rpos, d_rcnt;
        MR_TRY(d_rpos.alloc(sizeof(int32_t) * rpos.size()));
        MR_TRY(d_rcnt.alloc(sizeof(int32_t) * rcnt.size()));
        MR_TRY(p.prsi.alloc(sizeof(int32_t) * (size_t)std::max<int64_t>(n_rtiles * RT, 1)));
        MR_TRY(p.ptrow.alloc(sizeof(int32_t) * trow.size()));
        MR_HIP_TRY(hipMemcpyAsync(d_rpos.p, rpos.data(), sizeof(int32_t) * rpos.size(), hipMemcpyHostToDevice, s));
        MR_HIP_TRY(hipMemcpyAsync(d_rcnt.p, rcnt.data(), sizeof(int32_t) * rcnt.size(), hipMemcpyHostToDevice, s));
        MR_HIP_TRY(hipMemcpyAsync(p.ptrow.p, trow.data(), sizeof(int32_t) * trow.size(), hipMemcpyHostToDevice, s));
        MR_HIP_TRY(k::panel_rowsum_build(n, n_rtiles * RT, ib, p.prs.as<int32_t>(), p.pgidx.as<int32_t>(),
                                         d_rpos.as<int32_t>(), d_rcnt.as<int32_t>(), p.prsi.as<int32_t>(), s));
        MR_HIP_TRY(hipStreamSynchronize(s));   // the staging buffers and host vectors are freed on return
    }
    if (lrow.empty()) {
        p.prs.release();
        p.pgidx.release();
    } else {
        MR_TRY(p.plrow.alloc(sizeof(int32_t) * lrow.size()));
        MR_HIP_TRY(hipMemcpyAsync(p.plrow.p, lrow.data(), sizeof(int32_t) * lrow.size(), hipMemcpyHostToDevice, s));
    }
    MR_TRY(p.pjobs.alloc(sizeof(int32_t) * std::max<size_t>(jt.jobs.size(), 3)));
    MR_TRY(p.pwg.alloc(sizeof(int32_t) * jt.wg.size()));
    if (!jt.jobs.empty())
        MR_HIP_TRY(hipMemcpyAsync(p.pjobs.p, jt.jobs.data(), sizeof(int32_t) * jt.jobs.size(), hipMemcpyHostToDevice, s));
    MR_HIP_TRY(hipMemcpyAsync(p.pwg.p, jt.wg.data(), sizeof(int32_t) * jt.wg.size(), hipMemcpyHostToDevice, s));
    MR_HIP_TRY(hipStreamSynchronize(s));   // temporaries are freed on return
    p.ptiles = tiles;
    p.pparts = n_parts;
    p.psegs = n_seg;
    p.pn_wg = (int)jt.wg.size() - 1;
    p.pn_jobs = (int)(jt.jobs.size() / 3);
    p.prtiles = n_rtiles;
    p.prs_ib = ib;
    p.pn_long = (int)lrow.size();
    p.pn_built = true;
    p.pn_build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return MR_OK;
}

}  // namespace mr

namespace mr {

k::PanelView panel_view(mr_mat m)
{
    const MerbitPlan& p = m->merbit;
    k::PanelView v;
    v.S = p.ps;
    v.W = p.pw;
    v.nt = p.policy & 1;
    v.pcol = p.pcol.as<uint16_t>();
    v.pval = p.pval.p;
    v.tseg = p.ptseg.as<int32_t>();
    v.jobs = p.pjobs.as<int32_t>();
    v.wg = p.pwg.as<int32_t>();
    v.n_wg = p.pn_wg;
    v.n_cols = m->n_cols;
    v.n_rows = m->n_rows;
    v.items = p.prsi.as<int32_t>();
    v.rs_ib = p.prs_ib;
    v.trow = p.ptrow.as<int32_t>();
    v.n_rtiles = p.prtiles;
    v.lrow = p.plrow.as<int32_t>();
    v.n_long = p.pn_long;
    v.rs = p.prs.as<int32_t>();
    v.gidx = p.pgidx.as<int32_t>();
    return v;
}

}  // namespace mr
