"""Benchmark of the SpMV locality path on one MI355X.

Headline workload (the single-large-SpMV configuration): R-MAT scale 22
(4,194,304 rows), edge factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2,
duplicates merged, self loops kept, int32 CSR; FP32 values and x uniform in
[-1,1].  Pipeline: pGOrder (GPU placement loop + per-batch CPU Gorder, window
5) on the device-resident matrix, the MERBIT autotuner, then MERBIT SpMV on the
relabelled matrix.  One step = one y = A x over the whole matrix with A, x, y
resident in HBM; `value` is GFLOP/s = 2 nnz / step time (hipEvents on the
launch stream around the timed steps).

Further sections (reported under "extra", each on its own synthetic inputs):
  precision  3-D 27-point stencil 160^3 and R-MAT s21 ef32 with log-uniform
             values, stored BF16 / FP16 / FP32 (FP32 accumulation): max and
             relative-L2 error against the FP64 GPU product, time per dtype;
  batched    64 matrices (rows log-uniform in [2^10, 2^18], nnz/row
             log-uniform in [4, 64], half R-MAT, half banded, seed 3), each
             pGOrder-reordered, k = 1, 8, 32 right-hand sides in one
             persistent fused launch (autotuned, captured in a hipGraph);
  serving    32 resident reordered BF16 matrices, Poisson arrivals, k in
             {1, 4, 16}, every 1000th request uploads an R-MAT s18 that is
             reordered with pGOrder in the background; p50 / p99 latency;
  reference_scale  the shape of the reference's published MERBIT runs (symmetric
             R-MAT s24 ef4, all-ones values and x), unreordered and after
             pGOrder, beside the published sm_89 figures;
  cpu        the CPU baseline: the oracle's threaded CSR SpMV (FP32, and FP64
             beside it) on the same matrix, plus every reorderer on R-MAT s18 -- host
             (sequential Gorder, RCM, Rabbit Order, tile-aware order) and
             device (degree, DBG, gorder5, pGOrder) -- each with the SpMV
             time it yields.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import pathlib
import re
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "matrix-reorder_amd"))

HBM_PEAK_GBS = 8000.0
# rocprofv3 counter summaries of the headline kernels, by MERBIT variant
TRAFFIC_PROFILES = {"panel": ROOT / "profiles" / "merbit_panel_s22_fp32_pmc.md",
                    "lane_padded": ROOT / "profiles" / "merbit_lp_s22_fp32_pmc.md"}
_KERNELS = ROOT / "matrix-reorder_amd" / "csrc"
# the sources each summary measured (their hash is recorded in the summary)
TRAFFIC_SOURCES = {"panel": [_KERNELS / "kernels" / "merbit_panel.hip", _KERNELS / "kernels" / "launch_panel.hpp",
                             _KERNELS / "kernels" / "ldst.hpp", _KERNELS / "panel_plan.cpp"],
                   "lane_padded": [_KERNELS / "kernels" / "merbit_lp.hip"]}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def time_spmv(ctx, mr, mat, xd, yd, kernel, steps, warmup, stream=None):
    """Average ms per SpMV over `steps` launches on one stream (hipEvents), and
    the host wall time per step."""
    s = stream or ctx.stream()
    for _ in range(warmup):
        mat.spmv(xd, yd, kernel, s)
    ctx.sync(s)
    e0, e1 = ctx.event(), ctx.event()
    t0 = time.perf_counter()
    ctx.record(e0, s)
    for _ in range(steps):
        mat.spmv(xd, yd, kernel, s)
    ctx.record(e1, s)
    ctx.sync(s)
    wall = (time.perf_counter() - t0) * 1e3
    return ctx.elapsed(e0, e1) / steps, wall / steps


def time_per_launch(ctx, mat, xd, yd, kernel, launches, warmup=10):
    """ms of each of `launches` SpMVs, each bracketed by its own pair of
    hipEvents on one stream."""
    s = ctx.stream()
    for _ in range(warmup):
        mat.spmv(xd, yd, kernel, s)
    ev = [(ctx.event(), ctx.event()) for _ in range(launches)]
    for e0, e1 in ev:
        ctx.record(e0, s)
        mat.spmv(xd, yd, kernel, s)
        ctx.record(e1, s)
    ctx.sync(s)
    return np.array([ctx.elapsed(e0, e1) for e0, e1 in ev])


def csr_bytes(n_rows, n_cols, nnz, val_bytes, vec_bytes=4, k=1):
    """Algorithmic bytes of y = A X in canonical int32 CSR: every array once."""
    return nnz * (val_bytes + 4) + (n_rows + 1) * 4 + k * (n_cols + n_rows) * vec_bytes


def kernel_source_hash(policy):
    """First 16 hex digits of the SHA-256 over the sources of the MERBIT variant
    `policy` selects (kernel and plan files, in a fixed order)."""
    files = TRAFFIC_SOURCES["panel" if policy & 8 else "lane_padded"]
    h = hashlib.sha256()
    for f in files:
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def traffic_from_profile(policy):
    """HBM bytes per SpMV of the headline kernel(s) from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE scaled by the copy kernel).
    The summary names the MERBIT policy it was measured with and the hash of
    the sources it measured; a run that picked another policy, or runs changed
    sources, gets None rather than a figure of other code."""
    path = TRAFFIC_PROFILES["panel" if policy & 8 else "lane_padded"]
    if not path.exists():
        return None
    text = path.read_text()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    pm = re.search(r"^policy:\s*(\d+)", text, re.M)
    hm = re.search(r"^source_sha16:\s*([0-9a-f]+)", text, re.M)
    if not m or (pm and int(pm.group(1)) != policy):
        return None
    if not hm or hm.group(1) != kernel_source_hash(policy):
        return None
    return float(m.group(1))


def err_stats(y, yref):
    y = np.asarray(y, np.float64)
    yref = np.asarray(yref, np.float64)
    d = y - yref
    nrm = float(np.linalg.norm(yref))
    return {"max_abs": float(np.max(np.abs(d))) if d.size else 0.0,
            "rel_l2": float(np.linalg.norm(d) / nrm) if nrm > 0 else 0.0}


# ---------------------------------------------------------------------------
def section_main(mr, ctx, a, extra):
    log("generating R-MAT scale 22, edge factor 16, seed 2")
    t = time.perf_counter()
    rp, ci = mr.gen_rmat(22, 16, (0.57, 0.19, 0.19), seed=2)
    n, nnz = len(rp) - 1, len(ci)
    val = mr.gen_values(102, nnz).astype(np.float32)
    x = mr.gen_values(103, n).astype(np.float32)
    log(f"  n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    xd = ctx.upload(x)
    yd = ctx.empty(n, np.float32)
    if not a.only_main:
        log("unreordered matrix: MERBIT (default variant) and the RD headroom predictor")
        extra["merbit_unreordered_ms"], _ = time_spmv(ctx, mr, mat, xd, yd, mr.MERBIT, 30, 5)
        t = time.perf_counter()
        pred, feat, _ = mat.locality_predict(mr.MODEL_MERBIT)
        extra["rd_predicted_speedup"] = pred
        extra["rd_features"] = [round(float(v), 6) for v in feat]
        extra["rd_ms"] = (time.perf_counter() - t) * 1e3
        # the same prediction from the model refitted on MI355X timings
        pred_mi, feat_mi, _ = mat.locality_predict(mr.MODEL_MERBIT_MI355X)
        extra["rd_predicted_speedup_mi355x_model"] = pred_mi
        extra["rd_features_mi355x_model"] = [round(float(v), 6) for v in feat_mi]
        log(f"  {extra['merbit_unreordered_ms'] * 1e3:.1f} us/SpMV, predicted pGOrder speedup {pred:.3f} "
            f"(sm_89 model), {pred_mi:.3f} (MI355X model)")
        # pGOrder's CPU phase apart from the GPU loop it overlaps: order_on_cpu
        # on every slice on all host threads, the median of three runs
        gpu_order, final_order, _ = mat.pgorder_order(0, 5)
        runs = []
        for _ in range(3):
            fo, ms = mat.pgorder_cpu_phase(gpu_order, 0, 5, 0)
            runs.append(ms)
        if not np.array_equal(fo, final_order):
            raise RuntimeError("the CPU phase alone differs from pGOrder's own")
        extra["pgorder_cpu_phase_alone"] = {"median_ms": float(np.median(runs)), "runs_ms": [round(r, 1) for r in runs],
                                            "threads": host_cores()}
        log(f"  pGOrder CPU phase alone (order_on_cpu, {host_cores()} threads): median {np.median(runs):.0f} ms")
        del gpu_order, final_order, fo
    perm_path = pathlib.Path(a.perm_cache) if a.perm_cache else None
    if perm_path is not None and perm_path.exists():
        log(f"relabelling with the cached pGOrder permutation {perm_path}")
        perm = np.load(perm_path)
        mat.permute(perm)
        extra["pgorder"] = "cached permutation"
    else:
        log("pGOrder (width auto, window 5)")
        t = time.perf_counter()
        perm, st = mat.reorder_pgorder(0, 5)
        extra["pgorder"] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}
        extra["pgorder_wall_ms"] = (time.perf_counter() - t) * 1e3
        log(f"  {json.dumps(extra['pgorder'])}")
        if perm_path is not None:
            np.save(perm_path, perm)
    xp = ctx.empty(n, np.float32)
    mat.permute_vec(0, xd, xp)           # x in the new order, outside the timed region
    if not a.no_tune:
        log("MERBIT autotuner")
        best = mat.merbit_autotune(10)
        extra["autotune"] = best
        log(f"  {best}")
    if a.policy >= 0:
        mat.merbit_options(a.policy)
    cfg = {"policy": mat.merbit_policy(), **mat.merbit_info(), **mat.merbit_xc_info()}
    if cfg["policy"] & 8:
        cfg.update({f"panel_{k}": v for k, v in mat.merbit_panel_info().items()})
    log(f"MERBIT FP32 after pGOrder: {a.warmup} warmup + {a.steps} timed steps")
    ms, wall = time_spmv(ctx, mr, mat, xp, yd, mr.MERBIT, a.steps, a.warmup)
    gflops = 2.0 * nnz / (ms * 1e-3) / 1e9
    bytes_alg = csr_bytes(n, n, nnz, 4)
    achieved = bytes_alg / (ms * 1e-3) / 1e9
    log(f"  {ms * 1e3:.1f} us/SpMV (wall {wall * 1e3:.1f} us), {gflops:.1f} GFLOP/s, {achieved:.0f} GB/s")
    res = {"ms": ms, "gflops": gflops, "achieved": achieved, "bytes": bytes_alg, "n": n, "nnz": nnz, "cfg": cfg}
    # per-launch statistics: the median of 200 launches, each between its own
    # events, and the reference's figure, the mean over 400 back-to-back
    # launches (class.cu:1421-1432)
    per = time_per_launch(ctx, mat, xp, yd, mr.MERBIT, 200)
    extra["per_launch_us"] = {"median": float(np.median(per)) * 1e3, "p10": float(np.percentile(per, 10)) * 1e3,
                              "p90": float(np.percentile(per, 90)) * 1e3, "launches": int(len(per))}
    m400, _ = time_spmv(ctx, mr, mat, xp, yd, mr.MERBIT, 400, 10)
    extra["mean_us_over_400"] = m400 * 1e3
    # the stored format's bytes beside the canonical int32-CSR bytes (the
    # column panels keep 16-bit panel columns, and add partial sums)
    sb = mat.merbit_stored_bytes()
    extra["stored_format_bytes"] = sb
    extra["canonical_csr_bytes"] = bytes_alg
    extra["stored_format_GBs"] = sb / (ms * 1e-3) / 1e9 if sb > 0 else None
    log(f"  per launch: median {extra['per_launch_us']['median']:.1f} us (p10 {extra['per_launch_us']['p10']:.1f}, "
        f"p90 {extra['per_launch_us']['p90']:.1f}) over 200; mean over 400: {m400 * 1e3:.1f} us")
    log(f"  stored format {sb / 1e6:.1f} MB per SpMV ({extra['stored_format_GBs'] or 0:.0f} GB/s) vs canonical CSR "
        f"{bytes_alg / 1e6:.1f} MB")
    # copy-kernel reference of the same session (measured bandwidth; also the
    # calibration kernel of the rocprofv3 byte counters)
    nb = 1 << 30
    src, dst = ctx.empty(nb // 4, np.float32), ctx.empty(nb // 4, np.float32)
    s = ctx.stream()
    ctx.copy(dst, src, nb, s)
    e0, e1 = ctx.event(), ctx.event()
    ctx.record(e0, s)
    for _ in range(10):
        ctx.copy(dst, src, nb, s)
    ctx.record(e1, s)
    ctx.sync(s)
    cms = ctx.elapsed(e0, e1) / 10
    extra["copy_1GiB_GBs"] = 2.0 * nb / (cms * 1e-3) / 1e9
    src.free()
    dst.free()
    log(f"  copy kernel {extra['copy_1GiB_GBs']:.0f} GB/s (read + write)")
    if not a.only_main:
        if "merbit_unreordered_ms" in extra:
            extra["measured_speedup"] = extra["merbit_unreordered_ms"] / ms
        # the workgroup-tile kernel and the CSR kernels on the same matrix
        pol = mat.merbit_policy()
        mat.merbit_options(1)
        extra["merbit_workgroup_tiles_ms"], _ = time_spmv(ctx, mr, mat, xp, yd, mr.MERBIT, 30, 5)
        mat.merbit_options(pol)
        extra["csr_vector_ms"], _ = time_spmv(ctx, mr, mat, xp, yd, mr.CSR_VECTOR, 30, 5)
        extra["csr_scalar_ms"], _ = time_spmv(ctx, mr, mat, xp, yd, mr.CSR_SCALAR, 30, 5)
        for key in ("csr_vector", "csr_scalar"):
            extra[f"{key}_gbs"] = csr_bytes(n, n, nnz, 4) / (extra[f"{key}_ms"] * 1e-3) / 1e9
        log(f"  CSR vector {extra['csr_vector_ms'] * 1e3:.0f} us ({extra['csr_vector_gbs']:.0f} GB/s), "
            f"scalar {extra['csr_scalar_ms'] * 1e3:.0f} us ({extra['csr_scalar_gbs']:.0f} GB/s)")
        # lower-precision values of the same reordered matrix vs the FP64 product
        m64 = mat.convert(mr.F64)
        x64 = ctx.upload(np.ascontiguousarray(xp.to_host().astype(np.float64)))
        y64 = ctx.empty(n, np.float64)
        m64.spmv(x64, y64, mr.MERBIT)
        ctx.sync()
        yref = y64.to_host()
        m64.close()
        prec = {}
        for name, dt in (("f32", mr.F32), ("bf16", mr.BF16), ("f16", mr.F16)):
            m2 = mat if dt == mr.F32 else mat.convert(dt)
            if dt != mr.F32:
                m2.merbit_autotune(5)
            t2, _ = time_spmv(ctx, mr, m2, xp, yd, mr.MERBIT, 50, 5)
            e = err_stats(yd.to_host(), yref)
            prec[name] = {"ms": t2, "gflops": 2.0 * nnz / (t2 * 1e-3) / 1e9, **e}
            if dt != mr.F32:
                m2.close()
        extra["s22_values_vs_fp64"] = prec
        extra["bf16_speedup_vs_fp32"] = prec["f32"]["ms"] / prec["bf16"]["ms"]
        log(f"  BF16 {prec['bf16']['ms'] * 1e3:.1f} us ({extra['bf16_speedup_vs_fp32']:.2f}x FP32), "
            f"rel-L2 {prec['bf16']['rel_l2']:.2e}")
    mat.close()
    return res, (rp, ci, val, x)


def section_precision(mr, ctx, a, extra):
    """Stencil 160^3 and R-MAT s21 ef32, values log-uniform in [1e-3, 1] with
    random sign, stored BF16 / FP16 / FP32 (FP32 accumulation) vs FP64."""
    out = {}
    cases = [("stencil27_160", lambda: mr.gen_stencil27(160, 160, 160)),
             ("rmat_s21_ef32", lambda: mr.gen_rmat(21, 32, (0.57, 0.19, 0.19), seed=21))]
    for name, make in cases:
        t = time.perf_counter()
        rp, ci = make()
        n, nnz = len(rp) - 1, len(ci)
        val = mr.gen_values(303, nnz, kind=1, lo=1e-3, hi=1.0)
        x = mr.gen_values(304, n)
        log(f"precision sweep {name}: n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
        m64 = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F64)
        x64 = ctx.upload(x)
        y64 = ctx.empty(n, np.float64)
        m64.spmv(x64, y64, mr.MERBIT)
        ctx.sync()
        yref = y64.to_host()
        t64, _ = time_spmv(ctx, mr, m64, x64, y64, mr.MERBIT, 10, 2)
        res = {"n": n, "nnz": nnz, "f64_ms": t64}
        del x64, y64
        xd = ctx.upload(x.astype(np.float32))
        yd = ctx.empty(n, np.float32)
        for dname, dt in (("f32", mr.F32), ("f16", mr.F16), ("bf16", mr.BF16)):
            m = m64.convert(dt)
            m.merbit_autotune(5)
            ms, _ = time_spmv(ctx, mr, m, xd, yd, mr.MERBIT, 30, 3)
            e = err_stats(yd.to_host(), yref)
            res[dname] = {"ms": ms, "gflops": 2.0 * nnz / (ms * 1e-3) / 1e9,
                          "gbs_canonical": csr_bytes(n, n, nnz, 2 if dt != mr.F32 else 4) / (ms * 1e-3) / 1e9,
                          **e}
            log(f"  {dname}: {ms * 1e3:.1f} us, max {e['max_abs']:.2e}, rel-L2 {e['rel_l2']:.2e}")
            m.close()
        m64.close()
        out[name] = res
    extra["precision_sweep"] = out


def section_batched(mr, ctx, a, extra):
    t = time.perf_counter()
    shapes = mr.gen_batched_set(3, 64)
    mats = []
    t_re = 0.0
    for i, (rp, ci) in enumerate(shapes):
        val = mr.gen_values(3000 + i, len(ci)).astype(np.float32)
        m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
        t0 = time.perf_counter()
        m.reorder_pgorder(0, 5)
        t_re += time.perf_counter() - t0
        mats.append(m)
    tot_nnz = sum(len(c) for _, c in shapes)
    tot_rows = sum(len(r) - 1 for r, _ in shapes)
    log(f"batched: 64 matrices, {tot_rows} rows, {tot_nnz} nnz; pGOrder of all {t_re:.1f} s "
        f"(setup {time.perf_counter() - t:.1f} s)")
    res = {"matrices": 64, "rows": tot_rows, "nnz": tot_nnz, "pgorder_s": t_re}
    st = ctx.new_stream()
    for k in (1, 8, 32):
        xs, ys = [], []
        for m in mats:
            n = m.info()[0]
            xs.append(ctx.upload(mr.serve_rhs(k, n, k)))
            ys.append(ctx.empty((n, k), np.float32))
        b = mr.Batch(ctx, mats, k)
        b.set_io(xs, ys, stream=st)
        ctx.sync(st)
        tune = b.autotune(5)
        b.set_io(xs, ys, stream=st)
        b.capture(st)
        e0, e1 = ctx.event(), ctx.event()
        reps = 50
        b.launch(st)
        ctx.sync(st)
        ctx.record(e0, st)
        for _ in range(reps):
            b.launch(st)
        ctx.record(e1, st)
        ctx.sync(st)
        g_ms = ctx.elapsed(e0, e1) / reps
        ctx.record(e0, st)
        for _ in range(reps):
            b.run(st)
        ctx.record(e1, st)
        ctx.sync(st)
        e_ms = ctx.elapsed(e0, e1) / reps
        # the same work as one launch per matrix (reference style); one
        # untimed pass first builds each matrix's own plans
        def per_matrix():
            for j, m in enumerate(mats):
                if k == 1:
                    m.spmv(xs[j], ys[j], mr.MERBIT, st)
                else:
                    m.spmm(k, xs[j], ys[j], mr.MERBIT, stream=st)
        per_matrix()
        ctx.sync(st)
        ctx.record(e0, st)
        for _ in range(5):
            per_matrix()
        ctx.record(e1, st)
        ctx.sync(st)
        s_ms = ctx.elapsed(e0, e1) / 5
        info = b.info()
        res[f"k{k}"] = {"graph_ms": g_ms, "eager_ms": e_ms, "per_matrix_launches_ms": s_ms,
                        "gflops": info["flops"] / (g_ms * 1e-3) / 1e9,
                        "gbs_canonical": info["bytes"] / (g_ms * 1e-3) / 1e9,
                        "tiles": info["tiles"], "autotune": tune}
        log(f"  k={k}: graph {g_ms * 1e3:.1f} us, eager {e_ms * 1e3:.1f} us, per-matrix {s_ms * 1e3:.1f} us, "
            f"{res[f'k{k}']['gflops']:.0f} GFLOP/s, tune {tune}")
        b.close()
        del xs, ys
    extra["batched"] = res
    return mats


def section_serving(mr, ctx, a, extra, batch_mats):
    """32 resident BF16 matrices: R-MAT s22 and s21 ef32 reordered and tuned
    (single vectors through their best MERBIT variant), plus 30 of the batched
    set; Poisson loads with uploads of R-MAT s18 every 1000th.  Each rate runs
    with no deadline (every request completes, latency grows with the queue)
    and, at the overload rates, with a 2 ms queue deadline: requests queued
    longer are shed (timed_out) and achieved_rate is the sustained capacity."""
    t = time.perf_counter()
    resident = []
    for scale, ef, seed in ((22, 16, 2), (21, 32, 21)):
        rp, ci = mr.gen_rmat(scale, ef, (0.57, 0.19, 0.19), seed=seed)
        val = mr.gen_values(500 + scale, len(ci))
        m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.BF16)
        m.reorder_pgorder(0, 5)
        m.merbit_autotune(3)
        resident.append(m)
    for m in batch_mats[:30]:
        resident.append(m.convert(mr.BF16))
    log(f"serving: 32 resident BF16 matrices ({time.perf_counter() - t:.1f} s setup), large ones on policies "
        f"{[m.merbit_policy() for m in resident[:2]]}")

    def run(srv, rate, tag):
        nreq = int(min(max(rate * a.serve_seconds, 2000), 60000))
        r = srv.run_load(rate=rate, n_requests=nreq, seed=int(rate), ks=(1, 4, 16), upload_every=a.serve_upload_every,
                         upload_scale=18, upload_edge_factor=16)
        log(f"  {tag}{rate:.0f} req/s: p50 {r['p50_us']:.0f} us, p99 {r['p99_us']:.0f} us, {r['gflops']:.0f} GFLOP/s, "
            f"{r['completed']} done, {r['timed_out']} shed, {r['achieved_rate']:.0f} done/s, "
            f"heavy {r['completed_heavy']} done / {r['timed_out_heavy']} shed, p99 light {r['p99_light_us']:.0f} us, "
            f"heavy {r['p99_heavy_us']:.0f} us, "
            f"{r['uploads_admitted']} uploads ({r['upload_ms_mean']:.0f} ms each), "
            f"{r['mean_requests_per_batch']:.1f} req/batch")
        return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}

    res = {}
    cfg = {"slots": a.serve_slots, "heavy_slots": a.serve_heavy_slots, "max_jobs": 256, "coalesce": 1,
           "upload_threads": a.serve_upload_threads}
    res["config"] = dict(cfg, hw_queues=int(os.environ.get("GPU_MAX_HW_QUEUES", "4")))
    srv = mr.Server(ctx, **cfg)
    for m in resident:
        srv.add(m)
    for rate in a.serve_rates:
        res[f"{int(rate)}"] = run(srv, rate, "")
    srv.close()
    srv = mr.Server(ctx, timeout_ms=2.0, **cfg)
    for m in resident:
        srv.add(m)
    res["deadline_2ms"] = {f"{int(rate)}": run(srv, rate, "deadline 2 ms, ") for rate in a.serve_rates if rate >= 20000}
    # an upload of a matrix seen before: pGOrder replaced by the plan cache
    import tempfile
    with tempfile.TemporaryDirectory(prefix="mr_plans_") as plans:
        srv.set_plan_cache(plans)
        up = {}
        for key in ("first_upload_ms", "repeat_upload_ms"):
            st, mid, ms = srv.upload_wait(srv.upload_rmat(18, 16, seed=1818))
            up[key] = round(ms, 1) if st == 0 else f"status {st}"
        up["hits_misses"] = list(srv.plan_cache_stats())
        srv.set_plan_cache(None)
    res["plan_cache_s18_upload"] = up
    log(f"  R-MAT s18 upload with the plan cache: first {up['first_upload_ms']} ms, repeat "
        f"{up['repeat_upload_ms']} ms")
    srv.close()
    for m in resident:
        m.close()
    extra["serving"] = res


def section_reference_scale(mr, ctx, a, extra):
    """The shape of the reference's published MERBIT runs: a symmetric R-MAT
    s24, edge factor 4, (0.57, 0.19, 0.19), values and x all ones, FP32;
    MERBIT unreordered and after pGOrder at get_width's width (16384 here, as
    in the reference's runs)."""
    t = time.perf_counter()
    rp, ci = mr.gen_rmat(24, 4, (0.57, 0.19, 0.19), seed=24, symmetric=True)
    n, nnz = len(rp) - 1, len(ci)
    log(f"reference-scale R-MAT s24 ef4 symmetric: n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")
    m = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F32)
    del rp, ci
    xd = ctx.upload(np.ones(n, np.float32))
    yd = ctx.empty(n, np.float32)
    out = {"workload": "R-MAT s24 ef4 (0.57,0.19,0.19) symmetric, values 1, x 1, FP32", "n": n, "nnz": nnz}
    m.merbit_autotune(5)
    out["unreordered_ms"], _ = time_spmv(ctx, mr, m, xd, yd, mr.MERBIT, 50, 5)
    t = time.perf_counter()
    _, st = m.reorder_pgorder(0, 5)
    out["pgorder"] = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in st.items()}
    out["pgorder_wall_ms"] = (time.perf_counter() - t) * 1e3
    m.merbit_autotune(5)
    out["pgorder_ms"], _ = time_spmv(ctx, mr, m, xd, yd, mr.MERBIT, 50, 5)
    for key in ("unreordered", "pgorder"):
        ms = out[f"{key}_ms"]
        out[f"{key}_gflops"] = 2.0 * nnz / (ms * 1e-3) / 1e9
        out[f"{key}_gbs"] = csr_bytes(n, n, nnz, 4) / (ms * 1e-3) / 1e9
    out["reference_published_sm89"] = {"unreordered_gflops": 95.1, "gorder_gflops": 153.8}
    log(f"  unreordered {out['unreordered_ms'] * 1e3:.0f} us ({out['unreordered_gflops']:.0f} GFLOP/s), pGOrder "
        f"{out['pgorder_wall_ms'] / 1e3:.1f} s, after {out['pgorder_ms'] * 1e3:.0f} us ({out['pgorder_gflops']:.0f} GFLOP/s)")
    m.close()
    extra["reference_scale"] = out


def section_cpu(mr, ctx, a, data):
    """The oracle's threaded CSR SpMV on the headline matrix (unreordered), on
    the host's cores: FP32 (the reported baseline) and FP64, each the mean over
    up to 400 back-to-back iterations after warm-up (about 10 s in all) and
    the median of 11 single-iteration runs."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    rp, ci, val, x = data
    cores = max(1, min(host_cores(), 16))
    out = {}
    for name, fn, vt in (("fp32", oracle.spmv_f32_threads, np.float32), ("fp64", oracle.spmv_f64_threads, np.float64)):
        v, xx = val.astype(vt), x.astype(vt)
        _, ms = fn(rp, ci, v, xx, cores, 10)                      # warm-up
        per = ms / 10
        iters = max(3, min(400, int(5.0e3 / max(per, 1e-3))))
        _, ms = fn(rp, ci, v, xx, cores, iters)
        mean = ms / iters
        med = float(np.median([fn(rp, ci, v, xx, cores, 1)[1] for _ in range(11)]))
        out[name] = {"mean_ms": mean, "median_ms": med, "iters": iters,
                     "gflops": 2.0 * len(ci) / (mean * 1e-3) / 1e9,
                     "median_gflops": 2.0 * len(ci) / (med * 1e-3) / 1e9}
    f32 = out["fp32"]
    return {"value": round(f32["gflops"], 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"oracle threaded CSR SpMV (FP32, nnz-balanced row ranges) on the unreordered R-MAT s22 "
                      f"matrix, mean over {f32['iters']} iterations after 10 warm-ups, {f32['mean_ms']:.2f} ms/iter "
                      f"(median of 11 single iterations {f32['median_ms']:.2f} ms)",
            "fp32_median_gflops": round(f32["median_gflops"], 3),
            "fp64_gflops": round(out["fp64"]["gflops"], 3), "fp64_median_gflops": round(out["fp64"]["median_gflops"], 3),
            "fp64_iters": out["fp64"]["iters"],
            "cpu_model": cpu_model(), "host_threads": host_cores(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def section_cpu_reorders(mr, ctx, extra):
    """Every reorderer on R-MAT s18, each followed by the MERBIT SpMV (FP32) it
    yields: the host ones (sequential Gorder, RCM, Rabbit Order, the
    tile-aware order) and the device ones (degree, DBG, gorder5, pGOrder).
    Reorder times are the median of three runs."""
    rp, ci = mr.gen_rmat(18, 16, (0.57, 0.19, 0.19), seed=18)
    n, nnz = len(rp) - 1, len(ci)
    val = mr.gen_values(18, nnz).astype(np.float32)
    x = mr.gen_values(19, n).astype(np.float32)
    out = {"workload": "R-MAT s18 ef16 seed 18, FP32", "n": n, "nnz": nnz, "cores": host_cores()}
    xd = ctx.upload(x)
    yd = ctx.empty(n, np.float32)
    for name, fn in (("none", None), ("gorder_cpu_w5", lambda m: m.reorder_gorder_cpu(5)),
                     ("rcm", lambda m: m.reorder_rcm()), ("rabbit", lambda m: m.reorder_rabbit(False)),
                     ("tile_order_1024_512", lambda m: m.reorder_tile_order(1024, 512)),
                     ("degree", lambda m: m.reorder_degree()), ("dbg", lambda m: m.reorder_dbg()),
                     ("gorder5", lambda m: m.reorder_gorder5(0, 1, 5)),
                     ("pgorder", lambda m: m.reorder_pgorder(0, 5))):
        # the median of three runs, each on a fresh copy (one run when the
        # first takes over 60 s)
        res, walls = {}, []
        while True:
            m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
            if fn is None:
                break
            t = time.perf_counter()
            _, st = fn(m)
            walls.append((time.perf_counter() - t) * 1e3)
            res.update({k: v for k, v in st.items() if k.endswith("_ms")})
            if len(walls) == 3 or walls[0] > 60e3:
                break
            m.close()
        if walls:
            res["wall_ms"] = float(np.median(walls))
            res["wall_ms_runs"] = [round(w, 1) for w in walls]
        m.merbit_autotune(5)
        res["spmv_us"] = time_spmv(ctx, mr, m, xd, yd, mr.MERBIT, 50, 5)[0] * 1e3
        out[name] = res
        m.close()
    log("  reorderers on s18: " + ", ".join(
        f"{k} {v.get('wall_ms', 0):.0f} ms -> {v['spmv_us']:.1f} us" for k, v in out.items() if isinstance(v, dict)))
    extra["cpu_reorders"] = out


def gorder_cpu_s22_child():
    """Child process: the sequential CPU Gorder (gorder1, w = 5) on the headline
    R-MAT s22 matrix; prints one JSON line with its time."""
    import mrpy as mr

    rp, ci = mr.gen_rmat(22, 16, (0.57, 0.19, 0.19), seed=2)
    ctx = mr.Ctx(0)
    m = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F32)
    t = time.perf_counter()
    _, st = m.reorder_gorder_cpu(5)
    print(json.dumps({"wall_ms": (time.perf_counter() - t) * 1e3, **{k: v for k, v in st.items()}}), flush=True)
    m.close()
    ctx.close()


def section_cpu_gorder_s22(a, extra):
    """The reference's CPU Gorder (gorder1, window 5, one thread) on the
    headline s22 matrix, in a child process under a time limit, beside the
    pGOrder time of the main section."""
    import subprocess

    log(f"sequential CPU Gorder (w=5) on R-MAT s22, limit {a.cpu_gorder_timeout:.0f} s")
    t = time.perf_counter()
    try:
        # the child's progress lines (MR_PROGRESS) go straight to this stderr
        r = subprocess.run([sys.executable, str(pathlib.Path(__file__).resolve()), "--child-gorder-cpu"],
                           stdout=subprocess.PIPE, text=True, timeout=a.cpu_gorder_timeout,
                           env=dict(os.environ, MR_PROGRESS="1"))
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        out = json.loads(line[-1]) if r.returncode == 0 and line else {"error": f"exit status {r.returncode}"}
    except subprocess.TimeoutExpired:
        out = {"timed_out_after_s": a.cpu_gorder_timeout}
    out["child_wall_s"] = time.perf_counter() - t
    out["cpu_model"] = cpu_model()
    out["threads"] = 1
    out["pgorder_total_ms"] = extra.get("pgorder", {}).get("total_ms") if isinstance(extra.get("pgorder"), dict) else None
    log(f"  {out}")
    extra["cpu_gorder_s22"] = out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--only-main", action="store_true", help="headline SpMV only (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--perm-cache", default="", help="load (or save) the s22 pGOrder permutation here")
    ap.add_argument("--policy", type=int, default=-1, help="force a MERBIT load policy (after the autotuner)")
    ap.add_argument("--sections", default="precision,batched,serving,reference_scale,cpu_reorders")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[5000.0, 20000.0, 50000.0])
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-upload-every", type=int, default=1000)
    ap.add_argument("--cpu-gorder-timeout", type=float, default=900.0,
                    help="time limit of the cpu_gorder_s22 section (not in the default sections)")
    ap.add_argument("--child-gorder-cpu", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-heavy-slots", type=int, default=1)
    ap.add_argument("--serve-upload-threads", type=int, default=0, help="host threads of an upload's pGOrder (0 = half)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process (0 = keep the environment's)")
    a = ap.parse_args()
    # The serving section keeps four slot streams and an upload stream busy at
    # once.  With HIP's default of 4 hardware queues per process some of them
    # share a queue, and a light request's kernels wait behind another
    # stream's (at 20k req/s with a 2 ms deadline, p99 2.6-2.9 ms on 4 queues,
    # 1.9-2.1 ms on 8; the single-stream sections do not depend on it).  The
    # runtime reads it at its first call, which comes after this line.
    if a.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(a.hw_queues, 32))
    if a.child_gorder_cpu:
        gorder_cpu_s22_child()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (this build has no multi-GPU path; "
              f"the sharding design is documented instead)", file=sys.stderr)
        sys.exit(2)

    import mrpy as mr

    ctx = mr.Ctx(0)
    arch, cus = ctx.arch()
    log(f"device {arch}, {cus} CUs")
    extra = {}
    main_res, data = section_main(mr, ctx, a, extra)
    sections = set() if a.only_main else set(s for s in a.sections.split(",") if s)
    if "precision" in sections:
        section_precision(mr, ctx, a, extra)
    batch_mats = []
    if "batched" in sections or "serving" in sections:
        batch_mats = section_batched(mr, ctx, a, extra)
    if "serving" in sections:
        section_serving(mr, ctx, a, extra, batch_mats)
    for m in batch_mats:
        m.close()
    if "reference_scale" in sections:
        section_reference_scale(mr, ctx, a, extra)
    if "cpu_gorder_s22" in sections:
        section_cpu_gorder_s22(a, extra)
    cpu = None
    if not a.no_cpu and not a.only_main:
        log("CPU baseline (oracle threaded CSR SpMV, FP32 and FP64)")
        cpu = section_cpu(mr, ctx, a, data)
        log(f"  FP32 {cpu['value']} GFLOP/s (median {cpu['fp32_median_gflops']}), FP64 {cpu['fp64_gflops']} GFLOP/s "
            f"(median {cpu['fp64_median_gflops']}) on {cpu['cores']} cores")
        if "cpu_reorders" in sections:
            section_cpu_reorders(mr, ctx, extra)

    ms, n, nnz = main_res["ms"], main_res["n"], main_res["nnz"]
    cfg = main_res["cfg"]
    out = {
        "metric": "SpMV GFLOP/s + % HBM roofline after pGOrder; batched SpMV p50/p99 latency",
        "value": round(main_res["gflops"], 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "R-MAT scale 22 ef16 (0.57,0.19,0.19) seed 2, FP32 values, pGOrder w=5 + "
                               "autotuned MERBIT SpMV",
                   "n_rows": n, "nnz": nnz, "kernel": "merbit", **cfg},
        "roofline": {"bound": "hbm", "achieved": round(main_res["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(main_res["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(cfg["policy"])},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
