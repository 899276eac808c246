"""Benchmark: SpMV throughput after pGOrder on one MI355X.

Workload (the single-large-SpMV configuration): R-MAT scale 22 (4,194,304
rows), edge factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2, duplicates
merged, self loops kept, int32 CSR; FP32 values and x uniform in [-1,1].
Pipeline: pGOrder (GPU placement loop + per-batch CPU Gorder, window 5) on the
device-resident matrix, then MERBIT SpMV on the relabelled matrix.  One step =
one y = A x over the whole matrix with A, x, y resident in HBM.

Prints ONE JSON line on stdout; progress goes to stderr.  Extra figures
(unreordered MERBIT, measured vs RD-predicted speedup, CSR-vector and
BF16-value runs, pGOrder phase times) are under "extra".
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "matrix-reorder_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_spmv(ctx, mr, mat, xd, yd, kernel, steps, warmup):
    s = ctx.stream()
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
    ev = ctx.elapsed(e0, e1)
    return ev / steps, wall / steps


def algorithmic_bytes(n, nnz, val_bytes, vec_bytes=4):
    # nnz*(val + int32 col) + (n+1)*4 row pointer + n*|x| + n*|y|
    return nnz * (val_bytes + 4) + (n + 1) * 4 + n * vec_bytes + n * vec_bytes


def cpu_baseline(rp, ci, val32, x32, target_s=10.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    _, ms = oracle.spmv_f32_threads(rp, ci, val32, x32, cores, 2)
    per = ms / 2
    iters = max(3, min(400, int(target_s * 1e3 / max(per, 1e-3))))
    _, ms = oracle.spmv_f32_threads(rp, ci, val32, x32, cores, iters)
    per = ms / iters
    gflops = 2.0 * len(ci) / (per * 1e-3) / 1e9
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"oracle threaded CSR SpMV (FP32, nnz-balanced rows) on the same R-MAT s22 matrix, "
                      f"{iters} iterations, {per:.2f} ms/iter"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (this build has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)

    import mrpy as mr

    scale = a.scale
    log(f"generating R-MAT scale {scale}, edge factor 16, seed 2")
    t = time.perf_counter()
    rp, ci = mr.gen_rmat(scale, 16, (0.57, 0.19, 0.19), seed=2)
    n, nnz = len(rp) - 1, len(ci)
    val = mr.gen_values(102, nnz).astype(np.float32)
    x = mr.gen_values(103, n).astype(np.float32)
    log(f"n={n} nnz={nnz} ({time.perf_counter() - t:.1f} s)")

    ctx = mr.Ctx(0)
    arch, cus = ctx.arch()
    log(f"device {arch}, {cus} CUs")
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    xd = ctx.upload(x)
    yd = ctx.empty(n, np.float32)
    extra = {}

    if not a.no_extra:
        log("unreordered MERBIT")
        ms_orig, _ = time_spmv(ctx, mr, mat, xd, yd, mr.MERBIT, max(20, a.steps // 4), 5)
        extra["merbit_unreordered_ms"] = ms_orig
        log(f"  {ms_orig * 1e3:.1f} us/SpMV")
        log("RD headroom predictor on the unreordered matrix")
        t = time.perf_counter()
        pred, feat, rdst = mat.locality_predict(mr.MODEL_MERBIT)
        extra["rd_predicted_speedup"] = pred
        extra["rd_features"] = [round(float(v), 6) for v in feat]
        extra["rd_ms"] = (time.perf_counter() - t) * 1e3
        log(f"  predicted speedup {pred:.3f} ({extra['rd_ms']:.0f} ms)")

    log("pGOrder (width auto, window 5)")
    t = time.perf_counter()
    perm, st = mat.reorder_pgorder(0, 5)
    extra["pgorder"] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}
    extra["pgorder_wall_ms"] = (time.perf_counter() - t) * 1e3
    log(f"  {json.dumps(extra['pgorder'])}")
    # relabel x into the new order once (outside the timed region)
    xp = ctx.empty(n, np.float32)
    mat.permute_vec(0, xd, xp)
    mi = mat.merbit_info()
    extra["merbit_sigma"] = mi["sigma"]
    extra["merbit_build_ms"] = mi["build_ms"]

    log(f"MERBIT FP32 after pGOrder: {a.warmup} warmup + {a.steps} timed steps")
    ms, wall = time_spmv(ctx, mr, mat, xp, yd, mr.MERBIT, a.steps, a.warmup)
    gflops = 2.0 * nnz / (ms * 1e-3) / 1e9
    bytes_alg = algorithmic_bytes(n, nnz, 4)
    achieved = bytes_alg / (ms * 1e-3) / 1e9
    log(f"  {ms * 1e3:.1f} us/SpMV (wall {wall * 1e3:.1f} us), {gflops:.1f} GFLOP/s, {achieved:.0f} GB/s")

    if not a.no_extra:
        if "merbit_unreordered_ms" in extra:
            extra["measured_speedup"] = extra["merbit_unreordered_ms"] / ms
        for sg in (8, 32):
            mat.merbit_configure(sg)
            t_s, _ = time_spmv(ctx, mr, mat, xp, yd, mr.MERBIT, max(20, a.steps // 4), 5)
            extra[f"merbit_sigma{sg}_ms"] = t_s
        mat.merbit_configure(mi["sigma"])
        t_csr, _ = time_spmv(ctx, mr, mat, xp, yd, mr.CSR_VECTOR, max(20, a.steps // 4), 5)
        extra["csr_vector_ms"] = t_csr
        log(f"  CSR vector {t_csr * 1e3:.1f} us")
        mb = mat.convert(mr.BF16)
        t_bf, _ = time_spmv(ctx, mr, mb, xp, yd, mr.MERBIT, max(20, a.steps // 4), 5)
        extra["merbit_bf16_ms"] = t_bf
        extra["bf16_speedup_vs_fp32"] = ms / t_bf
        log(f"  MERBIT BF16 values {t_bf * 1e3:.1f} us ({ms / t_bf:.2f}x FP32)")
        mb.close()

    cpu = None
    if not a.no_cpu:
        log("CPU baseline (oracle threaded CSR SpMV, FP32)")
        cpu = cpu_baseline(rp, ci, val, x)
        log(f"  {cpu['value']} GFLOP/s on {cpu['cores']} cores")

    out = {
        "metric": "SpMV GFLOP/s + % HBM roofline after pGOrder; batched SpMV p50/p99 latency",
        "value": round(gflops, 3),
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
        "config": {"workload": f"R-MAT scale {scale} ef16 (0.57,0.19,0.19) seed 2, FP32, pGOrder w=5 + MERBIT SpMV",
                   "n_rows": n, "nnz": nnz, "kernel": "merbit", "sigma": mi["sigma"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    mat.close()
    ctx.close()


if __name__ == "__main__":
    main()
