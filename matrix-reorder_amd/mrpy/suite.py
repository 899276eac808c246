"""The reference's R-MAT s24 experiment, re-run on this library.

The reference publishes MERBIT times for 75 Graph500 R-MAT s24 matrices
(config/config3.txt ids 1-75: edge factor 4 / 8 / 12, a from 0.43 to 0.91,
b = c, d = 0.05), unreordered ("origin") and after its CPU Gorder ("gorder1"),
measured on sm_89 with FP32 values 1 and x = 1 (test_rmat_meribit.txt).  The
matrix files are external Graph500 outputs, so each one is generated here with
the same scale, edge factor and (a, b, c), symmetrised as the reference's
loader does, with duplicates merged.  The workflow is the pGOrder branch of
src/main.cu:122-140: locality_predict on the unreordered matrix, reorder,
spmv.  Per matrix it records:

  origin     MERBIT time on the generated order (autotuned variant);
  predicted  the RD headroom predictor's speedup (sm_89 and MI355X models);
  pgorder    pGOrder time (width 0 = get_width, 16384 at s24, as the
             reference runs it) and the MERBIT time after it;

beside the published origin / gorder1 times of the same id
(suite/rmat_s24_published.json).

Usage, from the package directory:
    python -m mrpy.suite --ids 1-25 --out /tmp/suite/rmat_1_25.json
    python -m mrpy.suite --summarize run1.json run2.json --out /tmp/suite/rmat_s24_mi355x.json
"""
from __future__ import annotations

import argparse
import json
import pathlib
import sys
import time

import numpy as np

PUBLISHED = pathlib.Path(__file__).resolve().parent.parent / "suite" / "rmat_s24_published.json"


def parse_ids(text: str) -> list:
    out = []
    for part in text.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        elif part:
            out.append(int(part))
    return out


def time_merbit(ctx, mr, mat, xd, yd, reps: int = 30, warmup: int = 5) -> float:
    s = ctx.stream()
    for _ in range(warmup):
        mat.spmv(xd, yd, mr.MERBIT, s)
    ctx.sync(s)
    e0, e1 = ctx.event(), ctx.event()
    ctx.record(e0, s)
    for _ in range(reps):
        mat.spmv(xd, yd, mr.MERBIT, s)
    ctx.record(e1, s)
    ctx.sync(s)
    return ctx.elapsed(e0, e1) / reps


def log(*a):
    print("[suite]", *a, file=sys.stderr, flush=True)


def run(ids, out_path: str, width: int = 0) -> list:
    import mrpy as mr

    pub = {m["id"]: m for m in json.loads(PUBLISHED.read_text())["matrices"]}
    ctx = mr.Ctx(0)
    results = []
    path = pathlib.Path(out_path)
    path.parent.mkdir(parents=True, exist_ok=True)
    for gid in ids:
        p = pub[gid]
        t = time.perf_counter()
        rp, ci = mr.gen_rmat(p["scale"], p["ef"], (p["a"], p["b"], p["b"]), seed=gid, symmetric=True)
        n, nnz = len(rp) - 1, len(ci)
        rec = {"id": gid, "scale": p["scale"], "ef": p["ef"], "a": p["a"], "b": p["b"], "n": n, "nnz": nnz,
               "published_nnz": 2 * p["file_nnz"], "gen_s": round(time.perf_counter() - t, 2)}
        mat = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F32)
        del rp, ci
        xd = ctx.upload(np.ones(n, np.float32))
        yd = ctx.empty(n, np.float32)
        mat.merbit_autotune(5)
        rec["origin_ms"] = time_merbit(ctx, mr, mat, xd, yd)
        sp, f_ref, _ = mat.locality_predict(mr.MODEL_MERBIT)
        sp_mi, f_mi, _ = mat.locality_predict(mr.MODEL_MERBIT_MI355X)
        rec["predicted_speedup_sm89_model"] = sp
        rec["predicted_speedup_mi355x_model"] = sp_mi
        rec["rd_features"] = [float(v) for v in f_ref]
        rec["rd_features_mi355x_model"] = [float(v) for v in f_mi]   # under that model's bin edges
        log(f"id {gid}: a={p['a']} ef={p['ef']} nnz={nnz} origin {rec['origin_ms'] * 1e3:.0f} us")
        t = time.perf_counter()
        _, st = mat.reorder_pgorder(width, 5)
        rec["pgorder_wall_ms"] = (time.perf_counter() - t) * 1e3
        rec["pgorder_gpu_ms"], rec["pgorder_cpu_ms"] = st["gpu_ms"], st["cpu_ms"]
        mat.merbit_autotune(5)
        rec["pgorder_ms"] = time_merbit(ctx, mr, mat, xd, yd)
        rec["origin_gflops"] = 2.0 * nnz / (rec["origin_ms"] * 1e-3) / 1e9
        rec["pgorder_gflops"] = 2.0 * nnz / (rec["pgorder_ms"] * 1e-3) / 1e9
        rec["measured_speedup"] = rec["origin_ms"] / rec["pgorder_ms"]
        for key in ("origin_merbit_ms", "gorder1_merbit_ms", "gorder1_reorder_ms"):
            rec[f"published_{key}"] = p[key]
        rec["published_speedup"] = p["origin_merbit_ms"] / p["gorder1_merbit_ms"]
        rec["published_origin_gflops"] = 2.0 * rec["published_nnz"] / (p["origin_merbit_ms"] * 1e-3) / 1e9
        rec["published_gorder1_gflops"] = 2.0 * rec["published_nnz"] / (p["gorder1_merbit_ms"] * 1e-3) / 1e9
        results.append(rec)
        mat.close()
        xd.free()
        yd.free()
        log(f"id {gid}: pGOrder {rec['pgorder_wall_ms'] / 1e3:.1f} s, after {rec['pgorder_ms'] * 1e3:.0f} us "
            f"({rec['pgorder_gflops']:.0f} GFLOP/s; published after Gorder {rec['published_gorder1_gflops']:.0f})")
        path.write_text(json.dumps(results, indent=1) + "\n")
    ctx.close()
    return results


def geomean(v) -> float:
    v = [float(x) for x in v]
    return float(np.exp(np.mean(np.log(v)))) if v else float("nan")


def summarize(paths, out_path: str = "") -> dict:
    """Merge run files into one record per id and the sweep-wide figures."""
    rows = {}
    for pth in paths:
        for r in json.loads(pathlib.Path(pth).read_text()):
            rows[r["id"]] = r
    rs = [rows[k] for k in sorted(rows)]
    summ = {
        "matrices": len(rs),
        "geomean_origin_gflops": geomean(r["origin_gflops"] for r in rs),
        "geomean_pgorder_gflops": geomean(r["pgorder_gflops"] for r in rs),
        "geomean_published_origin_gflops": geomean(r["published_origin_gflops"] for r in rs),
        "geomean_published_gorder1_gflops": geomean(r["published_gorder1_gflops"] for r in rs),
        "geomean_measured_speedup": geomean(r["measured_speedup"] for r in rs),
        "geomean_published_speedup": geomean(r["published_speedup"] for r in rs),
        "geomean_after_reorder_vs_published": geomean(r["pgorder_gflops"] / r["published_gorder1_gflops"] for r in rs),
        "mean_pgorder_s": float(np.mean([r["pgorder_wall_ms"] for r in rs]) / 1e3),
        "mean_published_gorder1_s": float(np.mean([r["published_gorder1_reorder_ms"] for r in rs]) / 1e3),
        "median_abs_log_err_sm89_model": float(np.median([abs(np.log(r["predicted_speedup_sm89_model"] /
                                                                     r["measured_speedup"])) for r in rs])),
        "median_abs_log_err_mi355x_model": float(np.median([abs(np.log(r["predicted_speedup_mi355x_model"] /
                                                                       r["measured_speedup"])) for r in rs])),
    }
    out = {"summary": summ, "matrices": rs}
    if out_path:
        pathlib.Path(out_path).write_text(json.dumps(out, indent=1) + "\n")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mrpy.suite")
    ap.add_argument("--ids", default="1-75")
    ap.add_argument("--out", default="")
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--summarize", nargs="*", default=None, help="merge these run files instead of running")
    a = ap.parse_args(argv)
    if a.summarize is not None:
        print(json.dumps(summarize(a.summarize, a.out)["summary"], indent=1))
        return 0
    if not a.out:
        ap.error("--out is required")
    run(parse_ids(a.ids), a.out, a.width)
    return 0


if __name__ == "__main__":
    sys.exit(main())
