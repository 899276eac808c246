"""Refit of the RD headroom predictor on MI355X timings.

The reference fits its predictor offline (ridge/train.py:596-665).  The
pipeline is a median imputer, a standard scaler and Ridge(alpha=1) on
log(speedup).  The inputs are ten sampled RD bin probabilities.  The training
orders are local_swap perturbations (src/class.cu:236-301) at the nine
(rho, near, window) settings of src/main.cu:27-29.  The exported constants
(include/model.cuh) were fitted on sm_89 timings.  This module repeats the
method with MI355X timings.

``collect`` runs on the GPU.  It builds a corpus of matrices and reorders each
one with pGOrder.  It then records, for other orders of the same matrix, the
sampled RD features and the measured SpMV times.  The other orders are the
generator's order, a random order, degree order, RCM, and local_swap
perturbations of the pGOrder order.  The target is
t(order) / t(pGOrder order): the time pGOrder would save on that order.
Features are recorded under two sets of bin edges: the reference's
(src/class.cu:1265-1277) and edges placed at MI355X cache sizes.

``fit`` runs on the CPU.  It applies the reference's pipeline to each
(kernel, edge set) pair and cross-validates with the folds grouped by matrix.
It picks the edge set with the lower cross-validated error and writes the
coefficients as a C++ header for ``mr_predict``.

Usage, from the package directory (``matrix-reorder_amd/``):
    python -m mrpy.refit collect --out /tmp/mi355x_rd_train.csv
    python -m mrpy.refit fit --csv model/mi355x_rd_train.csv \\
        --header csrc/model_mi355x.hpp --report model/mi355x_rd_fit.json
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import pathlib
import sys
import time

import numpy as np

# ---------------------------------------------------------------- constants
# (rho, near, window) of the reference's perturbation sweep (src/main.cu:27-29)
PARA = ((0.10, 64, 256), (0.10, 256, 1024), (0.20, 1024, 4096), (0.30, 4096, 16384), (0.30, 16384, 65536),
        (0.40, 65536, 262144), (0.50, 262144, 1048576), (0.70, 1048576, 4194304), (0.90, 0, 10000000))

# Bin edges in 128-byte sectors (32 FP32 elements: one MI355X cache line).
# 256 lines = the 32 KB vector L1 of a CU; 1024 = 128 KB; 8192 = 1 MB;
# 32768 = 4 MB, the L2 of one XCD; 131072 = 16 MB, a quarter of all L2;
# 262144 = 32 MB, the L2 of all eight XCDs.  The cap is the last edge.
MI355X_EDGES = (1, 32, 112, 256, 1024, 8192, 32768, 131072, 262144)
FEATURES = 10


def llround(v: float) -> int:
    """std::llround for the non-negative values local_swap passes it."""
    return int(math.floor(v + 0.5))


def local_swap(n: int, rho: float, near: int, window: int, seed: int) -> np.ndarray:
    """Controlled perturbation of the identity order (src/class.cu:236-301).

    The steps follow the reference:
      1. take the first round(rho * n) indices of a shuffled iota;
      2. for each index i, take the left interval [i-window, i-near] and the
         right interval [i+near, i+window], both clipped to [0, n);
      3. pick a side weighted by interval length, then a uniform j in that
         interval;
      4. swap entries i and j of the permutation, in list order.
    The random numbers come from numpy's PCG64 rather than std::mt19937_64,
    so the permutations are not the reference's bit for bit.  The draws
    depend only on i, so they are made up front; the swaps are applied in
    sequence."""
    rng = np.random.default_rng(seed)
    perm = np.arange(n, dtype=np.int64)
    k = llround(rho * n)
    if n == 0 or k == 0:
        return perm.astype(np.int32)
    il = rng.permutation(n)[:k].astype(np.int64)
    l1 = np.maximum(0, il - window)
    r1 = np.minimum(n - 1, il - near)
    l2 = np.maximum(0, il + near)
    r2 = np.minimum(n - 1, il + window)
    len1 = np.where(l1 <= r1, r1 - l1 + 1, 0)
    len2 = np.where(l2 <= r2, r2 - l2 + 1, 0)
    tot = len1 + len2
    u = rng.random(k)
    left = np.where((len1 > 0) & (len2 > 0), u < len1 / np.maximum(tot, 1), len1 > 0)
    lo = np.where(left, l1, l2)
    span = np.where(left, len1, len2)
    j = lo + np.floor(rng.random(k) * np.maximum(span, 1)).astype(np.int64)
    j = np.minimum(j, lo + np.maximum(span, 1) - 1)
    ok = (tot > 0) & (j != il)
    p = perm.tolist()
    for a, b in zip(il[ok].tolist(), j[ok].tolist()):
        p[a], p[b] = p[b], p[a]
    return np.asarray(p, dtype=np.int32)


# ---------------------------------------------------------------- corpus
def corpus():
    """Base matrices of the training set: (name, family, spec).  Sizes run
    from 2^17 to 2^24 rows, so x runs from 0.5 MB (resident in L2) to 64 MB
    (twice the L2 of all eight XCDs)."""
    out = []
    skews = {"a57": (0.57, 0.19, 0.19), "a45": (0.45, 0.15, 0.15), "a65": (0.65, 0.15, 0.15)}
    for scale in (17, 19, 21):
        for ef in (4, 16):
            for tag, abc in skews.items():
                out.append((f"rmat_s{scale}_ef{ef}_{tag}", "rmat",
                            dict(scale=scale, ef=ef, abc=abc, sym=False)))
    for tag, abc in skews.items():
        out.append((f"rmat_sym_s20_ef8_{tag}", "rmat", dict(scale=20, ef=8, abc=abc, sym=True)))
    for scale in (18, 20):
        out.append((f"uniform_s{scale}_ef8", "rmat", dict(scale=scale, ef=8, abc=(0.25, 0.25, 0.25), sym=False)))
    for n in (1 << 18, 1 << 21):
        for bw in (32, 2048, 131072):
            out.append((f"banded_n{n}_bw{bw}", "banded", dict(n=n, bw=bw, per_row=12)))
    for g in (48, 96, 128):
        out.append((f"stencil27_{g}", "stencil", dict(g=g)))
    # the sizes of the single-large-SpMV and precision workloads (other seeds)
    out.append(("rmat_s22_ef16_a57", "rmat", dict(scale=22, ef=16, abc=skews["a57"], sym=False)))
    out.append(("rmat_s22_ef8_a65", "rmat", dict(scale=22, ef=8, abc=skews["a65"], sym=False)))
    out.append(("banded_n4194304_bw2048", "banded", dict(n=1 << 22, bw=2048, per_row=12)))
    out.append(("stencil27_160", "stencil", dict(g=160)))
    # the scales of the reference's published R-MAT sweep (s24, symmetric)
    out.append(("rmat_s23_ef8_a57", "rmat", dict(scale=23, ef=8, abc=skews["a57"], sym=False)))
    out.append(("rmat_s23_ef16_a65", "rmat", dict(scale=23, ef=16, abc=skews["a65"], sym=False)))
    out.append(("rmat_sym_s24_ef4_a57", "rmat", dict(scale=24, ef=4, abc=skews["a57"], sym=True)))
    out.append(("rmat_sym_s24_ef4_a45", "rmat", dict(scale=24, ef=4, abc=skews["a45"], sym=True)))
    out.append(("rmat_sym_s24_ef8_a65", "rmat", dict(scale=24, ef=8, abc=skews["a65"], sym=True)))
    return out


def sweep_corpus():
    """The reference's published R-MAT s24 sweep (config/config3.txt ids
    1-75, symmetric, the same generator as mrpy/suite.py with seed = id)."""
    pub = json.loads((pathlib.Path(__file__).resolve().parent.parent / "suite" / "rmat_s24_published.json")
                     .read_text())["matrices"]
    return [(f"sweep_s24_id{m['id']}", "rmat", dict(scale=m["scale"], ef=m["ef"], abc=(m["a"], m["b"], m["b"]),
                                                    sym=True, seed=m["id"]))
            for m in pub]


def build(mr, family: str, spec: dict, seed: int):
    if family == "rmat":
        return mr.gen_rmat(spec["scale"], spec["ef"], spec["abc"], seed=spec.get("seed", seed), symmetric=spec["sym"])
    if family == "banded":
        return mr.gen_banded(spec["n"], spec["bw"], spec["per_row"], seed)
    if family == "stencil":
        g = spec["g"]
        return mr.gen_stencil27(g, g, g)
    raise ValueError(family)


# ---------------------------------------------------------------- collection
def _time(ctx, mat, xd, yd, kernel, reps, warmup):
    s = ctx.stream()
    for _ in range(warmup):
        mat.spmv(xd, yd, kernel, s)
    ctx.sync(s)
    best = math.inf
    for _ in range(3):           # the best of three averages
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0, s)
        for _ in range(reps):
            mat.spmv(xd, yd, kernel, s)
        ctx.record(e1, s)
        ctx.sync(s)
        best = min(best, ctx.elapsed(e0, e1) / reps)
    return best * 1e3            # microseconds


def collect(out_path: str, limit: int = 0, reps: int = 20, skip: int = 0, sweep: bool = False) -> None:
    """Measure the corpus (or, with sweep, the published s24 sweep: the
    generated and the pGOrder orders of each matrix only)."""
    import mrpy as mr

    ctx = mr.Ctx(0)
    p_ref = mr.rd_params()
    p_mi = mr.rd_params(bin_edges=MI355X_EDGES, rd_cap=MI355X_EDGES[-1])
    cols = (["graph", "family", "n", "nnz", "variant", "t_merbit_us", "t_csr_us", "t_merbit_pg_us", "t_csr_pg_us",
             "speedup_merbit", "speedup_csr", "pred_ref_merbit", "pred_ref_csr"]
            + [f"r{i + 1}" for i in range(FEATURES)] + [f"q{i + 1}" for i in range(FEATURES)])
    path = pathlib.Path(out_path)
    path.parent.mkdir(parents=True, exist_ok=True)
    items = list(enumerate(sweep_corpus() if sweep else corpus()))[skip:]
    if limit:
        items = items[:limit]
    t_start = time.perf_counter()
    with path.open("w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(cols)
        for k, (gi, (name, family, spec)) in enumerate(items):
            seed = 7001 + gi
            rp, ci = build(mr, family, spec, seed)
            n, nnz = len(rp) - 1, len(ci)
            val = mr.gen_values(seed + 1, nnz).astype(np.float32)
            x = mr.gen_values(seed + 2, n).astype(np.float32)
            xd = ctx.upload(x)
            yd = ctx.empty(n, np.float32)
            # the pGOrder order: reference point of the speedups
            m0 = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
            # pGOrder at get_width's width, as the reference calls gorder6
            pg, _ = m0.reorder_pgorder(0, 5)
            grp, gci, gval = m0.export()
            m0.close()
            # other orders, each given as a relabelling of the original
            # matrix (perm[old] = new)
            rng = np.random.default_rng(seed + 3)
            variants = [("pgorder", pg), ("original", None)]
            if not sweep:
                variants.append(("random", rng.permutation(n).astype(np.int32)))
                deg = np.diff(rp)
                order = np.argsort(-deg, kind="stable")
                dperm = np.empty(n, np.int32)
                dperm[order] = np.arange(n, dtype=np.int32)
                variants.append(("degree", dperm))
                m1 = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
                rperm, _ = m1.reorder_rcm()
                m1.close()
                variants.append(("rcm", rperm))
                for pi, (rho, near, win) in enumerate(PARA):
                    ls = local_swap(n, rho, near, win, seed * 16 + pi)
                    # perturb the pGOrder order: old -> pGOrder position -> swapped
                    inv_ls = np.empty(n, np.int32)
                    inv_ls[ls] = np.arange(n, dtype=np.int32)
                    variants.append((f"swap_{rho}_{near}_{win}", inv_ls[pg]))
            t_pg = None
            for vname, perm in variants:
                m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
                if perm is not None:
                    m.permute(perm)
                xv = ctx.empty(n, np.float32)
                m.permute_vec(0, xd, xv)
                tm = _time(ctx, m, xv, yd, mr.MERBIT, reps, 5)
                tc = _time(ctx, m, xv, yd, mr.CSR_VECTOR, reps, 3)
                fr, _ = m.rd_features(mr.RD_SAMPLED, p_ref)
                fq, _ = m.rd_features(mr.RD_SAMPLED, p_mi)
                m.close()
                xv.free()
                if vname == "pgorder":
                    t_pg = (tm, tc)
                row = [name, family, n, nnz, vname, f"{tm:.3f}", f"{tc:.3f}", f"{t_pg[0]:.3f}", f"{t_pg[1]:.3f}",
                       f"{tm / t_pg[0]:.6f}", f"{tc / t_pg[1]:.6f}",
                       f"{mr.predict(fr, mr.MODEL_MERBIT):.6f}", f"{mr.predict(fr, mr.MODEL_CSR):.6f}"]
                row += [f"{v:.9g}" for v in fr] + [f"{v:.9g}" for v in fq]
                w.writerow(row)
            fh.flush()
            xd.free()
            yd.free()
            print(f"[refit] {k + 1}/{len(items)} {name}: n={n} nnz={nnz} pgorder {t_pg[0]:.1f} us "
                  f"({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
    ctx.close()


# ---------------------------------------------------------------- fitting
def load_rows(csv_path: str):
    with open(csv_path, newline="") as fh:
        return list(csv.DictReader(fh))


def _pipeline(alpha: float):
    from sklearn.impute import SimpleImputer
    from sklearn.linear_model import Ridge
    from sklearn.pipeline import Pipeline
    from sklearn.preprocessing import StandardScaler

    return Pipeline([("imputer", SimpleImputer(strategy="median")), ("scaler", StandardScaler()),
                     ("model", Ridge(alpha=alpha))])


def raw_coefficients(pipe):
    """Raw-feature form exp(b + sum w_i r_i) of a fitted pipeline, as the
    reference exports it (ridge/train.py:624-665)."""
    sc = pipe.named_steps["scaler"]
    rg = pipe.named_steps["model"]
    coef_scaled = np.asarray(rg.coef_, np.float64).reshape(-1)
    coef = coef_scaled / sc.scale_
    b = float(rg.intercept_) - float(np.dot(coef_scaled, sc.mean_ / sc.scale_))
    return b, coef


def _metrics(y_log: np.ndarray, p_log: np.ndarray) -> dict:
    from scipy.stats import spearmanr

    d = p_log - y_log
    ss = float(np.sum((y_log - y_log.mean()) ** 2))
    return {"rows": int(len(y_log)), "mae_log": float(np.mean(np.abs(d))),
            "rmse_log": float(np.sqrt(np.mean(d * d))),
            "r2_log": float(1.0 - np.sum(d * d) / ss) if ss > 0 else 0.0,
            "median_abs_rel_err": float(np.median(np.abs(np.exp(d) - 1.0))),
            "spearman": float(spearmanr(y_log, p_log).statistic)}


def fit(csv_path: str, alpha: float = 1.0, folds: int = 5):
    """Fit each (kernel, edge set) pair; return the report and the chosen
    models."""
    from sklearn.model_selection import GroupKFold

    rows = load_rows(csv_path)
    groups = np.array([r["graph"] for r in rows])
    report = {"csv": str(csv_path), "rows": len(rows), "graphs": int(len(set(groups))), "alpha": alpha,
              "folds": folds, "target": "log(t(order) / t(pGOrder order))", "kernels": {}}
    chosen = {}
    for kernel in ("merbit", "csr"):
        y = np.log(np.array([float(r[f"speedup_{kernel}"]) for r in rows]))
        ref_pred = np.log(np.array([float(r[f"pred_ref_{kernel}"]) for r in rows]))
        kr = {"reference_model_sm89": _metrics(y, ref_pred)}
        best = None
        for fs, prefix, edges in (("reference_edges", "r", None), ("mi355x_edges", "q", MI355X_EDGES)):
            X = np.array([[float(r[f"{prefix}{i + 1}"]) for i in range(FEATURES)] for r in rows])
            cv = np.zeros(len(rows))
            for tr, te in GroupKFold(n_splits=folds).split(X, y, groups):
                p = _pipeline(alpha).fit(X[tr], y[tr])
                cv[te] = p.predict(X[te])
            full = _pipeline(alpha).fit(X, y)
            b, coef = raw_coefficients(full)
            raw = b + X @ coef
            kr[fs] = {"cv": _metrics(y, cv), "train": _metrics(y, raw),
                      "max_raw_vs_pipeline": float(np.max(np.abs(raw - full.predict(X)))),
                      "intercept": b, "coef": coef.tolist()}
            if best is None or kr[fs]["cv"]["mae_log"] < kr[best[0]]["cv"]["mae_log"]:
                best = (fs, edges, b, coef)
        kr["chosen"] = best[0]
        chosen[kernel] = best
        report["kernels"][kernel] = kr
    return report, chosen


def header_text(chosen: dict, csv_name: str) -> str:
    """C++ constants of the MI355X models.  Both kernels must share one edge
    set, since mr_locality_predict computes one feature vector."""
    fs = chosen["merbit"][0]
    edges = chosen["merbit"][1] or (1, 112, 512, 18432, 36864, 73728, 147456, 294912, 589824)

    def arr(v):
        return ",\n    ".join(f"{float(x):.17e}" for x in v)

    mb, mw = chosen["merbit"][2], chosen["merbit"][3]
    cb, cw = chosen["csr"][2], chosen["csr"][3]
    return f"""// model_mi355x.hpp -- RD headroom models fitted on MI355X timings.
// Written by mrpy/refit.py (fit) from {csv_name}; do not edit by hand.
// speedup = exp(b + sum_i w[i] * r[i]) over the ten sampled RD bin
// probabilities computed with kEdges; target t(order) / t(pGOrder order).
// Feature set: {fs}.
#pragma once
#include <cstdint>

namespace mr {{
namespace mi355x {{

constexpr int32_t kEdges[9] = {{{", ".join(str(int(e)) for e in edges)}}};
constexpr int32_t kRdCap = {int(edges[-1])};
constexpr double kMerbitB = {mb:.17e};
constexpr double kMerbitW[10] = {{
    {arr(mw)}}};
constexpr double kCsrB = {cb:.17e};
constexpr double kCsrW[10] = {{
    {arr(cw)}}};

}}  // namespace mi355x
}}  // namespace mr
"""


def fit_shared(csv_path: str, alpha: float = 1.0, folds: int = 5):
    """Fit, then refit the CSR model on the MERBIT model's edge set so that
    both share one feature vector."""
    report, chosen = fit(csv_path, alpha, folds)
    fs = chosen["merbit"][0]
    kr = report["kernels"]["csr"]
    chosen["csr"] = (fs, chosen["merbit"][1], kr[fs]["intercept"], np.asarray(kr[fs]["coef"]))
    kr["used"] = fs
    report["kernels"]["merbit"]["used"] = fs
    return report, chosen


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mrpy.refit")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect", help="measure the training set on the GPU")
    c.add_argument("--out", required=True)
    c.add_argument("--limit", type=int, default=0, help="first N matrices of the corpus only")
    c.add_argument("--reps", type=int, default=20)
    c.add_argument("--skip", type=int, default=0, help="start at this corpus index (seeds stay per index)")
    c.add_argument("--sweep", action="store_true",
                   help="the published R-MAT s24 sweep (generated and pGOrder orders) instead of the corpus")
    f = sub.add_parser("fit", help="fit the models on the CPU")
    f.add_argument("--csv", required=True)
    f.add_argument("--header", default="")
    f.add_argument("--report", default="")
    f.add_argument("--alpha", type=float, default=1.0)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        collect(a.out, a.limit, a.reps, a.skip, a.sweep)
        return 0
    report, chosen = fit_shared(a.csv, a.alpha)
    if a.header:
        pathlib.Path(a.header).write_text(header_text(chosen, pathlib.Path(a.csv).name))
    text = json.dumps(report, indent=1)
    if a.report:
        pathlib.Path(a.report).write_text(text + "\n")
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
