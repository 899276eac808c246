"""The benchmark configurations at full size, through the C-ABI on the device,
checked against the oracle:

  config 2  R-MAT s22 ef16 seed 2, FP32: pGOrder gives a permutation whose
            first 24 batches (GPU loop, and the CPU phase of batches 10-23)
            equal the oracle's bit for bit; y of the workgroup-tile,
            lane-padded and column-panel MERBIT kernels within 1e-5 of the
            FP64 oracle (condition-aware and relative L2);
  config 3  27-point stencil 160^3 and R-MAT s21 ef32 with log-uniform
            values: the FP64 GPU product within 1e-12 of the oracle; FP32 /
            FP16 / BF16 storage with FP32 accumulation within 1e-5 of the
            oracle on the stored values, and within the stated bound of the
            FP64 product (rel-L2 <= 1e-6 / 1e-3 / 1e-2);
  config 4  the 64-matrix batched set, each matrix pGOrder-reordered, k = 1,
            8, 32 in one autotuned fused launch replayed from a hipGraph:
            every job bit-equal to its single-matrix launch and within 1e-5
            of the oracle;
  config 5  32 resident reordered BF16 matrices served with coalescing and
            uploads: a Poisson load with an R-MAT s18 upload runs to the end,
            and results of requests in a mix (large matrices, k = 1, 4, 16)
            lie within 1e-5 of the oracle on the stored values.

The generators are the benchmark's (mr_gen_*: counter-based, seeded per
configuration)."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_spmv(ctx, mat, x):
    xd = ctx.upload(x)
    yd = ctx.empty(mat.info()[0], x.dtype)
    mat.spmv(xd, yd)
    ctx.sync()
    return yd.to_host()


def parity(oracle, y, yhat, absrow, tol):
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= tol and l2 <= tol, (worst, l2)
    return worst, l2


def stored_f64(mat, dt, mr):
    _, _, vs = mat.export()
    if dt == mr.BF16:
        return (vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    if dt == mr.F16:
        return vs.view(np.float16).astype(np.float64)
    return np.asarray(vs, np.float64)


def test_config2_s22_pgorder_batches_and_spmv(ctx, mr, oracle):
    rp, ci = mr.gen_rmat(22, 16, (0.57, 0.19, 0.19), seed=2)
    n, nnz = len(rp) - 1, len(ci)
    val = mr.gen_values(102, nnz).astype(np.float32)
    x = mr.gen_values(103, n).astype(np.float32)
    W = mr.get_width(n, nnz)
    assert W == 1024
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    perm, st = mat.reorder_pgorder(0, 5)
    print(f"[config2] n={n} nnz={nnz} pGOrder {st['total_ms']:.0f} ms", flush=True)
    assert mr.check_perm(perm)
    # the first NB batches: oracle GPU loop, then order_on_cpu on batches
    # log2(W) .. NB-1 (each refines its own slice), against the device
    NB = 24
    tp, ti = oracle.transpose(n, n, rp, ci)
    pl = oracle.pgorder_gpu_loop(n, rp, ci, W, max_batches=NB, tp=tp, ti=ti)
    ps, pls, os_, ol = oracle.pgorder_schedule(n, W)
    L = int(math.sqrt(n) / 4)
    final = pl.copy()
    for b in range(int(math.log2(W)), NB):
        final[os_[b]:os_[b] + ol[b]] = oracle.order_on_cpu(n, rp, ci, pl[os_[b]:os_[b] + ol[b]], b, 5, L, tp, ti)
    upto = int(ps[NB - 1] + pls[NB - 1])
    inv = np.argsort(perm)   # position -> vertex
    assert np.array_equal(inv[:upto], final[:upto].astype(np.int64))
    # hubs above the 10,000 cap were placed by their start score alone
    assert np.max(np.diff(rp)[inv[:upto]]) > 10000
    del tp, ti
    # y = A x through the relabelled matrix, back in the original order
    yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float64), x.astype(np.float64))
    xp = np.empty_like(x)
    xp[perm] = x
    for pol in (1, 4, 9):
        mat.merbit_options(pol)
        y = run_spmv(ctx, mat, xp)[perm]
        worst, l2 = parity(oracle, y, yhat, absrow, 1e-5)
        print(f"[config2] policy {pol}: max {worst:.2e} rel-L2 {l2:.2e}", flush=True)
    mat.close()


@pytest.mark.parametrize("name", ["stencil27_160", "rmat_s21_ef32"])
def test_config3_precision_sweep(ctx, mr, oracle, name):
    if name == "stencil27_160":
        rp, ci = mr.gen_stencil27(160, 160, 160)
    else:
        rp, ci = mr.gen_rmat(21, 32, (0.57, 0.19, 0.19), seed=21)
    n, nnz = len(rp) - 1, len(ci)
    val = mr.gen_values(303, nnz, kind=1, lo=1e-3, hi=1.0)
    x = mr.gen_values(304, n)
    yhat, absrow = oracle.spmv_f64(rp, ci, val, x)
    m64 = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F64)
    y64 = run_spmv(ctx, m64, x)
    parity(oracle, y64, yhat, absrow, 1e-12)
    m64.close()
    nrm = float(np.linalg.norm(yhat))
    bound = {mr.F32: 1e-6, mr.F16: 1e-3, mr.BF16: 1e-2}
    for dt in (mr.F32, mr.F16, mr.BF16):
        m = mr.Mat.from_csr(ctx, rp, ci, val, store=dt)
        vs = stored_f64(m, dt, mr)
        ys, ars = oracle.spmv_f64(rp, ci, vs, x.astype(np.float32).astype(np.float64))
        for pol in (4, 8):
            m.merbit_options(pol)
            y = run_spmv(ctx, m, x.astype(np.float32))
            parity(oracle, y, ys, ars, 1e-5)
            d = y.astype(np.float64) - yhat
            rel = float(np.linalg.norm(d)) / nrm
            print(f"[config3] {name} dtype {dt} policy {pol}: max-abs {np.max(np.abs(d)):.2e} "
                  f"rel-L2 {rel:.2e} vs FP64", flush=True)
            assert rel <= bound[dt], (dt, pol, rel)
        m.close()


def test_config4_batched_set(ctx, mr, oracle):
    shapes = mr.gen_batched_set(3, 64)
    mats, host = [], []
    for i, (rp, ci) in enumerate(shapes):
        val = mr.gen_values(3000 + i, len(ci)).astype(np.float32)
        m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
        perm, _ = m.reorder_pgorder(0, 5)
        assert mr.check_perm(perm)
        prp, pci, pv = m.export()
        mats.append(m)
        host.append((prp, pci, np.asarray(pv, np.float64)))
    print(f"[config4] {len(mats)} matrices, {sum(len(h[1]) for h in host)} nnz", flush=True)
    st = ctx.new_stream()
    for k in (1, 8, 32):
        rng = np.random.default_rng(40 + k)
        xs, ys, xh = [], [], []
        for prp, pci, _ in host:
            x = rng.uniform(-1, 1, (len(prp) - 1, k)).astype(np.float32)
            xh.append(x)
            xs.append(ctx.upload(x))
            ys.append(ctx.empty((len(prp) - 1, k), np.float32))
        b = mr.Batch(ctx, mats, k)
        b.set_io(xs, ys, stream=st)
        tune = b.autotune(3)
        b.set_io(xs, ys, stream=st)
        b.capture(st)
        b.launch(st)
        ctx.sync(st)
        got = [y.to_host() for y in ys]
        for j, m in enumerate(mats):
            prp, pci, pv = host[j]
            nr = len(prp) - 1
            m.merbit_configure(tune["sigma"])
            m.merbit_options(1)   # the workgroup-tile kernels the batch runs
            yd = ctx.empty((nr, k), np.float32)
            if k == 1:
                m.spmv(xs[j], yd)
            else:
                m.spmm(k, xs[j], yd)
            ctx.sync()
            assert np.array_equal(got[j].view(np.uint32), yd.to_host().view(np.uint32)), (k, j)
            for c in sorted({0, k - 1}):
                yh, ar = oracle.spmv_f64(prp, pci, pv, xh[j][:, c].astype(np.float64))
                parity(oracle, got[j][:, c], yh, ar, 1e-5)
        print(f"[config4] k={k}: tune {tune}, all jobs equal their single launches", flush=True)
        b.close()
    for m in mats:
        m.close()


def test_config5_serving_set(ctx, mr, oracle):
    resident, host = [], []
    for scale, ef, seed in ((22, 16, 2), (21, 32, 21)):
        rp, ci = mr.gen_rmat(scale, ef, (0.57, 0.19, 0.19), seed=seed)
        m = mr.Mat.from_csr(ctx, rp, ci, mr.gen_values(500 + scale, len(ci)), store=mr.BF16)
        m.reorder_pgorder(0, 5)
        resident.append(m)
    for i, (rp, ci) in enumerate(mr.gen_batched_set(3, 64)[:30]):
        m = mr.Mat.from_csr(ctx, rp, ci, mr.gen_values(3000 + i, len(ci)), store=mr.BF16)
        m.reorder_pgorder(0, 5)
        resident.append(m)
    for m in resident:
        prp, pci, pv = m.export()
        host.append((prp, pci, (pv.astype(np.uint32) << 16).view(np.float32).astype(np.float64), len(prp) - 1))
    # the load as the benchmark runs it (results hashed on the device only)
    srv = mr.Server(ctx, slots=4, max_jobs=256, coalesce=1)
    for m in resident:
        srv.add(m)
    print("[config5] 32 resident BF16 matrices", flush=True)
    rep = srv.run_load(rate=5000.0, n_requests=2000, seed=5, ks=(1, 4, 16), upload_every=1000,
                       upload_scale=18, upload_edge_factor=16)
    print(f"[config5] load: {rep}", flush=True)
    assert rep["completed"] + rep["rejected"] + rep["timed_out"] == 2000 - rep["uploads_admitted"] - \
        rep["uploads_rejected"]
    assert rep["uploads_admitted"] == 2 and rep["completed"] > 0 and rep["p50_us"] > 0
    srv.close()
    # a burst of mixed requests (coalesced and batched with each other),
    # results copied out
    srv = mr.Server(ctx, slots=4, max_jobs=256, coalesce=1, verify=2)
    ids = [srv.add(m) for m in resident]
    rng = np.random.default_rng(55)
    mix = [(0, 16), (1, 4), (0, 1), (2, 16), (1, 1), (5, 4), (0, 4), (17, 16), (1, 16), (29, 1)]
    mix += [(int(rng.integers(0, 32)), int(rng.choice([1, 4, 16]))) for _ in range(30)]
    reqs = [srv.submit(ids[m], k, 9000 + i) for i, (m, k) in enumerate(mix)]
    print(f"[config5] burst of {len(reqs)} requests submitted", flush=True)
    for i, r in enumerate(reqs):
        st, _, _ = srv.wait(r)
        assert st == 0, i
        m, k = mix[i]
        if i >= 10 and i % 5:
            continue
        prp, pci, pv, n = host[m]
        y = srv.result(r, n, k)
        X = mr.serve_rhs(9000 + i, n, k).astype(np.float64)
        for c in sorted({0, k - 1}):
            yh, ar = oracle.spmv_f64(prp, pci, pv, X[:, c])
            parity(oracle, y[:, c], yh, ar, 1e-5)
    srv.close()
    for m in resident:
        m.close()
