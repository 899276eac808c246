"""Batched SpMV / SpMM (mr_batch_*) on the device: one persistent fused launch
over many uneven matrices must give, job by job, the same bits as the
single-matrix MERBIT entry points (mr_spmv / mr_spmm) at the same tile width,
and lie within the stated tolerance of the FP64 oracle; hipGraph replays of
the fused launch must equal the eager launch on every replay."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_csr, rmat_csr

pytestmark = pytest.mark.gpu


def uneven_set(mr, seed=3):
    """Matrices of very different shapes: R-MAT, banded, ragged random, a
    matrix of empty rows, single-row matrices, a rectangular one."""
    rng = np.random.default_rng(seed)
    out = []
    for s in (10, 12, 9):
        rp, ci = rmat_csr(s, 8, seed=s)
        out.append((rp, ci, 1 << s))
    for n, bw, d in ((3000, 64, 12), (700, 8, 4), (20000, 300, 24)):
        rp, ci = mr.gen_banded(n, bw, d, seed=n)
        out.append((rp, ci, n))
    rp, ci = random_csr(rng, 1500, empty_frac=0.6, max_row=200)
    out.append((rp, ci, 1500))
    out.append((np.zeros(1001, np.int32), np.zeros(0, np.int32), 1000))
    out.append((np.array([0, 1], np.int32), np.array([0], np.int32), 1))
    out.append((np.array([0, 0], np.int32), np.zeros(0, np.int32), 1))
    rp, ci = random_csr(rng, 400, empty_frac=0.1, max_row=50, n_cols=2500)
    out.append((rp, ci, 2500))
    rows = [np.sort(rng.choice(9000, 8500, replace=False))] + [np.sort(rng.choice(9000, 3, replace=False))
                                                              for _ in range(30)]
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    out.append((rp, np.concatenate(rows).astype(np.int32), 9000))
    return out


def make_mats(ctx, mr, store, seed=5):
    rng = np.random.default_rng(seed)
    mats, host = [], []
    for rp, ci, nc in uneven_set(mr):
        val = rng.uniform(-1, 1, len(ci))
        mats.append(mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=store))
        host.append((rp, ci, val, nc))
    return mats, host


def vec_np(store):
    return np.float64 if store == 0 else np.float32


def io_for(ctx, mats, host, k, rng, store):
    vt = vec_np(store)
    xs, ys, xh = [], [], []
    for (rp, ci, val, nc) in host:
        x = rng.uniform(-1, 1, (nc, k)).astype(vt)
        xh.append(x)
        xs.append(ctx.upload(x))
        ys.append(ctx.empty((len(rp) - 1, k), vt))
    return xs, ys, xh


def single(ctx, mr, mat, xd, n_rows, k, vt):
    yd = ctx.empty((n_rows, k), vt)
    if k == 1:
        mat.spmv(xd, yd, mr.MERBIT)
    else:
        mat.spmm(k, xd, yd, mr.MERBIT)
    ctx.sync()
    return yd.to_host()


@pytest.mark.parametrize("sigma,nt,grid", [(8, 1, 0), (16, 0, 0), (32, 1, 0), (8, 0, 512), (16, 1, 97),
                                          (32, 0, 1)])
def test_batch_spmv_equals_single(ctx, mr, oracle, sigma, nt, grid):
    mats, host = make_mats(ctx, mr, mr.F32)
    rng = np.random.default_rng(sigma + grid)
    xs, ys, xh = io_for(ctx, mats, host, 1, rng, mr.F32)
    b = mr.Batch(ctx, mats, 1)
    b.configure(sigma, nt, grid)
    b.set_io(xs, ys)
    b.run()
    ctx.sync()
    for j, m in enumerate(mats):
        m.merbit_configure(sigma)
        m.merbit_options(1)   # the workgroup-tile kernel the batch runs
        rp, ci, val, nc = host[j]
        got = ys[j].to_host()
        want = single(ctx, mr, m, xs[j], len(rp) - 1, 1, np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"job {j}"
        yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float32).astype(np.float64),
                                       xh[j][:, 0].astype(np.float64))
        worst, l2, _ = oracle.spmv_parity(got[:, 0], yhat, absrow)
        assert worst <= 1e-5 and l2 <= 1e-5, (j, worst, l2)
    b.close()


def test_batch_spmv_f64_parity(ctx, mr, oracle):
    mats, host = make_mats(ctx, mr, mr.F64)
    rng = np.random.default_rng(2)
    xs, ys, xh = io_for(ctx, mats, host, 1, rng, mr.F64)
    b = mr.Batch(ctx, mats, 1)
    b.configure(16, 1, 0)
    b.set_io(xs, ys)
    b.run()
    ctx.sync()
    for j in range(len(mats)):
        rp, ci, val, nc = host[j]
        yhat, absrow = oracle.spmv_f64(rp, ci, val, xh[j][:, 0])
        worst, l2, _ = oracle.spmv_parity(ys[j].to_host()[:, 0], yhat, absrow)
        assert worst <= 1e-12 and l2 <= 1e-12, (j, worst, l2)
    b.close()


@pytest.mark.parametrize("k", [2, 3, 8, 32, 64])
def test_batch_spmm_equals_single(ctx, mr, oracle, k):
    mats, host = make_mats(ctx, mr, mr.F32)
    rng = np.random.default_rng(k)
    xs, ys, xh = io_for(ctx, mats, host, k, rng, mr.F32)
    b = mr.Batch(ctx, mats, k)
    b.configure(8, 0, 0 if k % 2 else 300)
    b.set_io(xs, ys)
    b.run()
    ctx.sync()
    for j, m in enumerate(mats):
        rp, ci, val, nc = host[j]
        got = ys[j].to_host()
        want = single(ctx, mr, m, xs[j], len(rp) - 1, k, np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"job {j}"
        for c in (0, k - 1):
            yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float32).astype(np.float64),
                                           xh[j][:, c].astype(np.float64))
            worst, l2, _ = oracle.spmv_parity(got[:, c], yhat, absrow)
            assert worst <= 1e-5 and l2 <= 1e-5, (j, c, worst, l2)
    b.close()


def test_batch_strided_io(ctx, mr):
    """ldx / ldy larger than k: only the first k columns are read / written."""
    mats, host = make_mats(ctx, mr, mr.F32)
    k, ld = 4, 7
    rng = np.random.default_rng(9)
    xs, ys, xw, sentinel = [], [], [], []
    for (rp, ci, val, nc) in host:
        x = rng.uniform(-1, 1, (nc, ld)).astype(np.float32)
        xw.append(x)
        xs.append(ctx.upload(x))
        y0 = np.full((len(rp) - 1, ld), 7.0, np.float32)
        sentinel.append(y0)
        ys.append(ctx.upload(y0))
    b = mr.Batch(ctx, mats, k)
    b.set_io(xs, ys, ldx=[ld] * len(mats), ldy=[ld] * len(mats))
    b.run()
    ctx.sync()
    for j, m in enumerate(mats):
        n = len(host[j][0]) - 1
        got = ys[j].to_host()
        assert np.all(got[:, k:] == 7.0)
        xd = ctx.upload(np.ascontiguousarray(xw[j][:, :k]))
        want = single(ctx, mr, m, xd, n, k, np.float32)
        assert np.array_equal(got[:, :k].view(np.uint32), want.view(np.uint32)), f"job {j}"
    b.close()


@pytest.mark.parametrize("k", [1, 8])
def test_batch_graph_replays(ctx, mr, k):
    """Capture once; replay several times with eager work and host syncs in
    between; every replay equals the eager result, and a replay after
    set_io reads the new buffers."""
    mats, host = make_mats(ctx, mr, mr.BF16)
    rng = np.random.default_rng(4)
    xs, ys, xh = io_for(ctx, mats, host, k, rng, mr.F32)
    st = ctx.new_stream()
    b = mr.Batch(ctx, mats, k)
    b.set_io(xs, ys, stream=st)
    b.run(st)
    ctx.sync(st)
    eager = [y.to_host() for y in ys]
    b.capture(st)
    for rep in range(4):
        for y in ys:
            y.zero()
        b.launch(st)
        ctx.sync(st)
        for j, y in enumerate(ys):
            assert np.array_equal(y.to_host().view(np.uint32), eager[j].view(np.uint32)), (rep, j)
        b.run(st)          # eager work between replays
        ctx.sync(st)
    xs2, ys2, _ = io_for(ctx, mats, host, k, np.random.default_rng(44), mr.F32)
    b.set_io(xs2, ys2, stream=st)
    b.launch(st)
    ctx.sync(st)
    got2 = [y.to_host() for y in ys2]
    b.configure(8, 1, 0)   # drops the graph
    b.set_io(xs2, ys2, stream=st)
    b.run(st)
    ctx.sync(st)
    for j, y in enumerate(ys2):
        assert np.array_equal(y.to_host().view(np.uint32), got2[j].view(np.uint32)), j
    b.close()


def test_batch_autotune_keeps_results(ctx, mr):
    mats, host = make_mats(ctx, mr, mr.F32)
    rng = np.random.default_rng(6)
    xs, ys, xh = io_for(ctx, mats, host, 1, rng, mr.F32)
    b = mr.Batch(ctx, mats, 1)
    b.set_io(xs, ys)
    cfg = b.autotune(2)
    assert cfg["sigma"] in (8, 16, 32) and cfg["nt"] in (0, 1) and cfg["grid"] >= 0 and cfg["us"] > 0
    b.run()
    ctx.sync()
    for j, m in enumerate(mats):
        m.merbit_configure(cfg["sigma"])
        m.merbit_options(1)
        want = single(ctx, mr, m, xs[j], len(host[j][0]) - 1, 1, np.float32)
        assert np.array_equal(ys[j].to_host().view(np.uint32), want.view(np.uint32)), j
    info = b.info()
    assert info["nnz"] == sum(len(h[1]) for h in host) and info["tiles"] > 0
    b.close()


def test_batch_errors(ctx, mr):
    mats, host = make_mats(ctx, mr, mr.F32)
    m64 = mr.Mat.from_csr(ctx, host[0][0], host[0][1], host[0][2], n_cols=host[0][3], store=mr.F64)
    with pytest.raises(mr.MRError) as e:
        mr.Batch(ctx, [mats[0], m64], 1)
    assert e.value.code == 5
    with pytest.raises(mr.MRError) as e:
        mr.Batch(ctx, mats, 65)
    assert e.value.code == 5
    b = mr.Batch(ctx, mats, 1)
    with pytest.raises(mr.MRError) as e:
        b.run()
    assert e.value.code == 1
    with pytest.raises(mr.MRError) as e:
        b.configure(12, 0, 0)
    assert e.value.code == 1
    b.close()
    b64 = mr.Batch(ctx, [m64], 1)
    with pytest.raises(mr.MRError) as e:
        b64.configure(32, 0, 0)
    assert e.value.code == 5
    b64.close()
    m64.close()
