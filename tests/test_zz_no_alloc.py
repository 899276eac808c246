"""mr_spmv_noalloc: the MERBIT SpMV whose carries live in the library's static
device scratch array.  Through the C-ABI its results equal mr_spmv_ws on the
same plan bit for bit, on two calls in a row where the second matrix is
smaller (the scratch keeps the first call's contents), for the lane-padded
x-cache plan, the workgroup-tile plan and FP64."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rmat_csr

pytestmark = pytest.mark.gpu


def _ws_result(ctx, mat, xd, n):
    ws = ctx.empty(mat.spmv_ws_bytes(), np.uint8)
    yd = ctx.empty(n, xd.dtype)
    mat.spmv_ws(xd, yd, ws)
    ctx.sync()
    return yd.to_host()


def _noalloc_result(ctx, mat, xd, n):
    yd = ctx.empty(n, xd.dtype)
    mat.spmv_noalloc(xd, yd)
    ctx.sync()
    return yd.to_host()


@pytest.mark.parametrize("store,policy", [("f32", 4), ("bf16", 5), ("f32", 1), ("f64", 0)])
def test_noalloc_equals_workspace_entry_point(ctx, mr, oracle, store, policy):
    dt = {"f32": mr.F32, "bf16": mr.BF16, "f64": mr.F64}[store]
    vt = np.float64 if store == "f64" else np.float32
    mats = []
    for scale, seed in ((14, 41), (11, 42)):      # the second call is smaller
        rp, ci = rmat_csr(scale, 16, seed=seed)
        n = len(rp) - 1
        rng = np.random.default_rng(seed)
        val = rng.uniform(-1, 1, len(ci))
        x = rng.uniform(-1, 1, n).astype(vt)
        mat = mr.Mat.from_csr(ctx, rp, ci, val, store=dt)
        mat.merbit_options(policy)
        mats.append((mat, rp, ci, val, x, n))
    # the workspace entry point first (it builds both plans), then the
    # no-allocation one on both matrices in a row, the smaller one second
    wants = []
    for mat, rp, ci, val, x, n in mats:
        xd = ctx.upload(x)
        wants.append((xd, _ws_result(ctx, mat, xd, n)))
    for (mat, rp, ci, val, x, n), (xd, want) in zip(mats, wants):
        print(f"[noalloc] store={store} policy={policy} n={n} nnz={len(ci)}", flush=True)
        got = _noalloc_result(ctx, mat, xd, n)
        assert got.dtype == want.dtype
        assert np.array_equal(got.view(np.uint64 if store == "f64" else np.uint32),
                              want.view(np.uint64 if store == "f64" else np.uint32))
        # and the product itself, against the oracle on the stored values
        _, _, vs = mat.export()
        if store == "bf16":
            v64 = (vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            v64 = np.asarray(vs, np.float64)
        yhat, absrow = oracle.spmv_f64(rp, ci, v64, x.astype(np.float64))
        worst, l2, _ = oracle.spmv_parity(got, yhat, absrow)
        tol = 1e-12 if store == "f64" else 1e-5
        assert worst <= tol and l2 <= tol, (worst, l2)
    for mat, *_ in mats:
        mat.close()


def test_noalloc_refuses_an_unbuilt_plan(ctx, mr):
    rp, ci = rmat_csr(10, 8, seed=43)
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F32)
    mat.merbit_options(4)
    xd = ctx.upload(np.ones(n, np.float32))
    yd = ctx.empty(n, np.float32)
    with pytest.raises(mr.MRError) as e:
        mat.spmv_noalloc(xd, yd)
    assert e.value.code == 1      # MR_ERR_INVALID_ARG
    mat.spmv(xd, yd)              # builds the plan; now it runs
    mat.spmv_noalloc(xd, yd)
    ctx.sync()
    mat.close()
