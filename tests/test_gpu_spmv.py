"""SpMV/SpMM on the device through the C-ABI against the FP64 oracle:
MERBIT (all tile widths), CSR scalar and CSR vector, FP64 at 1e-12
(condition-aware: max_i |y_i - yhat_i| / sum_j |a_ij x_j|, and relative L2),
FP32 at 1e-5; MERBIT metadata bit-exact; results bit-reproducible."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rmat_csr, spmv_cases

pytestmark = pytest.mark.gpu

CASES = spmv_cases()
IDS = [c[0] for c in CASES]


def spmv(ctx, mr, mat, x, kernel):
    xd = ctx.upload(x)
    yd = ctx.empty(mat.info()[0], x.dtype)
    mat.spmv(xd, yd, kernel)
    ctx.sync()
    return yd.to_host()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("kernel", ["merbit", "csr_scalar", "csr_vector"])
def test_spmv_f64_parity(ctx, mr, oracle, case, kernel):
    name, rp, ci, nc = case
    rng = np.random.default_rng(hash(name) % 1000)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F64)
    k = {"merbit": mr.MERBIT, "csr_scalar": mr.CSR_SCALAR, "csr_vector": mr.CSR_VECTOR}[kernel]
    y = spmv(ctx, mr, mat, x, k)
    yhat, absrow = oracle.spmv_f64(rp, ci, val, x)
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-12 and l2 <= 1e-12, (worst, l2)
    mat.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("sigma", [8, 16, 32])
def test_merbit_f32_all_sigmas(ctx, mr, oracle, case, sigma):
    name, rp, ci, nc = case
    rng = np.random.default_rng(11)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F32)
    mat.merbit_configure(sigma)
    y = spmv(ctx, mr, mat, x.astype(np.float32), mr.MERBIT)
    yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float32).astype(np.float64),
                                   x.astype(np.float32).astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    mat.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_merbit_metadata_bit_exact(ctx, mr, oracle, case):
    name, rp, ci, nc = case
    mat = mr.Mat.from_csr(ctx, rp, ci, None, n_cols=nc, store=mr.F32)
    for sigma in (8, 16, 32):
        mat.merbit_configure(sigma)
        flags, bx, by, info = mat.merbit_export()
        n, nnz = len(rp) - 1, len(ci)
        T = 256 * sigma
        assert info["items"] == n + nnz and info["n_blocks"] == -(-(n + nnz) // T)
        bits = np.zeros(len(flags) * 32, np.uint8)
        bits[(rp[1:].astype(np.int64) + np.arange(n))] = 1        # item rp[r+1] + r ends row r
        want = np.packbits(bits.reshape(-1, 32)[:, ::-1], axis=1, bitorder="big").view(">u4").ravel()
        assert np.array_equal(flags, want.astype(np.uint32))
        for b in range(info["n_blocks"] + 1):
            d = min(b * T, n + nnz)
            assert (bx[b], by[b]) == oracle.merge_search(rp, nnz, d)
    mat.close()


def test_merbit_bit_reproducible_and_workspace(ctx, mr):
    name, rp, ci, nc = CASES[0]
    rng = np.random.default_rng(3)
    mat = mr.Mat.from_csr(ctx, rp, ci, rng.uniform(-1, 1, len(ci)), n_cols=nc, store=mr.F32)
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    y1 = spmv(ctx, mr, mat, x, mr.MERBIT)
    y2 = spmv(ctx, mr, mat, x, mr.MERBIT)
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))
    ws = ctx.empty(mat.spmv_ws_bytes(), np.uint8)
    xd = ctx.upload(x)
    yd = ctx.empty(len(y1), np.float32)
    mat.spmv_ws(xd, yd, ws)
    ctx.sync()
    assert np.array_equal(yd.to_host().view(np.uint32), y1.view(np.uint32))
    mat.close()


@pytest.mark.parametrize("k", [2, 4, 8, 32])
@pytest.mark.parametrize("store", ["f64", "f32"])
def test_spmm(ctx, mr, oracle, k, store):
    name, rp, ci, nc = CASES[0]
    rng = np.random.default_rng(k)
    val = rng.uniform(-1, 1, len(ci))
    X = rng.uniform(-1, 1, (nc, k))
    dt = np.float64 if store == "f64" else np.float32
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F64 if store == "f64" else mr.F32)
    Xd = ctx.upload(X.astype(dt))
    Yd = ctx.empty((len(rp) - 1, k), dt)
    mat.spmm(k, Xd, Yd, mr.CSR_VECTOR)
    ctx.sync()
    Y = Yd.to_host()
    Yh = oracle.spmm_f64(rp, ci, val.astype(dt).astype(np.float64), X.astype(dt).astype(np.float64))
    tol = 1e-12 if store == "f64" else 1e-5
    assert np.max(np.abs(Y - Yh)) <= tol * max(1.0, np.max(np.abs(Yh))) * 16
    mat.close()


def test_value_conversion_rounding(ctx, mr):
    # values are rounded once from FP64 to the storage type (round to nearest even)
    rng = np.random.default_rng(5)
    n = 4096
    v = np.concatenate([rng.uniform(-1, 1, n - 6), [1 + 2.0 ** -8, 1 + 3 * 2.0 ** -9, 65504.0, 1e-8, -0.0, 1 + 2.0 ** -8 + 2.0 ** -30]])
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.zeros(n, np.int32)
    m64 = mr.Mat.from_csr(ctx, rp, ci, v, n_cols=1, store=mr.F64)
    f16 = m64.convert(mr.F16).export()[2]
    assert np.array_equal(f16.view(np.uint16), v.astype(np.float16).view(np.uint16))
    bf = m64.convert(mr.BF16).export()[2]
    # exact single rounding to 8 significant bits, ties to even (np.rint)
    m, e = np.frexp(v)
    r = np.ldexp(np.rint(m * 256.0), e - 8)
    want = (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    assert np.array_equal(bf.view(np.uint16), want)
    for m in (m64,):
        m.close()


def test_malformed_csr_rejected(ctx, mr):
    with pytest.raises(mr.MRError) as e:
        mr.Mat.from_csr(ctx, np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), n_cols=2)
    assert e.value.code == 6
    with pytest.raises(mr.MRError) as e:
        mr.Mat.from_csr(ctx, np.array([0, 1], np.int32), np.array([5], np.int32), n_cols=2)
    assert e.value.code == 6


def test_merbit_policies_bit_identical(ctx, mr):
    # the load policy (non-temporal streams, XCD block order) changes nothing
    # in the arithmetic: y is bit-identical for every policy and sigma pair
    name, rp, ci, nc = CASES[0]
    rng = np.random.default_rng(17)
    mat = mr.Mat.from_csr(ctx, rp, ci, rng.uniform(-1, 1, len(ci)), n_cols=nc, store=mr.F32)
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    for sigma in (8, 16, 32):
        mat.merbit_configure(sigma)
        ys = []
        for pol in range(4):
            mat.merbit_options(pol)
            ys.append(spmv(ctx, mr, mat, x, mr.MERBIT).view(np.uint32))
        assert all(np.array_equal(ys[0], yy) for yy in ys[1:])
    best = mat.merbit_autotune(3)
    assert best["sigma"] in (8, 16, 32) and 0 <= best["policy"] <= 3 and best["us"] > 0
    mat.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("k", [2, 3, 8, 16, 32, 64, 70])
def test_spmm_merbit(ctx, mr, oracle, case, k):
    name, rp, ci, nc = case
    rng = np.random.default_rng(100 + k)
    val = rng.uniform(-1, 1, len(ci))
    X = rng.uniform(-1, 1, (nc, k))
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F32)
    n = len(rp) - 1
    ldy = k + 3   # a padded leading dimension for Y
    Xd = ctx.upload(X.astype(np.float32))
    Yd = ctx.upload(np.full((n, ldy), 7.0, np.float32))
    mat.spmm(k, Xd, Yd, mr.MERBIT, ldx=k, ldy=ldy)
    ctx.sync()
    Y = Yd.to_host()
    Yh = oracle.spmm_f64(rp, ci, val.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64))
    absm = np.abs(sp_abs(rp, ci, val, nc)) @ np.abs(X) if n else np.zeros((0, k))
    err = np.abs(Y[:, :k] - Yh)
    assert np.all(err <= 1e-5 * np.maximum(absm, 1e-30) + 1e-30), float(np.max(err / np.maximum(absm, 1e-30)))
    assert np.all(Y[:, k:] == 7.0)   # padding columns untouched
    # bit-reproducible
    Yd2 = ctx.upload(np.zeros((n, ldy), np.float32))
    mat.spmm(k, Xd, Yd2, mr.MERBIT, ldx=k, ldy=ldy)
    ctx.sync()
    assert np.array_equal(Yd2.to_host()[:, :k].view(np.uint32), Y[:, :k].view(np.uint32))
    mat.close()


def sp_abs(rp, ci, val, nc):
    import scipy.sparse as sp
    return sp.csr_matrix((np.abs(val), ci, rp), shape=(len(rp) - 1, nc))


XC_SHAPES = [(8, 16), (16, 16), (8, 8), (16, 8), (4, 16)]


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("shape", XC_SHAPES, ids=lambda s: f"S{s[0]}W{s[1]}")
@pytest.mark.parametrize("cached", [-1, 0, 7])
def test_merbit_xcache_f32_parity(ctx, mr, oracle, case, shape, cached):
    """Wave-tile MERBIT with the LDS cache of hot columns: every tile shape,
    cache sizes from none to the LDS capacity, FP32 within 1e-5 of the FP64
    oracle; the fix-up joins rows spanning many tiles."""
    name, rp, ci, nc = case
    rng = np.random.default_rng(17)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F32)
    mat.merbit_xc_configure(shape[0], shape[1], cached)
    mat.merbit_options(5)
    info = mat.merbit_xc_info()
    assert info["items_per_lane"] == shape[0] and info["waves"] == shape[1]
    n = len(rp) - 1
    assert info["tiles"] == -(-(n + len(ci)) // (64 * shape[0]))
    if cached >= 0:
        assert info["cached_cols"] <= cached
    y = spmv(ctx, mr, mat, x, mr.MERBIT)
    yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float32).astype(np.float64), x.astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    mat.close()


@pytest.mark.parametrize("store", ["bf16", "f16"])
def test_merbit_xcache_low_precision_and_repro(ctx, mr, oracle, store):
    """BF16 / FP16 values with FP32 accumulation through the x-cache kernel:
    within 1e-5 of the FP64 oracle on the stored values, identical bits on
    every run, identical through the caller-workspace entry point."""
    rp, ci = rmat_csr(13, 16, seed=5)
    n = len(rp) - 1
    rng = np.random.default_rng(3)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, n).astype(np.float32)
    st = mr.BF16 if store == "bf16" else mr.F16
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=st)
    mat.merbit_options(4)
    _, _, vs = mat.export()
    if st == mr.BF16:
        v64 = (vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    else:
        v64 = vs.view(np.float16).astype(np.float64)
    y1 = spmv(ctx, mr, mat, x, mr.MERBIT)
    y2 = spmv(ctx, mr, mat, x, mr.MERBIT)
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))
    yhat, absrow = oracle.spmv_f64(rp, ci, v64, x.astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y1, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    ws = ctx.empty(mat.spmv_ws_bytes() // 4 + 1, np.float32)
    xd = ctx.upload(x)
    yd = ctx.empty(n, np.float32)
    mat.spmv_ws(xd, yd, ws)
    ctx.sync()
    assert np.array_equal(yd.to_host().view(np.uint32), y1.view(np.uint32))
    mat.close()


def test_merbit_xcache_rejected_for_f64(ctx, mr):
    rp, ci = rmat_csr(8, 4, seed=1)
    mat = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F64)
    with pytest.raises(mr.MRError) as e:
        mat.merbit_options(4)
    assert e.value.code == 5
    with pytest.raises(mr.MRError) as e:
        mat.merbit_xc_configure(16, 16, -1)
    assert e.value.code == 5
    mat.close()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("k,ldx", [(2, 4), (3, 4), (4, 4), (8, 12), (13, 16), (16, 16), (33, 36), (5, 5)])
@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_spmm_lane_padded(ctx, mr, oracle, case, k, ldx, store):
    """SpMM on the lane-padded tiles (policy bit 2) with the LDS cache of the
    first X rows: 16 columns per pass, ldx a multiple of 4 (ldx = 5 takes the
    workgroup-tile SpMM instead); 1e-5 of the FP64 oracle on the stored
    values, padding columns of Y untouched, bit-reproducible."""
    check_lp_spmm(ctx, mr, oracle, case, k, ldx, store, None)


SHAPE_CASES = [c for c in CASES if c[0] in ("rmat_s14", "long_rows", "stencil27")]


@pytest.mark.parametrize("case", SHAPE_CASES, ids=[c[0] for c in SHAPE_CASES])
@pytest.mark.parametrize("shape", [(8, 16), (16, 16), (8, 8), (16, 8)], ids=lambda s: f"S{s[0]}W{s[1]}")
@pytest.mark.parametrize("k,ldx", [(4, 4), (6, 8), (16, 16), (33, 36)])
@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_spmm_lane_padded_shapes(ctx, mr, oracle, case, shape, k, ldx, store):
    """The same on every tile shape (items per lane, waves): 16-item tiles
    with more than 8 columns launch 8 waves whatever the plan's count."""
    check_lp_spmm(ctx, mr, oracle, case, k, ldx, store, shape)


def check_lp_spmm(ctx, mr, oracle, case, k, ldx, store, shape):
    name, rp, ci, nc = case
    rng = np.random.default_rng(200 + k)
    val = rng.uniform(-1, 1, len(ci))
    X = rng.uniform(-1, 1, (nc, ldx)).astype(np.float32)
    st = mr.F32 if store == "f32" else mr.BF16
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=st)
    mat.merbit_options(4)
    if shape is not None:
        mat.merbit_xc_configure(shape[0], shape[1], -1)
    _, _, vs = mat.export()
    v64 = vs.astype(np.float64) if st == mr.F32 else (vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    n = len(rp) - 1
    ldy = k + 2
    Xd = ctx.upload(X)
    Yd = ctx.upload(np.full((n, ldy), 7.0, np.float32))
    mat.spmm(k, Xd, Yd, mr.MERBIT, ldx=ldx, ldy=ldy)
    ctx.sync()
    Y = Yd.to_host()
    Yh = oracle.spmm_f64(rp, ci, v64, X[:, :k].astype(np.float64))
    absm = np.abs(sp_abs(rp, ci, v64, nc)) @ np.abs(X[:, :k].astype(np.float64)) if n else np.zeros((0, k))
    err = np.abs(Y[:, :k] - Yh)
    assert np.all(err <= 1e-5 * np.maximum(absm, 1e-30) + 1e-30), float(np.max(err / np.maximum(absm, 1e-30)))
    assert np.all(Y[:, k:] == 7.0)
    Yd2 = ctx.upload(np.zeros((n, ldy), np.float32))
    mat.spmm(k, Xd, Yd2, mr.MERBIT, ldx=ldx, ldy=ldy)
    ctx.sync()
    assert np.array_equal(Yd2.to_host()[:, :k].view(np.uint32), Y[:, :k].view(np.uint32))
    mat.close()


def large_csr_cases():
    """Matrices of at least 2^20 nonzeros with every row class: rows over 4096
    entries (one workgroup each, dispatched first), every shorter class, empty
    rows; a wide one and a square one of 20000 columns."""
    rng = np.random.default_rng(29)
    out = []
    for name, n, nc, n_long in (("wide", 100_000, 300_000, 24), ("square_20000", 20_000, 20_000, 90)):
        lens = rng.choice([0, 0, 0, 1, 2, 3, 5, 9, 17, 40, 64, 65, 300, 4096], size=n,
                          p=[0.2, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.05, 0.05, 0.04, 0.04, 0.01, 0.008, 0.002])
        lens[rng.choice(n, n_long, replace=False)] = rng.integers(4097, min(nc, 30_000), n_long)
        rows = [np.sort(rng.choice(nc, int(k), replace=False)) for k in lens]
        rp = np.zeros(n + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        ci = np.concatenate(rows).astype(np.int32)
        assert len(ci) >= 1 << 20
        out.append((name, rp, ci, nc))
    return out


LARGE_CASES = large_csr_cases()


@pytest.mark.parametrize("case", LARGE_CASES, ids=[c[0] for c in LARGE_CASES])
@pytest.mark.parametrize("kernel", ["csr_scalar", "csr_vector"])
@pytest.mark.parametrize("store", ["f64", "f32", "bf16"])
def test_csr_large_parity(ctx, mr, oracle, case, kernel, store):
    name, rp, ci, nc = case
    rng = np.random.default_rng(31)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc)
    dt = {"f64": mr.F64, "f32": mr.F32, "bf16": mr.BF16}[store]
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=dt)
    k = {"csr_scalar": mr.CSR_SCALAR, "csr_vector": mr.CSR_VECTOR}[kernel]
    if store == "f64":
        y = spmv(ctx, mr, mat, x, k)
        yhat, absrow = oracle.spmv_f64(rp, ci, val, x)
        tol = 1e-12
    else:
        x32 = x.astype(np.float32)
        y = spmv(ctx, mr, mat, x32, k)
        _, _, vs = mat.export()
        v64 = ((vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64) if store == "bf16"
               else np.asarray(vs, np.float64))
        yhat, absrow = oracle.spmv_f64(rp, ci, v64, x32.astype(np.float64))
        tol = 1e-5
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= tol and l2 <= tol, (worst, l2)
    y2 = spmv(ctx, mr, mat, x if store == "f64" else x.astype(np.float32), k)
    assert np.array_equal(np.asarray(y2).view(np.uint8), np.asarray(y).view(np.uint8))
    mat.close()
