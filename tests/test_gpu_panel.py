"""Column-panel MERBIT (policy bit 3) through the C-ABI: y against the FP64
oracle on the stored values (FP32 accumulation, 1e-5 condition-aware and
relative L2) for F32 / F16 / BF16 values, every tile shape and streaming
hint; bit-reproducible; the workspace and no-allocation entry points equal
mr_spmv; the plan's segment, tile and partial counts equal a numpy
restatement of the layout (32768-column panels, segments cut at tile ends)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import csr_rows, random_csr, rmat_csr, spmv_cases

pytestmark = pytest.mark.gpu

PANEL = 32768


def panel_cases():
    rng = np.random.default_rng(17)
    cases = [c for c in spmv_cases() if c[0] != "rmat_s14"]
    rp, ci = rmat_csr(16, 16, seed=3)                       # two panels
    cases.append(("rmat_s16", rp, ci, 1 << 16))
    rp, ci = random_csr(rng, 3000, empty_frac=0.4, max_row=90, n_cols=100_000)   # four panels, ragged
    cases.append(("random_wide", rp, ci, 100_000))
    # rows whose run inside one panel is longer than any tile (cut at tile ends)
    rows = [np.sort(rng.choice(140_000, 6000, replace=False)) if i % 7 == 2 else
            np.sort(rng.choice(140_000, int(rng.integers(0, 40)), replace=False)) for i in range(80)]
    rp, ci = csr_rows(rows, 140_000)
    cases.append(("long_runs", rp, ci, 140_000))
    # non-square, columns up to the last id
    rows = [np.unique(np.concatenate([rng.integers(0, 200_000, 12), [199_999]])) for _ in range(500)]
    rp, ci = csr_rows(rows, 200_000)
    cases.append(("nonsquare_tail", rp, ci, 200_000))
    # rows around the row-sum tile size (512 items: partials plus an end
    # marker): one entry in each of 510, 511 (a tile exactly), 512 and 600
    # panels, between short and empty rows
    nc = 600 * PANEL
    rows = []
    for i in range(400):
        if i in (7, 100, 101, 250, 399):
            k = {7: 600, 100: 511, 101: 512, 250: 510, 399: 600}[i]
            pan = np.sort(rng.choice(600, k, replace=False))
            rows.append(pan * PANEL + rng.integers(0, PANEL, k))
        else:
            rows.append(np.sort(rng.choice(nc, int(rng.integers(0, 9)), replace=False)))
    rp, ci = csr_rows(rows, nc)
    cases.append(("rowsum_long_rows", rp, ci, nc))
    # runs of empty rows (folded into the end marker before them, at most 512
    # rows per row-sum tile): 1, 511, 512, 513 and 1500 in a row, one at the
    # start and one at the end of the matrix
    nc = 70_000
    rows = [np.zeros(0, np.int64)] * 700
    for run in (1, 511, 512, 513, 1500, 3):
        rows += [np.sort(rng.choice(nc, int(rng.integers(1, 30)), replace=False)) for _ in range(int(rng.integers(1, 40)))]
        rows += [np.zeros(0, np.int64)] * run
    rows += [np.sort(rng.choice(nc, 600, replace=False))]          # a long row, then more empties
    rows += [np.zeros(0, np.int64)] * 900
    rp, ci = csr_rows(rows, nc)
    cases.append(("empty_row_runs", rp, ci, nc))
    return cases


CASES = panel_cases()
IDS = [c[0] for c in CASES]


def layout_counts(rp, ci, S):
    """(segments, tiles, partials) of the panel layout: per panel, rows in
    order, each row's run of nonzeros in the panel contiguous; a panel starts
    on a tile of 64*S items; a run crossing a tile end is cut there."""
    T = 64 * S
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    pan = ci.astype(np.int64) // PANEL
    if len(ci) == 0:
        return 0, 0, 0
    start = np.ones(len(ci), bool)
    start[1:] = (rows[1:] != rows[:-1]) | (pan[1:] != pan[:-1])
    segs = int(start.sum())
    s_idx = np.nonzero(start)[0]
    s_len = np.diff(np.append(s_idx, len(ci)))
    s_pan = pan[s_idx]
    order = np.argsort(s_pan, kind="stable")
    tiles = parts = 0
    for p in np.unique(s_pan):
        lens = s_len[order][s_pan[order] == p]
        off = np.concatenate([[0], np.cumsum(lens)[:-1]])
        tiles += -(-int(lens.sum()) // T)
        parts += int(np.sum((off + lens - 1) // T - off // T + 1))
    return segs, tiles, parts


def stored_f64(mat, store):
    _, _, vs = mat.export()
    if store == "bf16":
        return (vs.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    if store == "f16":
        return vs.view(np.float16).astype(np.float64)
    return np.asarray(vs, np.float64)


def run(ctx, mat, x):
    xd = ctx.upload(x)
    yd = ctx.empty(mat.info()[0], np.float32)
    mat.spmv(xd, yd)
    ctx.sync()
    return yd.to_host()


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("store", ["f32", "bf16", "f16"])
def test_panel_parity(ctx, mr, oracle, case, store):
    name, rp, ci, nc = case
    rng = np.random.default_rng(5)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    dt = {"f32": mr.F32, "bf16": mr.BF16, "f16": mr.F16}[store]
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=dt)
    mat.merbit_options(8)
    y = run(ctx, mat, x)
    yhat, absrow = oracle.spmv_f64(rp, ci, stored_f64(mat, store), x.astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    info = mat.merbit_panel_info()
    assert (info["segments"], info["tiles"], info["partials"]) == layout_counts(rp, ci, info["items_per_lane"])
    assert np.array_equal(run(ctx, mat, x).view(np.uint32), y.view(np.uint32))
    mat.close()


@pytest.mark.parametrize("shape", [(8, 16), (8, 8), (16, 8)], ids=lambda s: f"S{s[0]}W{s[1]}")
@pytest.mark.parametrize("nt", [0, 1])
@pytest.mark.parametrize("name", ["rmat_s16", "long_runs", "random_wide"])
def test_panel_shapes_and_entry_points(ctx, mr, oracle, shape, nt, name):
    _, rp, ci, nc = next(c for c in CASES if c[0] == name)
    rng = np.random.default_rng(9)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=mr.F32)
    mat.merbit_options(8 | nt)
    mat.merbit_panel_configure(*shape)
    info = mat.merbit_panel_info()
    assert (info["items_per_lane"], info["waves"]) == shape
    assert (info["segments"], info["tiles"], info["partials"]) == layout_counts(rp, ci, shape[0])
    y = run(ctx, mat, x)
    yhat, absrow = oracle.spmv_f64(rp, ci, val.astype(np.float32).astype(np.float64), x.astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    # caller workspace (one float per partial) and the static scratch
    assert mat.spmv_ws_bytes() == 4 * max(info["partials"], 1)
    xd = ctx.upload(x)
    ws = ctx.empty(mat.spmv_ws_bytes() // 4, np.float32)
    yd = ctx.empty(len(y), np.float32)
    mat.spmv_ws(xd, yd, ws)
    ctx.sync()
    assert np.array_equal(yd.to_host().view(np.uint32), y.view(np.uint32))
    yd2 = ctx.empty(len(y), np.float32)
    mat.spmv_noalloc(xd, yd2)
    ctx.sync()
    assert np.array_equal(yd2.to_host().view(np.uint32), y.view(np.uint32))
    mat.close()


def test_panel_policy_rules(ctx, mr):
    rp, ci = rmat_csr(10, 8, seed=2)
    m64 = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F64)
    with pytest.raises(mr.MRError) as e:
        m64.merbit_options(8)
    assert e.value.code == 5
    m64.close()
    m = mr.Mat.from_csr(ctx, rp, ci, None, store=mr.F32)
    with pytest.raises(mr.MRError) as e:
        m.merbit_options(12)
    assert e.value.code == 1
    with pytest.raises(mr.MRError):
        m.merbit_panel_configure(4, 16)
    with pytest.raises(mr.MRError):
        m.merbit_panel_configure(16, 16)      # its LDS staging would not fit beside the panel
    assert m.merbit_panel_info()["partials"] == -1
    m.close()


def test_panel_after_reorder_and_spmm(ctx, mr, oracle):
    """The plan follows a relabel (pGOrder), and SpMM on a panel-policy matrix
    runs on the wave tiles: both against the oracle."""
    rp, ci = rmat_csr(15, 16, seed=8)
    n = len(rp) - 1
    rng = np.random.default_rng(4)
    val = rng.uniform(-1, 1, len(ci))
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    mat.merbit_options(8)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    run(ctx, mat, x)                       # plan on the original order
    perm, _ = mat.reorder_pgorder(0, 5)
    prp, pci, pval = mat.export()
    xp = np.empty_like(x)
    xp[perm] = x
    y = run(ctx, mat, xp)
    yhat, absrow = oracle.spmv_f64(prp, pci, np.asarray(pval, np.float64), xp.astype(np.float64))
    worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
    assert worst <= 1e-5 and l2 <= 1e-5, (worst, l2)
    info = mat.merbit_panel_info()
    assert (info["segments"], info["tiles"], info["partials"]) == layout_counts(prp, pci, info["items_per_lane"])
    k = 6
    X = rng.uniform(-1, 1, (n, 8)).astype(np.float32)
    Xd = ctx.upload(X)
    Yd = ctx.empty((n, k), np.float32)
    mat.spmm(k, Xd, Yd, ldx=8, ldy=k)
    ctx.sync()
    Y = Yd.to_host()
    for j in range(k):
        yh, ar = oracle.spmv_f64(prp, pci, np.asarray(pval, np.float64), X[:, j].astype(np.float64))
        w, l, _ = oracle.spmv_parity(Y[:, j], yh, ar)
        assert w <= 1e-5 and l <= 1e-5, (j, w, l)
    mat.close()


class _At:
    """A device pointer `off` bytes into a DeviceArray (what mr_spmv reads is .ptr)."""

    def __init__(self, base, off):
        import ctypes
        self.ptr = ctypes.c_void_p(base.ptr.value + off)


@pytest.mark.parametrize("store", ["f32", "bf16"])
@pytest.mark.parametrize("name", ["rmat_s16", "random_wide", "nonsquare_tail"])
def test_panel_unaligned_x_and_y(ctx, mr, name, store):
    """x and y 4 bytes past a 16-byte boundary: the x panels load lane by
    lane (no 16-byte vectors, no next-panel prefetch), y stores are plain;
    the result equals the aligned call bit for bit."""
    _, rp, ci, nc = next(c for c in CASES if c[0] == name)
    rng = np.random.default_rng(13)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    dt = {"f32": mr.F32, "bf16": mr.BF16}[store]
    mat = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=nc, store=dt)
    mat.merbit_options(9)
    y = run(ctx, mat, x)
    n = mat.info()[0]
    xb = ctx.upload(np.concatenate([[np.float32(7.0)], x]).astype(np.float32))
    yb = ctx.empty(n + 1, np.float32)
    mat.spmv(_At(xb, 4), _At(yb, 4))
    ctx.sync()
    yu = yb.to_host()[1:]
    assert np.array_equal(yu.view(np.uint32), y.view(np.uint32))
    mat.close()


def test_stored_bytes_of_each_variant(ctx, mr):
    """mr_merbit_stored_bytes: -1 before a plan exists; for each variant the
    count its plan implies (column panels: item streams, tile bases, partials
    written and read back, the row-sum stream of whole 512-item tiles and its
    tile rows, x and y)."""
    _, rp, ci, nc = next(c for c in CASES if c[0] == "random_wide")
    n, nnz = len(rp) - 1, len(ci)
    rng = np.random.default_rng(21)
    mat = mr.Mat.from_csr(ctx, rp, ci, rng.uniform(-1, 1, nnz), n_cols=nc, store=mr.F32)
    xy = (nc + n) * 4
    mat.merbit_options(9)
    assert mat.merbit_stored_bytes() == -1
    x = rng.uniform(-1, 1, nc).astype(np.float32)
    run(ctx, mat, x)
    info = mat.merbit_panel_info()
    base = info["tiles"] * 64 * info["items_per_lane"] * 6 + info["tiles"] * 4 + 2 * info["partials"] * 4 + xy
    rest = mat.merbit_stored_bytes() - base            # row-sum tiles * 2048 + (tiles + 1) * 4
    assert rest >= 2052 and (rest - 4) % 2052 == 0, rest
    assert (rest - 4) // 2052 >= -(-(info["partials"]) // 512)
    mat.merbit_options(1)
    run(ctx, mat, x)
    wi = mat.merbit_info()
    assert mat.merbit_stored_bytes() == (nnz * 8 + wi["n_blocks"] * 256 * wi["sigma"] // 8 + (wi["n_blocks"] + 1) * 8
                                         + 2 * wi["n_blocks"] * 8 + xy)
    mat.merbit_options(5)
    run(ctx, mat, x)
    xi = mat.merbit_xc_info()
    assert mat.merbit_stored_bytes() == (xi["tiles"] * 64 * xi["items_per_lane"] * 8 + xi["tiles"] * 4
                                         + xi["cached_cols"] * 4 + 2 * xi["tiles"] * 8 + xy)
    mat.close()
