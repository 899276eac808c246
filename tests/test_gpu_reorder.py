"""Reordering on the device against the oracle, bit for bit: the relabel
P A P^T (row/column permutation with sorted columns), the pGOrder GPU
placement loop and its per-batch CPU Gorder phase, CPU Gorder (gorder1) and
RCM, and the GRAPH-style workflow (reorder, multiply, un-permute)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import csr_rows, random_csr, rmat_csr

pytestmark = pytest.mark.gpu


def relabel_cases():
    rng = np.random.default_rng(31)
    out = []
    rp, ci = random_csr(rng, 500, empty_frac=0.3, max_row=80)
    out.append(("random_500", rp, ci))
    n = 12000
    rows = []
    for i in range(n):
        if i == 7:
            rows.append(np.sort(rng.choice(n, 10000, replace=False)))   # > 8192: global bitonic
        elif i % 1000 == 3:
            rows.append(np.sort(rng.choice(n, 3000, replace=False)))    # LDS bitonic
        else:
            rows.append(np.sort(rng.choice(n, int(rng.integers(0, 70)), replace=False)))
    out.append(("long_rows_12000", *csr_rows(rows, n)))
    out.append(("rmat_s14", *rmat_csr(14, 16, seed=1)))
    return out


@pytest.mark.parametrize("case", relabel_cases(), ids=lambda c: c[0])
def test_permute_bit_exact(ctx, mr, oracle, case):
    name, rp, ci = case
    n = len(rp) - 1
    rng = np.random.default_rng(32)
    val = rng.uniform(-1, 1, len(ci))
    perm = rng.permutation(n).astype(np.int32)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F64)
    mat.permute(perm)
    grp, gci, gval = mat.export()
    orp, oci, oval = oracle.permute(n, rp, ci, val, perm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci) and np.array_equal(gval, oval)
    assert np.array_equal(mat.perm(), perm)
    # composition: a second relabel accumulates perm[old] = new
    p2 = rng.permutation(n).astype(np.int32)
    mat.permute(p2)
    assert np.array_equal(mat.perm(), p2[perm])
    mat.close()


@pytest.mark.parametrize("store", ["F64", "F32", "BF16"])
def test_permute_long_rows_all_widths(ctx, mr, oracle, store):
    """Several rows past 4096 entries (sorted together by one radix sort),
    repeated columns inside them (equal new columns keep their input order),
    and 8-, 4- and 2-byte values (values exact in every width)."""
    rng = np.random.default_rng(57)
    n = 30000
    long_len = {11: 4097, 500: 9000, 8191: 20000, 20000: 5000, 29999: 12000}
    rows = []
    for i in range(n):
        if i in long_len:
            r = np.sort(rng.choice(n, long_len[i], replace=True))   # repeats kept
        else:
            r = np.sort(rng.choice(n, int(rng.integers(0, 40)), replace=False))
        rows.append(r)
    rp, ci = csr_rows(rows, n)
    val = rng.integers(-64, 64, len(ci)) / 8.0
    perm = rng.permutation(n).astype(np.int32)
    mat = mr.Mat.from_csr(ctx, rp, ci, val, store=getattr(mr, store))
    mat.permute(perm)
    grp, gci, gval = mat.export()
    orp, oci, oval = oracle.permute(n, rp, ci, val, perm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    if store == "BF16":   # exported as raw bfloat16 bits
        gval = (gval.astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(np.asarray(gval, np.float64), oval)
    mat.close()


def test_permute_rejects_non_permutation(ctx, mr):
    rp, ci = random_csr(np.random.default_rng(1), 10)
    mat = mr.Mat.from_csr(ctx, rp, ci)
    with pytest.raises(mr.MRError) as e:
        mat.permute(np.zeros(10, np.int32))
    assert e.value.code == 4
    mat.close()


def pgorder_cases():
    rng = np.random.default_rng(41)
    out = [("rmat_s14_w1024", *rmat_csr(14, 16, seed=1), 1024)]
    rp, ci = random_csr(rng, 3000, empty_frac=0.4, max_row=12)
    out.append(("random_3000_w64", rp, ci, 64))
    rp, ci = random_csr(rng, 700, empty_frac=0.1, max_row=30)
    out.append(("random_700_w32", rp, ci, 32))
    out.append(("rmat_s12_w128", *rmat_csr(12, 8, seed=4), 128))
    rp, ci = random_csr(rng, 20, empty_frac=0.2, max_row=5)
    out.append(("tiny_20_w32", rp, ci, 32))
    out.append(("no_edges_5000", np.zeros(5001, np.int32), np.zeros(0, np.int32), 1024))
    # hubs above the 10000 cap of the GPU window score
    n = 30000
    rows = [np.sort(rng.choice(n, 12000, replace=False)) if i in (5, 9000) else
            np.sort(rng.choice(n, int(rng.integers(0, 6)), replace=False)) for i in range(n)]
    out.append(("hubs_30000_w1024", *csr_rows(rows, n), 1024))
    # hundreds of batches: the column lists are compacted every 32 batches
    out.append(("rmat_s13_w32_compacted", *rmat_csr(13, 16, seed=13), 32))
    out.append(("rmat_s15_w64_compacted", *rmat_csr(15, 16, seed=15), 64))
    out.extend(selection_path_cases())
    out.extend(wide_cases())
    return out


def wide_cases():
    """Widths above 1024: from the batch where |P| or the window passes 1024
    the selection is spread over the grid (tie key found digit by digit, two
    stable radix sorts, a multi-workgroup fill) and the window's prefix sums
    come from a device scan.

    width 0 takes get_width: R-MAT s20 with two edges per row gives 16384,
    one edge per row (nnz < n, so nnz / n = 0) gives 262144, the widths of
    every published R-MAT s24 run.  The explicit widths run the wide path on
    small graphs for many batches: ties of one exact count larger than the
    need (the id digits), counts of 131072 (an inexact bin: the count digits
    too), batches filled from untouched vertices, hubs above the cap, and the
    column-list compaction every 32 batches."""
    rng = np.random.default_rng(43)
    out = [("rmat_s20_ef2_width16384", *rmat_csr(20, 2, seed=7), 0),
           ("rmat_s20_ef1_width262144", *rmat_csr(20, 1, seed=7), 0),
           ("rmat_s14_w2048", *rmat_csr(14, 16, seed=1), 2048),
           ("rmat_s17_ef4_w2048_compacted", *rmat_csr(17, 4, seed=17), 2048)]
    rp, ci = random_csr(rng, 9000, empty_frac=0.4, max_row=12)
    out.append(("random_9000_w4096", rp, ci, 4096))
    out.append(("no_edges_9000_w2048", np.zeros(9001, np.int32), np.zeros(0, np.int32), 2048))
    n = 30000
    rows = [np.sort(rng.choice(n, 12000, replace=False)) if i in (5, 9000) else
            np.sort(rng.choice(n, int(rng.integers(0, 6)), replace=False)) for i in range(n)]
    out.append(("hubs_30000_w4096", *csr_rows(rows, n), 4096))
    for name, rp, ci, _ in selection_path_cases():
        out.append((name.replace("_w1024", "_w2048"), rp, ci, 2048))
    return out


def selection_path_cases():
    """Graphs built to reach the rarer selection paths of the GPU loop.

    ties: row 0 points at columns 1-3, each held by 9000 one-entry rows, so a
    window with row 0 gives ~27000 touched vertices in three equal-count
    groups -- more than 8192 ties of an exact count (the id-histogram path).
    dense: 6000 rows share 64 columns; once a 1024-row window holds them, the
    other rows reach count 64 * 1024 = 65536, beyond the exact bins, with
    more than 4096 candidates (the repeated-minimum path).  dense_mixed adds
    1000 rows holding one of those columns: the same inexact bin with at most
    4096 candidates (one sort of the candidates).
    """
    rng = np.random.default_rng(7)
    n = 30000
    rows = [np.zeros(0, np.int64) for _ in range(n)]
    rows[0] = np.array([1, 2, 3])
    zs = rng.permutation(np.arange(10, n))[:27000]
    for g in range(3):
        for z in zs[g * 9000:(g + 1) * 9000]:
            rows[z] = np.array([1 + g])
    out = [("ties_30000_w1024", *csr_rows(rows, n), 1024)]
    n = 6200
    block = np.arange(6100, 6164)
    rows = [block] * 6000 + [np.zeros(0, np.int64)] * 200
    out.append(("dense_6000x64_w1024", *csr_rows(rows, n), 1024))
    rows = [block] * 5000 + [block[:1]] * 1000 + [np.zeros(0, np.int64)] * 200
    out.append(("dense_mixed_w1024", *csr_rows(rows, n), 1024))
    return out


@pytest.mark.parametrize("case", pgorder_cases(), ids=lambda c: c[0])
def test_pgorder_bit_exact(ctx, mr, oracle, case):
    name, rp, ci, W = case
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    gpu_order, final_order, st = mat.pgorder_order(W, 5)
    if W == 0:
        W = oracle.get_width(n, len(ci))
        assert st["width"] == W, (st["width"], W)
    og, of, operm = oracle.pgorder(n, rp, ci, W, 5)
    assert np.array_equal(gpu_order, og), "GPU placement loop differs"
    assert np.array_equal(final_order, of), "CPU Gorder phase differs"
    perm, st2 = mat.reorder_pgorder(W, 5)
    assert np.array_equal(perm, operm)
    grp, gci, _ = mat.export()
    orp, oci, _ = oracle.permute(n, rp, ci, None, operm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    mat.close()


CPU_PHASE_CASES = ("rmat_s12_w128", "tiny_20_w32", "no_edges_5000", "hubs_30000_w1024", "ties_30000_w1024",
                   "rmat_s20_ef2_width16384", "rmat_s14_w2048", "hubs_30000_w4096")


@pytest.mark.parametrize("case", [c for c in pgorder_cases() if c[0] in CPU_PHASE_CASES], ids=lambda c: c[0])
@pytest.mark.parametrize("threads", [1, 4])
def test_pgorder_cpu_phase_alone(ctx, mr, oracle, case, threads):
    """mr_pgorder_cpu_phase on the GPU loop's placement gives
    mr_pgorder_order's final order, on one thread and on four."""
    name, rp, ci, W = case
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    gpu_order, final_order, _ = mat.pgorder_order(W, 5)
    alone, ms = mat.pgorder_cpu_phase(gpu_order, W, 5, threads)
    assert np.array_equal(alone, final_order)
    assert ms >= 0.0
    with pytest.raises(ValueError):
        mat.pgorder_cpu_phase(gpu_order[:-1], W, 5, threads)
    if len(gpu_order) > 1:
        bad = gpu_order.copy()
        bad[0] = bad[1]
        with pytest.raises(mr.MRError) as e:
            mat.pgorder_cpu_phase(bad, W, 5, threads)
        assert e.value.code == 4
        bad[0] = len(gpu_order)
        with pytest.raises(mr.MRError):
            mat.pgorder_cpu_phase(bad, W, 5, threads)
    mat.close()


@pytest.mark.parametrize("scale,ef", [(10, 8), (12, 16)])
def test_gorder_cpu_bit_exact(ctx, mr, oracle, scale, ef):
    rp, ci = rmat_csr(scale, ef, seed=2)
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci)
    perm, _ = mat.reorder_gorder_cpu(5)
    _, operm = oracle.gorder1(n, rp, ci, 5)
    assert np.array_equal(perm, operm)
    mat.close()


def test_rcm_bit_exact(ctx, mr, oracle):
    rng = np.random.default_rng(51)
    for n in (30, 400):
        rp, ci = random_csr(rng, n, empty_frac=0.2, max_row=6)
        mat = mr.Mat.from_csr(ctx, rp, ci)
        perm, _ = mat.reorder_rcm()
        # the reference hands RCM its COO sorted by (col, row) = rows of M in order:
        # COO row_ind = column of M, col_ind = row of M
        rows = ci.copy()
        cols = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
        order = oracle.rcm(n, rows, cols)
        want = np.empty(n, np.int32)
        want[order] = np.arange(n, dtype=np.int32)
        assert np.array_equal(perm, want)
        mat.close()


def test_graph_workflow_unpermutes(ctx, mr, oracle):
    rp, ci = rmat_csr(13, 16, seed=3)
    n = len(rp) - 1
    rng = np.random.default_rng(61)
    val = rng.uniform(-1, 1, len(ci))
    x = rng.uniform(-1, 1, n)
    g = mr.Graph(ctx, mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F64))
    g.gorder6()
    yhat, absrow = oracle.spmv_f64(rp, ci, val, x)
    for method in ("merbit", "csr", "csr_scalar"):
        y = g.spmv(x, method)
        worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
        assert worst <= 1e-12 and l2 <= 1e-12, (method, worst)


def rabbit_cases():
    rng = np.random.default_rng(71)
    out = [("rmat_s12", *rmat_csr(12, 8, seed=12))]
    rp, ci = random_csr(rng, 900, empty_frac=0.2, max_row=10)
    out.append(("random_900", rp, ci))
    # the reference's own demo graph (rabbitorder/src/main.cpp:9-11): COO
    # (row_ind, col_ind) with 12 vertices; as M's CSR, row = col_ind
    rows = [[] for _ in range(12)]
    for r, c in zip([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11], [2, 2, 4, 5, 5, 2, 8, 9, 10, 8, 1]):
        rows[c].append(r)
    out.append(("demo_12", *csr_rows([sorted(r) for r in rows], 12)))
    return out


@pytest.mark.parametrize("case", rabbit_cases(), ids=lambda c: c[0])
def test_rabbit_matches_oracle(ctx, mr, oracle, case):
    name, rp, ci = case
    n = len(rp) - 1
    ids, _ = oracle.rabbit(n, rp, ci)
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    perm, st = mat.reorder_rabbit(False)
    assert np.array_equal(perm, ids.astype(np.int32))
    grp, gci, _ = mat.export()
    orp, oci, _ = oracle.permute(n, rp, ci, None, perm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    mat.close()
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    perm2, _ = mat.reorder_rabbit(True)            # the reference call site's scatter
    inv = np.empty(n, np.int32)
    inv[ids] = np.arange(n, dtype=np.int32)
    assert np.array_equal(perm2, inv)
    mat.close()


def degree_cases():
    rng = np.random.default_rng(81)
    out = [("rmat_s14", *rmat_csr(14, 16, seed=1)), ("rmat_s17_hubs", *rmat_csr(17, 16, seed=17))]
    rp, ci = random_csr(rng, 3000, empty_frac=0.3, max_row=40)
    out.append(("random_3000", rp, ci))
    out.append(("no_edges_700", np.zeros(701, np.int32), np.zeros(0, np.int32)))
    rp = np.array([0, 2, 3, 6, 6, 8], np.int32)
    out.append(("hand_5", rp, np.array([1, 2, 2, 2, 3, 4, 0, 2], np.int32)))
    return out


@pytest.mark.parametrize("grouped", [False, True], ids=["degree", "dbg"])
@pytest.mark.parametrize("case", degree_cases(), ids=lambda c: c[0])
def test_degree_relabels_bit_exact(ctx, mr, oracle, case, grouped):
    name, rp, ci = case
    n = len(rp) - 1
    want = oracle.order_degree(n, rp, ci, grouped)
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    perm, st = mat.reorder_dbg() if grouped else mat.reorder_degree()
    assert np.array_equal(perm, want)
    grp, gci, _ = mat.export()
    orp, oci, _ = oracle.permute(n, rp, ci, None, want)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    mat.close()


def gorder5_cases():
    rng = np.random.default_rng(91)
    out = [("rmat_s14_w1024", *rmat_csr(14, 16, seed=1), 0, 1, 5)]
    rp, ci = random_csr(rng, 2000, empty_frac=0.3, max_row=12)
    out.append(("random_2000_v64_w2", rp, ci, 64, 2, 5))
    rp, ci = random_csr(rng, 700, empty_frac=0.1, max_row=30)
    out.append(("random_700_v10_w3", rp, ci, 10, 3, 3))
    out.append(("rmat_s12_v128", *rmat_csr(12, 8, seed=4), 128, 1, 5))
    out.append(("no_edges_3000", np.zeros(3001, np.int32), np.zeros(0, np.int32), 1024, 1, 5))
    # hubs above gorder6's 10000 cap: gorder5 counts them
    n = 30000
    rows = [np.sort(rng.choice(n, 12000, replace=False)) if i in (5, 9000) else
            np.sort(rng.choice(n, int(rng.integers(0, 6)), replace=False)) for i in range(n)]
    out.append(("hubs_30000", *csr_rows(rows, n), 1024, 1, 5))
    out.append(("rmat_s13_v32_compacted", *rmat_csr(13, 16, seed=13), 32, 1, 5))
    out.extend((name, rp, ci, W, 1, 5) for name, rp, ci, W in selection_path_cases())
    # batches and windows above 1024 (the wide selection), a batch size that is
    # not a power of two
    out.append(("rmat_s14_v2048_w2", *rmat_csr(14, 16, seed=1), 2048, 2, 5))
    rp, ci = random_csr(rng, 9000, empty_frac=0.3, max_row=12)
    out.append(("random_9000_v1500_w3", rp, ci, 1500, 3, 5))
    return out


@pytest.mark.parametrize("case", gorder5_cases(), ids=lambda c: c[0])
def test_gorder5_bit_exact(ctx, mr, oracle, case):
    name, rp, ci, vsize, w1, w2 = case
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    g, f, st = mat.gorder5_order(vsize, w1, w2)
    og, of, operm = oracle.gorder5(n, rp, ci, vsize or mr.get_width(n, len(ci)), w1, w2)
    assert np.array_equal(g, og), "GPU placement loop differs"
    assert np.array_equal(f, of), "CPU phase differs"
    perm, _ = mat.reorder_gorder5(vsize, w1, w2)
    assert np.array_equal(perm, operm)
    grp, gci, _ = mat.export()
    orp, oci, _ = oracle.permute(n, rp, ci, None, operm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    mat.close()


def test_gorder5_argument_limits(ctx, mr):
    rp, ci = rmat_csr(10, 8, seed=1)
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    for args in ((1 << 23, 1, 5), (1 << 22, 5, 5), (64, 0, 5), (64, 1, 0)):
        with pytest.raises(mr.MRError):
            mat.reorder_gorder5(*args)
    mat.close()


@pytest.mark.parametrize("scale,ef,w1,w2", [(10, 8, 1024, 512), (12, 16, 256, 64), (11, 4, 7, 1000)])
def test_tile_order_bit_exact(ctx, mr, oracle, scale, ef, w1, w2):
    rp, ci = rmat_csr(scale, ef, seed=scale)
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    perm, st = mat.reorder_tile_order(w1, w2)
    _, operm = oracle.tile_order(n, rp, ci, w1, w2)
    assert np.array_equal(perm, operm)
    grp, gci, _ = mat.export()
    orp, oci, _ = oracle.permute(n, rp, ci, None, operm)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    mat.close()
