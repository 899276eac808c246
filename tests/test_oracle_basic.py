"""The oracle pinned against the reference's own artefacts and against
mathematics that does not share its code: reference-built heap traces, the
reference model headers, the reference RCM, scipy.sparse, brute force."""
from __future__ import annotations

import json
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN, random_csr


def test_heap_matches_reference_traces(oracle):
    traces = json.loads((GOLDEN / "heap_traces.json").read_text())
    assert len(traces) >= 30
    for t in traces:
        pops = oracle.heap_trace(t["n"], t["buckets"], np.array(t["ops"], np.int32))
        assert pops.tolist() == t["pops"], f"trace n={t['n']}"


def test_heap_refresh_traces_match_reference(oracle):
    """refresh and PMAX resets (the tile-aware order's heap operations)
    against pop orders of the reference Integer_Heap."""
    traces = json.loads((GOLDEN / "heap_refresh_traces.json").read_text())
    assert len(traces) >= 20 and any(o[0] == 3 for t in traces for o in t["ops"])
    for t in traces:
        pops = oracle.heap_trace(t["n"], t["buckets"], np.array(t["ops"], np.int32))
        assert pops.tolist() == t["pops"], f"trace n={t['n']}"


def test_heap_hand_worked_order(oracle):
    # all in bucket 0: a push goes right after the head, so the list reads
    # 0, n-1, ..., 1; raising node 2 moves it to bucket 1 first
    n = 5
    ops = [(0, v, 0) for v in range(n)] + [(1, 2, 1)] + [(2, 0, 0)] * n
    assert oracle.heap_trace(n, 64, np.array(ops, np.int32)).tolist() == [2, 0, 4, 3, 1]
    # a lazy decrement is applied when the node reaches the head of the top bucket
    ops = [(0, 0, 0), (0, 1, 0), (1, 0, 2), (1, 1, 1), (1, 0, -2), (2, 0, 0), (2, 0, 0)]
    assert oracle.heap_trace(2, 64, np.array(ops, np.int32)).tolist() == [1, 0]


@pytest.mark.parametrize("n,nnz,w", [(16384, 228000, 1024), (4194304, 65_000_000, 1024), (1 << 24, 4 << 24, 16384),
                                     (1 << 26, 2 << 26, 262144), (1000, 10, 262144), (100, 1000, 1024)])
def test_get_width_known_answers(oracle, n, nnz, w):
    # round(log2(sqrt(n / (nnz // n)))): <10 -> 1024, <12 -> 16384, else 262144; nnz < n divides by 0 -> inf
    assert oracle.get_width(n, nnz) == w


def test_transpose_and_permute_vs_scipy(oracle):
    rng = np.random.default_rng(1)
    for n in (1, 7, 50, 300):
        rp, ci = random_csr(rng, n)
        val = rng.uniform(-1, 1, len(ci))
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        tp, ti = oracle.transpose(n, n, rp, ci)
        At = A.T.tocsr()
        At.sort_indices()
        assert np.array_equal(tp, At.indptr) and np.array_equal(ti, At.indices)
        perm = rng.permutation(n).astype(np.int32)
        inv = np.argsort(perm)
        rp2, ci2, v2 = oracle.permute(n, rp, ci, val, perm)
        B = A[inv][:, inv].tocsr()
        B.sort_indices()
        assert np.array_equal(rp2, B.indptr) and np.array_equal(ci2, B.indices) and np.array_equal(v2, B.data)
        # and back
        rp3, ci3, v3 = oracle.permute(n, rp2, ci2, v2, inv.astype(np.int32))
        assert np.array_equal(rp3, rp) and np.array_equal(ci3, ci) and np.array_equal(v3, val)


def test_check_perm(oracle):
    assert oracle.check_perm(np.arange(10))
    assert not oracle.check_perm(np.array([0, 1, 1, 3]))
    assert not oracle.check_perm(np.array([1, 2, 3, 4]))


def test_spmv_f64_vs_scipy(oracle):
    rng = np.random.default_rng(2)
    for n in (1, 33, 400):
        rp, ci = random_csr(rng, n, empty_frac=0.4)
        val = rng.uniform(-1, 1, len(ci))
        x = rng.uniform(-1, 1, n)
        y, absrow = oracle.spmv_f64(rp, ci, val, x)
        ref = sp.csr_matrix((val, ci, rp), shape=(n, n)) @ x
        worst, l2, _ = oracle.spmv_parity(y, ref, absrow)
        assert worst <= 1e-14 and l2 <= 1e-14
    X = rng.uniform(-1, 1, (n, 5))
    Y = oracle.spmm_f64(rp, ci, val, X)
    assert np.allclose(Y, sp.csr_matrix((val, ci, rp), shape=(n, n)) @ X, rtol=1e-13, atol=1e-13)


def _brute_merge(rp):
    """sequential merge of (row ends, nonzeros): 1 = row end"""
    n = len(rp) - 1
    bits, x, y = [], 0, 0
    while y < n:
        if x < rp[y + 1]:
            bits.append(0); x += 1
        else:
            bits.append(1); y += 1
    return np.array(bits, np.uint8)


def test_merbit_reference_tiles(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 5, 100, 700):
        rp, ci = random_csr(rng, n, empty_frac=0.5, max_row=40)
        nnz = int(rp[-1])
        bits = _brute_merge(rp)
        assert len(bits) == n + nnz
        assert np.array_equal(np.flatnonzero(bits), rp[1:] + np.arange(n))   # bit (rp[r+1] + r)
        sigma = 14
        tx, ty, desc = oracle.merbit_ref_tiles(rp, nnz, sigma)
        flags = []
        for lane, d in enumerate(desc):
            cnt = min(sigma, n + nnz - lane * sigma)
            flags.extend(((int(d) >> (18 + j)) & 1) for j in range(cnt))
        assert np.array_equal(np.array(flags, np.uint8), bits)
        cum = np.concatenate([[0], np.cumsum(bits)])
        for t in range(len(tx) - 1):
            d0 = t * 32 * sigma
            assert ty[t] & 0x7fffffff == cum[d0] and tx[t] == d0 - cum[d0]
            assert oracle.merge_search(rp, nnz, d0) == (d0 - cum[d0], cum[d0])
            end = min((t + 1) * 32 * sigma, n + nnz)
            assert bool(ty[t] >> 31) == (cum[end] == cum[d0])
        assert tx[-1] == nnz and ty[-1] == n
        val = rng.uniform(-1, 1, nnz)
        x = rng.uniform(-1, 1, n)
        y = oracle.merbit_spmv(rp, ci, val, x, sigma)
        assert np.allclose(y, sp.csr_matrix((val, ci, rp), shape=(n, n)) @ x, rtol=1e-12, atol=1e-13)


def test_predictor_constants_match_model_files(oracle):
    for name, model in (("model_merbit_sampled.json", 0), ("model_csr_sampled.json", 1)):
        d = json.loads((GOLDEN / name).read_text())["ridge_raw_finite_features"]
        f = np.zeros(10)
        assert oracle.predict(f, model) == math.exp(d["intercept"])
        for i in range(10):
            e = np.zeros(10)
            e[i] = 1.0
            assert oracle.predict(e, model) == math.exp(d["intercept"] + d["coefficients"][i])


def test_predictor_matches_reference_headers(oracle):
    rows = json.loads((GOLDEN / "predict_ref.json").read_text())
    for r in rows:
        assert oracle.predict(r["feat"], 0) == r["merbit"]
        assert oracle.predict(r["feat"], 1) == r["csr"]


def test_predictor_preprocessed_equals_raw_when_complete(oracle):
    # model.cuh folds the scaler into the raw coefficients: both forms agree
    rng = np.random.default_rng(4)
    for _ in range(20):
        f = rng.dirichlet(np.ones(10))
        a = oracle.predict(f, 0)
        b = oracle.predict_preprocessed(f, np.zeros(10, np.uint8))
        assert abs(a - b) <= 1e-12 * a


def test_rcm_matches_reference(oracle):
    for g in json.loads((GOLDEN / "rcm_ref.json").read_text()):
        out = oracle.rcm(g["n"], np.array(g["rows"], np.int32), np.array(g["cols"], np.int32))
        assert out.tolist() == g["perm"]


def _csr_with_column_counts(counts):
    """n x n pattern whose column j holds counts[j] entries (rows 0..counts[j]-1)."""
    n = len(counts)
    rows = [[j for j in range(n) if counts[j] > i] for i in range(n)]
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], np.int32)
    return rp, ci


def test_degree_orders_hand_worked(oracle):
    # rows 0..4: [1,2], [2], [2,3,4], [], [0,2]; column counts 1,1,4,1,1 differ
    # from row counts 2,1,3,0,2 -- the reference counts row_ind, the columns of M
    rp = np.array([0, 2, 3, 6, 6, 8], np.int32)
    ci = np.array([1, 2, 2, 2, 3, 4, 0, 2], np.int32)
    assert oracle.order_degree(5, rp, ci).tolist() == [1, 2, 0, 3, 4]
    # column counts 2,3,0,4,1,6,7: degree order 6,5,3,1,0,4,2; DBG groups
    # ceil(log2(c+1)) = 2,2,0,3,1,3,3 -> 3,5,6,0,1,4,2
    rp, ci = _csr_with_column_counts([2, 3, 0, 4, 1, 6, 7])
    assert oracle.order_degree(7, rp, ci).tolist() == [4, 3, 6, 2, 5, 1, 0]
    assert oracle.order_degree(7, rp, ci, grouped=True).tolist() == [3, 4, 6, 0, 5, 1, 2]


def test_degree_orders_properties(oracle):
    """Along the new order the key never increases and equal keys keep id
    order; the DBG key is the bit length of the count (= ceil(log2(c+1)))."""
    rng = np.random.default_rng(12)
    rp, ci = random_csr(rng, 3000, empty_frac=0.2, max_row=60)
    n = len(rp) - 1
    cnt = np.bincount(ci, minlength=n)
    for grouped, key in ((False, cnt), (True, np.array([int(c).bit_length() for c in cnt]))):
        perm = oracle.order_degree(n, rp, ci, grouped)
        assert np.array_equal(np.sort(perm), np.arange(n))
        order = np.empty(n, np.int64)
        order[perm] = np.arange(n)
        k = key[order]
        assert np.all(np.diff(k) <= 0)
        same = np.diff(k) == 0
        assert np.all(np.diff(order)[same] > 0)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_cpu_spmv_baselines_match_scipy(oracle, threads):
    """The CPU baseline's threaded CSR SpMV (FP32 and FP64, nnz-balanced row
    ranges) against scipy, including empty rows and more threads than rows
    carry nonzeros."""
    import scipy.sparse as sp
    rng = np.random.default_rng(threads)
    a = sp.random(300, 250, density=0.03, format="csr", random_state=threads)
    a.data = rng.uniform(-1, 1, a.nnz)
    x = rng.uniform(-1, 1, 250)
    rp, ci = a.indptr.astype(np.int32), a.indices.astype(np.int32)
    y64, _ = oracle.spmv_f64_threads(rp, ci, a.data, x, threads, 2)
    assert np.allclose(y64, a @ x, rtol=1e-12, atol=1e-12)
    y32, _ = oracle.spmv_f32_threads(rp, ci, a.data.astype(np.float32), x.astype(np.float32), threads, 2)
    assert np.allclose(y32, a @ x, rtol=1e-4, atol=1e-5)
