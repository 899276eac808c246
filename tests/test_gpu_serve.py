"""Serving loop (mr_serve_*) on the device: results of requests batched with
others equal the same requests served alone and the single-matrix entry points
(bit-exact without coalescing; within the FP32 tolerance 1e-5, condition-aware,
with coalescing), uploads run pGOrder and are admitted or evicted by the memory
budget, and the admission / timeout / cancel paths report their status."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rmat_csr

pytestmark = pytest.mark.gpu

OK, INVALID, ADMISSION, TIMEOUT, CANCELLED = 0, 1, 8, 11, 12


def resident(ctx, mr, store=None):
    store = mr.BF16 if store is None else store
    rng = np.random.default_rng(12)
    mats, host = [], []
    shapes = [rmat_csr(11, 16, seed=21), rmat_csr(9, 8, seed=22), mr.gen_banded(5000, 40, 10, seed=23),
              mr.gen_banded(800, 8, 4, seed=24), rmat_csr(12, 4, seed=25)]
    for rp, ci in shapes:
        n = len(rp) - 1
        val = rng.uniform(-1, 1, len(ci))
        m = mr.Mat.from_csr(ctx, rp, ci, val, n_cols=n, store=store)
        _, _, vstored = m.export()
        mats.append(m)
        host.append((rp, ci, vstored, n))
    return mats, host


def stored_f64(mr, v, store):
    if store == mr.BF16:
        return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return np.asarray(v, np.float64)


def alone_library(ctx, mr, mat, n, k, seed):
    """The same request through mr_spmv / mr_spmm (MERBIT) on its own."""
    x = mr.serve_rhs(seed, n, k)
    mat.merbit_options(1)   # the workgroup-tile kernel the server's batches run
    xd = ctx.upload(x)
    yd = ctx.empty((n, k), np.float32)
    if k == 1:
        mat.spmv(xd, yd, mr.MERBIT)
    else:
        mat.spmm(k, xd, yd, mr.MERBIT)
    ctx.sync()
    return yd.to_host()


def request_mix(n_mats, count, seed):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, n_mats)), int(rng.choice([1, 4, 16])), int(rng.integers(1, 2**62)))
            for _ in range(count)]


def test_rhs_generator_matches_host(ctx, mr):
    """The device right-hand sides are the host generator's values: a request
    on an identity matrix returns X itself."""
    n = 3000
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    eye = mr.Mat.from_csr(ctx, rp, ci, np.ones(n), n_cols=n, store=mr.F32)
    srv = mr.Server(ctx, verify=2, coalesce=0)
    mid = srv.add(eye)
    for k, seed in ((1, 5), (4, 6), (16, 7)):
        r = srv.submit(mid, k, seed)
        st, lat, h = srv.wait(r)
        assert st == OK and lat > 0
        y = srv.result(r, n, k)
        assert np.array_equal(y, mr.serve_rhs(seed, n, k))
        assert h == mr.serve_hash(y)
    srv.close()
    eye.close()


def test_mix_equals_alone_bit_exact(ctx, mr):
    mats, host = resident(ctx, mr)
    srv = mr.Server(ctx, verify=1, coalesce=0, slots=3, max_jobs=16)
    ids = [srv.add(m) for m in mats]
    mix = request_mix(len(ids), 120, seed=3)
    reqs = [srv.submit(ids[m], k, s) for m, k, s in mix]     # queued together: batched with each other
    mixed = [srv.wait(r) for r in reqs]
    assert all(st == OK for st, _, _ in mixed)
    for i in range(0, len(mix), 7):
        m, k, s = mix[i]
        st, _, h_alone = srv.wait(srv.submit(ids[m], k, s))  # alone: nothing else queued
        assert st == OK and h_alone == mixed[i][2], f"request {i}"
        y = alone_library(ctx, mr, mats[m], host[m][3], k, s)
        assert mr.serve_hash(y) == mixed[i][2], f"request {i} vs library"
    srv.close()


def test_mix_coalesced_within_tolerance(ctx, mr, oracle):
    mats, host = resident(ctx, mr)
    srv = mr.Server(ctx, verify=2, coalesce=1, slots=2)
    ids = [srv.add(m) for m in mats]
    mix = request_mix(len(ids), 60, seed=4)
    reqs = [srv.submit(ids[m], k, s) for m, k, s in mix]
    for i, r in enumerate(reqs):
        st, _, _ = srv.wait(r)
        assert st == OK
        m, k, s = mix[i]
        rp, ci, v, n = host[m]
        got = srv.result(r, n, k)
        alone = alone_library(ctx, mr, mats[m], n, k, s)
        x = mr.serve_rhs(s, n, k).astype(np.float64)
        vals = stored_f64(mr, v, mr.BF16)
        for c in sorted({0, k - 1}):
            yhat, absrow = oracle.spmv_f64(rp, ci, vals, x[:, c])
            for y in (got[:, c], alone[:, c]):
                worst, l2, _ = oracle.spmv_parity(y, yhat, absrow)
                assert worst <= 1e-5 and l2 <= 1e-5, (i, c, worst, l2)
            denom = np.where(absrow > 0, absrow, 1.0)
            assert np.max(np.abs(got[:, c] - alone[:, c]) / denom) <= 1e-5
    srv.close()


def test_run_load_with_uploads(ctx, mr):
    mats, host = resident(ctx, mr)
    srv = mr.Server(ctx, verify=0, coalesce=1)
    for m in mats:
        srv.add(m)
    rep = srv.run_load(rate=3000.0, n_requests=400, seed=9, ks=(1, 4, 16), upload_every=100, upload_scale=12)
    assert rep["uploads_admitted"] == 4 and rep["uploads_rejected"] == 0
    assert rep["completed"] == 396 and rep["rejected"] == 0 and rep["timed_out"] == 0
    assert 0 < rep["p50_us"] <= rep["p90_us"] <= rep["p99_us"] <= rep["max_us"]
    assert rep["gflops"] > 0 and rep["batches"] >= 1 and rep["upload_ms_mean"] > 0
    srv.close()


def test_upload_admission_and_eviction(ctx, mr):
    mats, host = resident(ctx, mr)
    # a budget that holds one scale-12 upload but not two
    srv = mr.Server(ctx, hbm_budget=2_500_000)
    srv.add(mats[0])
    st, first, ms = srv.upload_wait(srv.upload_rmat(12, 16, seed=1))
    assert st == OK and first >= 1 and ms > 0
    r = srv.submit(first, 4, 77)
    assert srv.wait(r)[0] == OK
    st, second, _ = srv.upload_wait(srv.upload_rmat(12, 16, seed=2))
    assert st == OK and second == first + 1
    with pytest.raises(mr.MRError) as e:       # the first upload was evicted
        srv.submit(first, 1, 1)
    assert e.value.code == INVALID
    assert srv.wait(srv.submit(second, 1, 3))[0] == OK
    st, _, _ = srv.upload_wait(srv.upload_rmat(15, 16, seed=3))   # larger than the whole budget
    assert st == ADMISSION
    srv.close()


def test_request_larger_than_arena_is_refused(ctx, mr):
    n = 300_000
    rp, ci = mr.gen_banded(n, 4, 3, seed=5)
    big = mr.Mat.from_csr(ctx, rp, ci, None, n_cols=n, store=mr.BF16)
    srv = mr.Server(ctx, slot_bytes=24 << 20, slots=1)
    mid = srv.add(big)
    st_big = srv.wait(srv.submit(mid, 16, 1))[0]     # X + Y = 38 MB > arena
    st_small = srv.wait(srv.submit(mid, 1, 2))[0]
    assert st_big == ADMISSION and st_small == OK
    srv.close()
    big.close()


def test_timeout_and_cancel(ctx, mr):
    mats, host = resident(ctx, mr)
    srv = mr.Server(ctx, timeout_ms=1e-9)
    mid = srv.add(mats[0])
    assert srv.wait(srv.submit(mid, 1, 1))[0] == TIMEOUT
    srv.close()
    srv = mr.Server(ctx, slots=1)
    mid = srv.add(mats[0])
    reqs = [srv.submit(mid, 16, i) for i in range(50)]
    cancelled = 0
    try:
        srv.cancel(reqs[-1])
        cancelled = 1
    except mr.MRError as e:   # already dispatched
        assert e.code == INVALID
    st = srv.wait(reqs[-1])[0]
    assert st == (CANCELLED if cancelled else OK)
    assert all(srv.wait(r)[0] == OK for r in reqs[:-1])
    srv.close()


def test_server_argument_errors(ctx, mr):
    mats, host = resident(ctx, mr)
    with pytest.raises(mr.MRError):
        mr.Server(ctx, slots=0)
    with pytest.raises(mr.MRError):
        mr.Server(ctx, sigma=12)
    srv = mr.Server(ctx)
    with pytest.raises(mr.MRError) as e:
        srv.upload_rmat(10)            # no matrix yet: the dtype is not fixed
    assert e.value.code == INVALID
    srv.add(mats[0])
    with pytest.raises(mr.MRError) as e:
        srv.submit(0, 65, 1)
    assert e.value.code == 5
    with pytest.raises(mr.MRError) as e:
        srv.submit(7, 1, 1)
    assert e.value.code == INVALID
    f32 = mr.Mat.from_csr(ctx, host[1][0], host[1][1], None, n_cols=host[1][3], store=mr.F32)
    with pytest.raises(mr.MRError) as e:
        srv.add(f32)
    assert e.value.code == 5
    srv.close()
    f32.close()


@pytest.mark.parametrize("policy", [4, 8, 9])
def test_tuned_routes_match_library(ctx, mr, policy):
    """Matrices tuned to the lane-padded variant (policy bit 2) are served
    through its SpMV / SpMM kernels, those tuned to the column-panel variant
    (bit 3) through the panel kernels for single vectors and the lane-padded
    SpMM for several; without coalescing every result equals the library's
    mr_spmv / mr_spmm on the same matrix bit for bit."""
    rp, ci = rmat_csr(12, 16, seed=31)
    n = len(rp) - 1
    rng = np.random.default_rng(9)
    big = mr.Mat.from_csr(ctx, rp, ci, rng.uniform(-1, 1, len(ci)), n_cols=n, store=mr.BF16)
    big.merbit_options(policy)
    mats, host = resident(ctx, mr)
    srv = mr.Server(ctx, verify=1, coalesce=0, slots=2, max_jobs=8)
    bid = srv.add(big)
    ids = [srv.add(m) for m in mats]
    mix = [(bid, k, 100 + i) for i, k in enumerate([1, 4, 16, 3, 8, 1, 16])]
    mix += [(ids[i % len(ids)], 4, 500 + i) for i in range(6)]
    reqs = [srv.submit(m, k, sd) for m, k, sd in mix]
    got = [srv.wait(r) for r in reqs]
    assert all(st == OK for st, _, _ in got)
    for (mid, k, sd), (_, _, h) in zip(mix, got):
        if mid != bid:
            continue
        x = mr.serve_rhs(sd, n, k)
        ldx = 1 if k == 1 else -(-k // 4) * 4       # the server's X layout
        xp = np.zeros((n, ldx), np.float32)
        xp[:, :k] = x
        xd = ctx.upload(xp)
        yd = ctx.empty((n, k), np.float32)
        if k == 1:
            big.spmv(xd, yd, mr.MERBIT)
        else:
            big.spmm(k, xd, yd, mr.MERBIT, ldx=ldx, ldy=k)
        ctx.sync()
        assert mr.serve_hash(yd.to_host()) == h, (k, sd)
    srv.close()
    big.close()


def test_concurrent_adds_of_two_dtypes(ctx, mr):
    """Two threads add matrices of different storage dtypes to one server at
    the same time: exactly one add succeeds and the other is refused
    (MR_ERR_UNSUPPORTED), so the fused kernels never read a matrix as the
    wrong dtype."""
    import threading
    rp, ci = rmat_csr(12, 16, seed=61)
    n = len(rp) - 1
    for rep in range(8):
        a = mr.Mat.from_csr(ctx, rp, ci, None, n_cols=n, store=mr.F32)
        b = mr.Mat.from_csr(ctx, rp, ci, None, n_cols=n, store=mr.BF16)
        srv = mr.Server(ctx)
        codes = {}
        gate = threading.Barrier(2)

        def add(name, m):
            gate.wait()
            try:
                srv.add(m)
                codes[name] = 0
            except mr.MRError as e:
                codes[name] = e.code

        ts = [threading.Thread(target=add, args=("a", a)), threading.Thread(target=add, args=("b", b))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert sorted(codes.values()) == [0, 5], (rep, codes)
        srv.close()
        a.close()
        b.close()


def test_light_requests_first_and_heavy_deadlines(ctx, mr):
    """Requests streaming 32 MiB or more (matrix and vectors) are heavy: a
    light request queued behind twenty heavy ones is dispatched at the next
    free slot (FIFO would finish it last), and with a deadline a heavy request
    predicted to finish late (streamed bytes / the learned rate) is shed at
    once, while a generous deadline serves it."""
    n = 1_000_000
    rp, ci = mr.gen_banded(n, 64, 6, seed=31)          # 6 M nonzeros: 36 MB of BF16 items
    heavy = mr.Mat.from_csr(ctx, rp, ci, None, n_cols=n, store=mr.BF16)
    mats, _ = resident(ctx, mr)
    srv = mr.Server(ctx, slots=1, max_jobs=1, coalesce=0)
    hid, lid = srv.add(heavy), srv.add(mats[1])
    hreqs = [srv.submit(hid, 4, i) for i in range(20)]
    lreq = srv.submit(lid, 1, 99)
    st_l, lat_l, _ = srv.wait(lreq)
    res = [srv.wait(r) for r in hreqs]
    assert st_l == OK and all(r[0] == OK for r in res)
    assert lat_l < 0.5 * res[-1][1], (lat_l, res[-1][1])
    srv.close()
    srv = mr.Server(ctx, timeout_ms=1e-3)
    hid = srv.add(heavy)
    assert srv.wait(srv.submit(hid, 1, 1))[0] == TIMEOUT
    srv.close()
    srv = mr.Server(ctx, timeout_ms=60_000.0)
    hid = srv.add(heavy)
    reqs = [srv.submit(hid, k, i) for i, k in enumerate((1, 4, 16))]
    assert all(srv.wait(r)[0] == OK for r in reqs)
    srv.close()
    heavy.close()
    for m in mats:
        m.close()
