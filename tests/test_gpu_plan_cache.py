"""The on-disk plan cache: the device fingerprint equals the host one, a
cached pGOrder gives the same permutation and matrix as pGOrder itself, stale
or corrupt entries are replaced, tuned MERBIT options travel with the entry,
and the serving loop's uploads find their entries."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_csr, rmat_csr

pytestmark = pytest.mark.gpu


def test_device_fingerprint_equals_host(ctx, mr):
    rng = np.random.default_rng(8)
    cases = [random_csr(rng, 500, empty_frac=0.3, max_row=30), rmat_csr(14, 16, seed=1),
             (np.zeros(1001, np.int32), np.zeros(0, np.int32)), rmat_csr(17, 16, seed=2)]
    for rp, ci in cases:
        mat = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
        assert mat.hash() == mr.csr_hash(rp, ci)
        mat.close()


def _plain(ctx, mr, rp, ci, val, W):
    m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    perm, _ = m.reorder_pgorder(W, 5)
    out = perm, m.export()
    m.close()
    return out


def test_cached_pgorder_hit_and_miss(ctx, mr, tmp_path):
    rp, ci = rmat_csr(14, 16, seed=1)
    val = np.random.default_rng(3).uniform(-1, 1, len(ci))
    want_perm, (wrp, wci, wval) = _plain(ctx, mr, rp, ci, val, 1024)
    for expect in (0, 1, 1):
        m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
        perm, st, hit = m.reorder_pgorder_cached(tmp_path, 1024, 5)
        assert hit == expect
        assert np.array_equal(perm, want_perm) and np.array_equal(m.perm(), want_perm)
        grp, gci, gval = m.export()
        assert np.array_equal(grp, wrp) and np.array_equal(gci, wci) and np.array_equal(gval, wval)
        m.close()
    files = sorted(p.name for p in tmp_path.iterdir())
    assert files == [f"{mr.csr_hash(rp, ci):016x}-pgorder-w1024-k5.mrplan"]
    # another width is another entry
    m = mr.Mat.from_csr(ctx, rp, ci, val, store=mr.F32)
    perm, _, hit = m.reorder_pgorder_cached(tmp_path, 64, 5)
    assert hit == 0 and np.array_equal(perm, _plain(ctx, mr, rp, ci, val, 64)[0])
    m.close()
    assert len(list(tmp_path.iterdir())) == 2


def test_corrupt_entries_are_replaced(ctx, mr, tmp_path):
    rp, ci = rmat_csr(13, 16, seed=5)
    want, _ = _plain(ctx, mr, rp, ci, None, 1024)
    m = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    assert m.reorder_pgorder_cached(tmp_path, 1024, 5)[2] == 0
    m.close()
    path = next(tmp_path.iterdir())
    good = path.read_bytes()
    bad_perm = bytearray(good)
    bad_perm[88 + 4 * 17] ^= 1                           # one permutation entry
    swapped = bytearray(good)
    a, b = 88 + 4 * 3, 88 + 4 * 9                        # two entries trade places: still a permutation
    swapped[a:a + 4], swapped[b:b + 4] = good[b:b + 4], good[a:a + 4]
    for blob in (bytes(bad_perm), bytes(swapped), good[:-4], good + b"\0", b"MRPLAN01"):
        path.write_bytes(blob)
        m = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
        perm, _, hit = m.reorder_pgorder_cached(tmp_path, 1024, 5)
        assert hit == 2 and np.array_equal(perm, want)
        m.close()
        assert path.read_bytes() == good


def test_saved_options_travel_with_the_entry(ctx, mr, tmp_path):
    rp, ci = rmat_csr(14, 16, seed=6)
    m = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    with pytest.raises(mr.MRError):
        m.plan_save()                                     # not reordered through the cache
    m.reorder_pgorder_cached(tmp_path, 1024, 5)
    m.merbit_options(4)
    m.merbit_configure(16)              # rebuilds the block plan, so before the x-cache plan
    m.merbit_xc_configure(16, 8, 1000)
    m.plan_save()
    opts = lambda mm: (mm.merbit_policy(), mm.merbit_info()["sigma"],  # noqa: E731
                       tuple(mm.merbit_xc_info()[k] for k in ("items_per_lane", "waves", "cached_cols")))
    want = opts(m)
    assert want == (4, 16, (16, 8, 1000))
    m.permute(np.arange(m.info()[0], dtype=np.int32)[::-1].copy())
    with pytest.raises(mr.MRError):
        m.plan_save()                                     # relabelled since
    m.close()
    m = mr.Mat.from_csr(ctx, rp, ci, store=mr.F32)
    assert opts(m) != want
    _, _, hit = m.reorder_pgorder_cached(tmp_path, 1024, 5)
    assert hit == 1
    assert opts(m) == want
    x = np.random.default_rng(1).uniform(-1, 1, m.info()[0]).astype(np.float32)
    xd, yd = ctx.upload(x), ctx.empty(m.info()[0], np.float32)
    m.spmv(xd, yd)
    ctx.sync()
    m.close()


def test_serve_uploads_use_the_cache(ctx, mr, tmp_path):
    srv = mr.Server(ctx, verify=0)
    rp, ci = rmat_csr(10, 8, seed=2)
    first = mr.Mat.from_csr(ctx, rp, ci, store=mr.BF16)   # the first matrix fixes the served dtype
    srv.add(first)
    srv.set_plan_cache(tmp_path)
    hashes = []
    for _ in range(2):
        st, mid, ms = srv.upload_wait(srv.upload_rmat(13, 16, seed=11))
        assert st == 0
        r = srv.submit(mid, 4, 99)
        st, _, h = srv.wait(r)
        assert st == 0
        hashes.append(h)
    assert srv.plan_cache_stats() == (1, 1)
    assert hashes[0] == hashes[1]
    srv.set_plan_cache(None)
    st, mid, _ = srv.upload_wait(srv.upload_rmat(13, 16, seed=11))
    assert st == 0 and srv.plan_cache_stats() == (1, 1)
    srv.close()
    first.close()
