"""Reuse-distance features and the predictor on the device against the
oracle: sampled HyperLogLog RD (sample positions, status, estimate bits and
bins identical), exact RD, locality_predict."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rmat_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,ef,seed", [(14, 16, 1), (12, 4, 9)])
def test_sampled_rd_bit_exact(ctx, mr, oracle, scale, ef, seed):
    rp, ci = rmat_csr(scale, ef, seed=seed)
    mat = mr.Mat.from_csr(ctx, rp, ci)
    p = mr.rd_params()
    est, st, pos = mat.rd_sampled_raw(p)
    ref = oracle.rd_sampled(ci, samples=p.samples, sector=32, rd_cap=589824, hll_p=12, max_scan=0,
                            seed=oracle.hll_seed(0, 0))
    assert np.array_equal(pos, ref["pos"])
    assert np.array_equal(st, ref["status"])
    assert np.array_equal(est.view(np.uint32), ref["est"].view(np.uint32))
    feat, stats = mat.rd_features(mr.RD_SAMPLED, p)
    assert np.array_equal(feat, ref["prob"])
    assert [stats["found"], stats["cold_samples"], stats["large"], stats["unresolved"]] == ref["stat4"].tolist()
    mat.close()


def test_sampled_rd_small_cap_and_max_scan(ctx, mr, oracle):
    # long gaps between reuses with a small cap exercise status 2; a short
    # max_scan exercises status 3
    rng = np.random.default_rng(71)
    n = 200000
    stream = np.concatenate([rng.permutation(n)[:60000], rng.permutation(n)[:60000]]).astype(np.int32)
    rp = np.zeros(len(stream) + 1, np.int32)
    rp[1:] = np.arange(1, len(stream) + 1)
    rp_sq = np.zeros(n + 1, np.int32)
    rp_sq[1:len(stream) + 1] = np.arange(1, len(stream) + 1)
    rp_sq[len(stream) + 1:] = len(stream)
    mat = mr.Mat.from_csr(ctx, rp_sq, stream, n_cols=n)
    edges = [1, 16, 64, 128]
    for cap, ms in ((200, 0), (0, 3000)):
        p = mr.rd_params(sector_size=1, samples=4000, rd_cap=cap, max_scan=ms, bin_edges=edges)
        est, st, pos = mat.rd_sampled_raw(p)
        ref = oracle.rd_sampled(stream, samples=4000, sector=1, rd_cap=cap, hll_p=12, max_scan=ms,
                                seed=oracle.hll_seed(0, 0), edges=np.array(edges, np.int32))
        assert np.array_equal(pos, ref["pos"]) and np.array_equal(st, ref["status"])
        assert np.array_equal(est.view(np.uint32), ref["est"].view(np.uint32))
        assert (st == (2 if cap else 3)).any()
    mat.close()


def test_exact_rd_and_predict(ctx, mr, oracle):
    rp, ci = rmat_csr(14, 16, seed=1)
    n = len(rp) - 1
    mat = mr.Mat.from_csr(ctx, rp, ci)
    feat, st = mat.rd_features(mr.RD_EXACT)
    counts, reuses, prob = oracle.rd_exact(ci, n)
    assert np.array_equal(feat, prob) and st["reuses"] == reuses
    sp, f10, _ = mat.locality_predict(mr.MODEL_MERBIT)
    ref = oracle.rd_sampled(ci, seed=oracle.hll_seed(0, 0))
    assert np.array_equal(f10, ref["prob"])
    assert sp == oracle.predict(ref["prob"], 0)
    mat.close()


@pytest.mark.parametrize("model", [2, 3])
def test_locality_predict_mi355x_models(ctx, mr, oracle, model):
    """The MI355X models compute their features with their own bin edges and
    cap; the features equal the oracle's sampled RD under those parameters."""
    rp, ci = rmat_csr(14, 16, seed=1)
    mat = mr.Mat.from_csr(ctx, rp, ci)
    p = mr.rd_params(model)
    sp, f10, _ = mat.locality_predict(model)
    edges = np.array([p.bin_edges[i] for i in range(p.n_edges)], np.int32)
    ref = oracle.rd_sampled(ci, rd_cap=p.rd_cap, seed=oracle.hll_seed(0, 0), edges=edges)
    assert np.array_equal(f10, ref["prob"])
    assert sp == mr.predict(f10, model)
    mat.close()
