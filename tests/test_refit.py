"""The MI355X predictor refit (mrpy/refit.py): the committed model header is
what the committed training set yields, the C-ABI models agree with the
fitted scikit-learn pipelines, and the local_swap perturbation keeps its
contract (src/class.cu:236-301)."""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "matrix-reorder_amd"
CSV = PKG / "model" / "mi355x_rd_train.csv"
HEADER = PKG / "csrc" / "model_mi355x.hpp"

pytest.importorskip("sklearn")


@pytest.fixture(scope="module")
def refit():
    from mrpy import refit as r
    return r


@pytest.fixture(scope="module")
def fitted(refit):
    return refit.fit_shared(str(CSV))


def test_header_is_reproduced_from_the_training_set(refit, fitted):
    _, chosen = fitted
    assert refit.header_text(chosen, CSV.name) == HEADER.read_text()


def _features(refit, rows, prefix):
    return np.array([[float(r[f"{prefix}{i + 1}"]) for i in range(refit.FEATURES)] for r in rows])


@pytest.mark.parametrize("kernel,model", [("merbit", 2), ("csr", 3)])
def test_c_abi_model_equals_fitted_pipeline(refit, fitted, mr, kernel, model):
    """mr_predict against the pipeline's own predict (imputer -> scaler ->
    ridge): a wrong sign or scale in the raw-coefficient export, or a constant
    copied into the wrong slot, changes the prediction."""
    report, chosen = fitted
    fs = chosen[kernel][0]
    prefix = "q" if fs == "mi355x_edges" else "r"
    rows = refit.load_rows(str(CSV))
    X = _features(refit, rows, prefix)
    y = np.log(np.array([float(r[f"speedup_{kernel}"]) for r in rows]))
    pipe = refit._pipeline(report["alpha"]).fit(X, y)
    want = np.exp(pipe.predict(X))
    got = np.array([mr.predict(x, model) for x in X[:64]])
    np.testing.assert_allclose(got, want[:64], rtol=1e-12)


def test_rd_params_of_each_model(refit, mr, fitted):
    ref = mr.rd_params()
    for model in (mr.MODEL_MERBIT, mr.MODEL_CSR):
        p = mr.rd_params(model)
        assert [p.bin_edges[i] for i in range(9)] == [ref.bin_edges[i] for i in range(9)]
        assert p.rd_cap == ref.rd_cap
    # the MI355X models use the edge set the refit chose (by cross-validated
    # MAE) for both kernels
    _, chosen = fitted
    if chosen["merbit"][0] == "mi355x_edges":
        edges, cap = refit.MI355X_EDGES, refit.MI355X_EDGES[-1]
    else:
        edges, cap = tuple(ref.bin_edges[i] for i in range(9)), ref.rd_cap
    for model in (mr.MODEL_MERBIT_MI355X, mr.MODEL_CSR_MI355X):
        p = mr.rd_params(model)
        assert p.n_edges == 9
        assert tuple(p.bin_edges[i] for i in range(9)) == edges
        assert p.rd_cap == cap
        assert p.samples == ref.samples and p.hll_p == ref.hll_p and p.sector_size == ref.sector_size
    with pytest.raises(mr.MRError):
        mr.rd_params(7)
    with pytest.raises(mr.MRError):
        mr.predict(np.zeros(10), 4)


def test_local_swap_is_a_permutation(refit):
    for n, (rho, near, win) in [(5000, p) for p in refit.PARA] + [(1, (0.9, 0, 10)), (0, (0.5, 1, 2))]:
        p = refit.local_swap(n, rho, near, win, 11)
        assert np.array_equal(np.sort(p), np.arange(n))


def test_local_swap_single_swap_distance(refit):
    """With one swap (rho * n rounds to 1) exactly two entries change, at a
    distance inside [near, window]."""
    n = 1000
    for seed in range(40):
        for near, win in ((3, 7), (0, 1), (64, 256), (500, 600)):
            p = refit.local_swap(n, 0.001, near, win, seed)
            moved = np.flatnonzero(p != np.arange(n))
            if len(moved) == 0:          # j == i (possible only when near == 0)
                assert near == 0
                continue
            assert len(moved) == 2
            d = int(moved[1] - moved[0])
            assert near <= d <= win
            assert p[moved[0]] == moved[1] and p[moved[1]] == moved[0]


def test_local_swap_degenerate_intervals(refit):
    n = 300
    assert np.array_equal(refit.local_swap(n, 0.0, 1, 10, 3), np.arange(n))
    # near beyond the matrix: no index has a partner
    assert np.array_equal(refit.local_swap(n, 0.5, 1000, 2000, 3), np.arange(n))
    assert refit.llround(0.5) == 1 and refit.llround(2.4999) == 2
