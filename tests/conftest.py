"""Shared fixtures.  `-m "not gpu"` tests run on the CPU (oracle vs golden
vectors, host logic, library exports); `-m gpu` tests call the HIP library
through the C-ABI on a real device."""
from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "matrix-reorder_amd"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def mr():
    import mrpy
    mrpy.lib()
    return mrpy


@pytest.fixture(scope="session")
def ctx(mr):
    c = mr.Ctx(0)
    yield c
    c.close()


def random_csr(rng, n, density=None, max_row=None, empty_frac=0.3, n_cols=None):
    """Small random CSR with sorted, unique columns and some empty rows."""
    n_cols = n if n_cols is None else n_cols
    rows = []
    for i in range(n):
        if rng.random() < empty_frac:
            rows.append(np.zeros(0, np.int32))
            continue
        k = int(rng.integers(1, (max_row or max(2, n_cols // 3)) + 1))
        k = min(k, n_cols)
        rows.append(np.sort(rng.choice(n_cols, k, replace=False)).astype(np.int32))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
    return rp, ci


def rmat_csr(scale, ef=16, seed=1, abc=(0.57, 0.19, 0.19)):
    import mrpy
    return mrpy.gen_rmat(scale, ef, abc, seed)


@pytest.fixture(autouse=True)
def _announce_gpu_case(request):
    """Print (and flush) the id of every GPU test before it touches the device."""
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def csr_rows(rows, n_cols):
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.asarray(r, np.int32) for r in rows]) if rows else np.zeros(0, np.int32)
    return rp, ci.astype(np.int32)


def spmv_cases():
    """(name, rp, ci, n_cols) parity cases: the small R-MAT of the plain-method
    configuration, ragged and empty rows, rows longer than a MERBIT block,
    degenerate sizes."""
    rng = np.random.default_rng(7)
    cases = []
    rp, ci = rmat_csr(14, 16, seed=1)
    cases.append(("rmat_s14", rp, ci, 1 << 14))
    rp, ci = random_csr(rng, 1000, empty_frac=0.5, max_row=60)
    cases.append(("random_1000", rp, ci, 1000))
    cases.append(("all_empty", np.zeros(5001, np.int32), np.zeros(0, np.int32), 5000))
    cases.append(("one_row_empty", np.zeros(2, np.int32), np.zeros(0, np.int32), 1))
    cases.append(("one_row_one_nnz", np.array([0, 1], np.int32), np.array([0], np.int32), 1))
    long_rows = [np.sort(rng.choice(30000, 29000, replace=False)) if i == 3 else
                 (np.sort(rng.choice(30000, 7000, replace=False)) if i % 17 == 5 else
                  np.sort(rng.choice(30000, int(rng.integers(0, 9)), replace=False))) for i in range(60)]
    rp, ci = csr_rows(long_rows, 30000)
    cases.append(("long_rows", rp, ci, 30000))
    import mrpy
    rp, ci = mrpy.gen_stencil27(12, 10, 9)
    cases.append(("stencil27", rp, ci, len(rp) - 1))
    return cases
