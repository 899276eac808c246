"""Plan-cache key on the host: the fingerprint of a CSR pattern follows its
documented formula (a numpy restatement), is stable, and changes with every
kind of edit of the pattern."""
from __future__ import annotations

import numpy as np

from conftest import random_csr, rmat_csr

M64 = (1 << 64) - 1


def _mix(z):
    z = np.asarray(z, np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _sum(a, salt):
    a = np.asarray(a, np.int32).view(np.uint32).astype(np.uint64)
    i = np.arange(len(a), dtype=np.uint64)
    s = np.uint64(salt)
    with np.errstate(over="ignore"):
        t = _mix(s + i * np.uint64(0x9E3779B97F4A7C15) + _mix(a ^ s))
        return int(np.sum(t, dtype=np.uint64))


def fingerprint(rp, ci, n_cols):
    n = len(rp) - 1
    with np.errstate(over="ignore"):
        h = int(_mix(np.uint64(0x4D52504C414E3031 ^ n)))
        for v in (n_cols, len(ci), _sum(rp, 0x726F775F70747273), _sum(ci, 0x636F6C5F69647873)):
            h = int(_mix(np.uint64((h ^ v) & M64)))
    return h


def test_fingerprint_formula(mr):
    rng = np.random.default_rng(5)
    for n in (1, 7, 500, 5000):
        rp, ci = random_csr(rng, n, empty_frac=0.3, max_row=20)
        assert mr.csr_hash(rp, ci) == fingerprint(rp, ci, n)
    rp, ci = rmat_csr(12, 8, seed=3)
    assert mr.csr_hash(rp, ci) == fingerprint(rp, ci, len(rp) - 1)
    assert mr.csr_hash(np.zeros(4, np.int32), np.zeros(0, np.int32)) == fingerprint(np.zeros(4), [], 3)


def test_fingerprint_sensitivity(mr):
    rp, ci = rmat_csr(12, 8, seed=4)
    n = len(rp) - 1
    base = mr.csr_hash(rp, ci)
    assert mr.csr_hash(rp.copy(), ci.copy()) == base
    seen = {base}
    c2 = ci.copy()
    c2[100] = (c2[100] + 1) % n                 # one column index
    seen.add(mr.csr_hash(rp, c2))
    c3 = ci.copy()
    c3[[10, 11]] = c3[[11, 10]]                  # two entries trade places
    seen.add(mr.csr_hash(rp, c3))
    r2 = rp.copy()
    k = int(np.flatnonzero(np.diff(rp) > 1)[0])
    r2[k + 1] -= 1                               # a nonzero moves to the next row
    seen.add(mr.csr_hash(r2, ci))
    seen.add(mr.csr_hash(rp, ci, n_cols=n + 1))  # one more column
    assert len(seen) == 5
