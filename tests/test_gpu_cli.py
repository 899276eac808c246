"""The MatrixMarket loader (BASEGRAPH, src/class.cu:5-216) and the C++
driver lib/mr_main over include/mr_graph.hpp (src/main.cu's workflow): the
multiplied matrix is the transpose of the file's COO with symmetric entries
doubled, every reorder type runs end to end and its spmv matches the
load-time product, and the log records have the reference's fields."""
from __future__ import annotations

import os
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import rmat_csr

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
MR_MAIN = ROOT / "matrix-reorder_amd" / "lib" / "mr_main"


def write_mtx(path, rows, cols, vals, n, symmetric=False, pattern=False, basezero=False):
    kind = "pattern" if pattern else "real"
    sym = "symmetric" if symmetric else "general"
    head = f"%%MatrixMarket matrix coordinate {kind} {sym}" + (" basezero" if basezero else "")
    off = 0 if basezero else 1
    lines = [head, "% written by the test", f"{n} {n} {len(rows)}"]
    for r, c, v in zip(rows, cols, vals):
        lines.append(f"{r + off} {c + off}" if pattern else f"{r + off} {c + off} {v!r}")
    path.write_text("\n".join(lines) + "\n")


def expected_csr(rows, cols, vals, n, symmetric):
    """Rows of the multiplied matrix = file columns; symmetric files add the
    mirrored entry of every line (diagonal ones too); columns sorted, equal
    columns in file order."""
    src, dst, vv = list(cols), list(rows), list(vals)
    if symmetric:
        src += list(rows)
        dst += list(cols)
        vv += list(vals)
    order = sorted(range(len(src)), key=lambda e: (src[e], dst[e], e))
    rp = np.zeros(n + 1, np.int32)
    for e in src:
        rp[e + 1] += 1
    return np.cumsum(rp).astype(np.int32), np.array([dst[e] for e in order], np.int32), \
        np.array([vv[e] for e in order], np.float64)


@pytest.mark.parametrize("symmetric,pattern,basezero,ones", [(False, False, False, False), (True, False, False, False),
                                                             (True, True, False, True), (False, False, True, True)])
def test_mtx_loader(ctx, mr, tmp_path, symmetric, pattern, basezero, ones):
    rng = np.random.default_rng(5)
    n = 300
    m = 2000
    rows = rng.integers(0, n, m).tolist()
    cols = rng.integers(0, n, m).tolist()
    vals = rng.uniform(-1, 1, m).round(6).tolist()
    path = tmp_path / "a.mtx"
    write_mtx(path, rows, cols, vals, n, symmetric, pattern, basezero)
    mat = mr.Mat.load_mtx(ctx, path, store=mr.F64, force_ones=ones)
    rp, ci, val = mat.export()
    erp, eci, ev = expected_csr(rows, cols, vals, n, symmetric)
    assert np.array_equal(rp, erp) and np.array_equal(ci, eci)
    assert np.array_equal(val, np.ones_like(ev) if (ones or pattern) else ev)
    mat.close()


def run_main(tmp_path, gid, typ, *extra):
    env = dict(os.environ, MR_CONFIG=str(tmp_path / "config.txt"), MR_LOG=str(tmp_path / "log.txt"), MR_DEVICE="0")
    print(f"[gpu-case] mr_main {gid} {typ} {' '.join(extra)}", flush=True)
    res = subprocess.run([str(MR_MAIN), str(gid), typ, *extra], env=env, capture_output=True, text=True, timeout=120)
    return res


def test_driver_runs_every_type(tmp_path):
    assert MR_MAIN.exists(), "build() links lib/mr_main"
    rp, ci = rmat_csr(12, 8, seed=9)
    n = len(rp) - 1
    # a symmetric pattern file of the lower triangle (as Graph500 R-MAT files are)
    rows, cols = [], []
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            if ci[q] <= i:
                rows.append(i)
                cols.append(int(ci[q]))
    write_mtx(tmp_path / "g.mtx", rows, cols, [1.0] * len(rows), n, symmetric=True, pattern=True)
    (tmp_path / "config.txt").write_text(f"7\t{tmp_path / 'g.mtx'}\t{n}\t{n}\t{len(rows)}\n")
    fields = {"origin": 10, "degree": 10, "dbg": 10, "rcm": 10, "rabbitorder": 10, "gorder1": 10, "gorder5": 10,
              "gorder6": 8, "order": 8}
    for typ, nf in fields.items():
        extra = ("256", "64") if typ == "order" else ()
        res = run_main(tmp_path, 7, typ, *extra)
        assert res.returncode == 0, (typ, res.stdout, res.stderr)
        assert "total error = 0" in res.stdout, (typ, res.stdout)
        last = (tmp_path / "log.txt").read_text().strip().splitlines()[-1].split("&")
        assert last[0] == "7" and len(last) == nf, (typ, last)
        if typ not in ("origin", "gorder6", "order"):
            assert last[1] == typ
        for f in last[1:]:
            if f not in ("", typ, "origin", "order", "256", "64", "gorder5"):
                assert float(f) >= 0.0, (typ, last)
    res = run_main(tmp_path, 99, "origin")
    assert res.returncode == 1 and "not found" in res.stderr


def test_driver_reads_gap_edge_list(tmp_path):
    """A config entry naming a GAP edge list (.el) goes through mr_gap_load."""
    assert MR_MAIN.exists(), "build() links lib/mr_main"
    rp, ci = rmat_csr(11, 8, seed=21)
    n = len(rp) - 1
    lines = [f"{i} {int(ci[q])}\n" for i in range(n) for q in range(rp[i], rp[i + 1])]
    (tmp_path / "g.el").write_text("".join(lines))
    (tmp_path / "config.txt").write_text(f"3\t{tmp_path / 'g.el'}\t{n}\t{n}\t{len(lines)}\n")
    for typ in ("origin", "gorder6"):
        res = run_main(tmp_path, 3, typ)
        assert res.returncode == 0, (typ, res.stdout, res.stderr)
        assert "total error = 0" in res.stdout, (typ, res.stdout)
