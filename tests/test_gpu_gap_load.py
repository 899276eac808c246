"""GAP-benchmark graph files (mr_gap_load) against a numpy restatement of the
reference builder's steps (parsb/include/builder.h:216-415): n = largest id +
1, edges sorted by (u, v, w), one per (u, v) keeping the smallest weight,
self-loops dropped, optional reverse edges with the same weight; the formats
of parsb/include/reader.h (.el .net .wel .gr .graph .sg .wsg)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def expected_csr(edges, symmetrize):
    """The builder's CSR of out-neighbours for (u, v, w) edges."""
    n = max(max(u, v) for u, v, _ in edges) + 1 if edges else 0
    el, seen = [], set()
    for u, v, w in sorted(edges):
        if (u, v) in seen or u == v:
            continue
        seen.add((u, v))
        el.append((u, v, w))
    if symmetrize:
        el += [(v, u, w) for u, v, w in el if (v, u) not in seen]
        el.sort()
    rp = np.zeros(n + 1, np.int64)
    for u, _, _ in el:
        rp[u + 1] += 1
    rp = np.cumsum(rp).astype(np.int32)
    ci = np.array([v for _, v, _ in el], np.int32)
    val = np.array([w for _, _, w in el], np.float64)
    return rp, ci, val


def random_edges(rng, n, m, weighted):
    out = []
    for _ in range(m):
        u, v = int(rng.integers(0, n)), int(rng.integers(0, n))
        w = float(rng.integers(1, 50)) / 4 if weighted else 1.0
        out.append((u, v, w))
    out += [(3, 3, 1.0), out[0], (out[1][0], out[1][1], out[1][2] + 1.0)]   # self-loop, duplicates
    return out


def load(ctx, mr, path, sym):
    mat = mr.Mat.load_gap(ctx, path, symmetrize=sym, store=mr.F64)
    rp, ci, val = mat.export()
    mat.close()
    return rp, ci, val


def check(got, want):
    assert np.array_equal(got[0], want[0]), "row pointers differ"
    assert np.array_equal(got[1], want[1]), "columns differ"
    assert np.array_equal(np.asarray(got[2], np.float64), want[2]), "values differ"


@pytest.mark.parametrize("sym", [False, True])
def test_edge_list(ctx, mr, tmp_path, sym):
    rng = np.random.default_rng(5)
    edges = random_edges(rng, 60, 300, weighted=False)
    for suffix in (".el", ".net"):
        p = tmp_path / f"g{suffix}"
        p.write_text("".join(f"{u} {v}\n" for u, v, _ in edges))
        check(load(ctx, mr, p, sym), expected_csr(edges, sym))


@pytest.mark.parametrize("sym", [False, True])
def test_weighted_edge_list(ctx, mr, tmp_path, sym):
    rng = np.random.default_rng(6)
    edges = random_edges(rng, 40, 250, weighted=True)
    p = tmp_path / "g.wel"
    p.write_text("".join(f"{u}\t{v}\t{w}\n" for u, v, w in edges))
    check(load(ctx, mr, p, sym), expected_csr(edges, sym))


def test_dimacs(ctx, mr, tmp_path):
    rng = np.random.default_rng(7)
    edges = random_edges(rng, 30, 120, weighted=True)
    lines = ["c a DIMACS shortest-path file\n", "p sp 30 123\n"]
    lines += [f"a {u + 1} {v + 1} {w}\n" for u, v, w in edges]
    p = tmp_path / "g.gr"
    p.write_text("".join(lines))
    check(load(ctx, mr, p, False), expected_csr(edges, False))


@pytest.mark.parametrize("weighted", [False, True])
def test_metis(ctx, mr, tmp_path, weighted):
    rng = np.random.default_rng(8)
    n = 25
    adj = {u: sorted({int(x) for x in rng.integers(0, n, int(rng.integers(0, 6)))}) for u in range(n)}
    adj[4] = []   # an isolated vertex: an empty line
    edges, lines = [], ["% METIS graph\n", f"{n} 0 {1 if weighted else 0}\n"]
    for u in range(n):
        toks = []
        for v in adj[u]:
            w = float(u + v + 1)
            edges.append((u, v, w if weighted else 1.0))
            toks.append(f"{v + 1} {int(w)}" if weighted else f"{v + 1}")
        lines.append(" ".join(toks) + "\n")
        if u == 10:
            lines.append("% a comment between adjacency lines\n")
    p = tmp_path / "g.graph"
    p.write_text("".join(lines))
    check(load(ctx, mr, p, False), expected_csr(edges, False))


@pytest.mark.parametrize("weighted", [False, True])
def test_serialized(ctx, mr, tmp_path, weighted):
    rng = np.random.default_rng(9)
    n = 40
    rp, ci, val = expected_csr(random_edges(rng, n, 200, weighted=True), False)
    n = len(rp) - 1
    if not weighted:
        val = np.ones(len(ci))
    head = np.array([0], np.uint8).tobytes() + np.array([len(ci), n], np.int64).tobytes()
    body = rp.astype(np.int64).tobytes()
    if weighted:
        pairs = np.stack([ci, val.astype(np.int32)], axis=1).astype(np.int32)
        body += pairs.tobytes()
    else:
        body += ci.astype(np.int32).tobytes()
    p = tmp_path / ("g.wsg" if weighted else "g.sg")
    p.write_bytes(head + body)
    want_val = np.floor(val) if weighted else val
    check(load(ctx, mr, p, False), (rp, ci, want_val))


def test_errors(ctx, mr, tmp_path):
    cases = [("g.txt", "0 1\n", mr.MRError, 5),
             ("g.el", "0 1\n2 x\n", mr.MRError, 7),
             ("g.wel", "0 1 2.5\n1 2\n", mr.MRError, 7),
             ("g.graph", "3 2 11\n2\n1\n\n", mr.MRError, 5),
             ("g.el", "0 -1\n", mr.MRError, 7)]
    for name, text, exc, code in cases:
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(exc) as e:
            mr.Mat.load_gap(ctx, p)
        assert e.value.code == code, (name, text, e.value)
    p = tmp_path / "t.sg"
    p.write_bytes(np.array([0], np.uint8).tobytes() + np.array([10, 5], np.int64).tobytes())   # no offsets
    with pytest.raises(mr.MRError) as e:
        mr.Mat.load_gap(ctx, p)
    assert e.value.code == 7
    with pytest.raises(mr.MRError) as e:
        mr.Mat.load_gap(ctx, tmp_path / "missing.el")
    assert e.value.code == 7
