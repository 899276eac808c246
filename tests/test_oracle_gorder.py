"""The Gorder-family oracle pinned by hand-derived schedules and by dense
matrix algebra (numpy products, not the oracle's loops): every batch of the
pGOrder GPU loop must be the first |P| unplaced vertices by (2-hop window
count desc, id asc), and every Gorder pop must carry the maximal score."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import random_csr


def dense(n, rp, ci):
    A = np.zeros((n, n), np.int64)
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            A[i, ci[q]] += 1
    return A


def test_schedule_hand_derived(oracle):
    # W = 1024: 32-wide doubling batches, then W-wide batches; O covers the
    # placed prefix until W, then equals P
    ps, pl, os_, ol = oracle.pgorder_schedule(5000, 1024)
    assert len(ps) == 5 + 5   # ceil(5000/1024) + log2(1024/32)
    assert list(zip(ps[:7], pl[:7])) == [(0, 32), (32, 32), (64, 64), (128, 128), (256, 256), (512, 512), (1024, 1024)]
    assert list(zip(os_[:7], ol[:7])) == [(0, 32), (0, 64), (0, 128), (0, 256), (0, 512), (0, 1024), (1024, 1024)]
    assert (ps[9], pl[9], os_[9], ol[9]) == (4096, 904, 4096, 904)
    # the union of the P ranges is exactly [0, n)
    cover = np.zeros(5000, int)
    for s, l in zip(ps, pl):
        cover[s:s + max(l, 0)] += 1
    assert (cover == 1).all()
    ps, pl, os_, ol = oracle.pgorder_schedule(100, 64)
    assert len(ps) == 2 + 1
    assert list(zip(ps, pl)) == [(0, 32), (32, 32), (64, 36)]
    assert list(zip(os_, ol)) == [(0, 32), (0, 64), (64, 36)]


def brute_gpu_loop(n, A, W, cap=10000):
    ps, pl, os_, ol = None, None, None, None
    import oracle
    ps, pl, os_, ol = oracle.pgorder_schedule(n, W)
    rowlen = A.sum(1)
    collen = A.sum(0)
    placed = np.zeros(n, bool)
    placement = np.full(n, -1, np.int64)
    score = rowlen.copy()
    for b in range(len(ps)):
        cand = np.flatnonzero(~placed)
        order = cand[np.lexsort((cand, -score[cand]))]
        take = order[: max(min(pl[b], n), 0)]
        placement[ps[b]: ps[b] + len(take)] = take
        placed[take] = True
        win = placement[os_[b]: os_[b] + max(min(ol[b], n - os_[b]), 0)]
        w = np.zeros(n, np.int64)
        w[win[rowlen[win] < cap]] = 1
        c = (collen < cap).astype(np.int64)
        score = ((w @ A) * c) @ A.T
        score[placed] = 0
    return placement


@pytest.mark.parametrize("seed,n,W", [(0, 40, 32), (1, 150, 64), (2, 300, 64), (3, 257, 128), (4, 20, 32)])
def test_gpu_loop_equals_dense_brute_force(oracle, seed, n, W):
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=6, empty_frac=0.3)
    got = oracle.pgorder_gpu_loop(n, rp, ci, W)
    want = brute_gpu_loop(n, dense(n, rp, ci), W)
    assert np.array_equal(got.astype(np.int64), want)


def test_gpu_loop_ties_and_isolated(oracle):
    # no edges at all: every count is 0, so placement is the identity
    n = 100
    rp = np.zeros(n + 1, np.int32)
    got = oracle.pgorder_gpu_loop(n, rp, np.zeros(0, np.int32), 32)
    assert got.tolist() == list(range(n))


def gorder1_scores(A, order, pid, window, L):
    n = A.shape[0]
    collen = A.sum(0)
    w = np.zeros(n, np.int64)
    w[order[max(0, pid - window):pid]] = 1
    cl = (collen < L).astype(np.int64)
    return (w * cl) @ A.T + w @ A + ((w @ A) * cl) @ A.T


@pytest.mark.parametrize("seed,n", [(5, 60), (6, 200), (7, 121)])
def test_gorder1_pops_carry_max_score(oracle, seed, n):
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=8, empty_frac=0.2)
    A = dense(n, rp, ci)
    order, perm = oracle.gorder1(n, rp, ci, 5)
    assert oracle.check_perm(perm) and np.array_equal(perm[order], np.arange(n))
    L = int(math.sqrt(n) / 4)
    rowlen = A.sum(1)
    assert order[0] == int(np.argmax(rowlen))          # start: first maximal row length
    placed = np.zeros(n, bool)
    for pid in range(n):
        s = gorder1_scores(A, order, pid, 5, L)
        free = ~placed
        assert s[order[pid]] == s[free].max(), f"pop {pid}"
        placed[order[pid]] = True


def test_gorder1_empty_graph_hand_worked(oracle):
    # all scores stay 0: the start (vertex 0) is popped first, then bucket 0
    # is read from its head: pushes insert right after the head -> n-1, ..., 1
    n = 7
    order, _ = oracle.gorder1(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), 5)
    assert order.tolist() == [0, 6, 5, 4, 3, 2, 1]


@pytest.mark.parametrize("seed,n", [(8, 120), (9, 300)])
def test_order_on_cpu_pops_carry_max_score(oracle, seed, n):
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=10, empty_frac=0.1)
    A = dense(n, rp, ci)
    rowlen, collen = A.sum(1), A.sum(0)
    sl = rng.permutation(n)[: n // 2].astype(np.uint32)
    degree = max(3, int(math.sqrt(n) / 4) + 2)
    out = oracle.order_on_cpu(n, rp, ci, sl, 3, 5, degree)
    assert sorted(out.tolist()) == sorted(sl.tolist())
    # start: first slice vertex with strictly maximal out-list length
    first = int(sl[0])
    for v in sl:
        if collen[v] > collen[first]:
            first = int(v)
    assert out[0] == first
    inb = np.zeros(n, bool)
    inb[sl] = True
    c = (collen < degree).astype(np.int64)
    placed = np.zeros(n, bool)
    for pid in range(len(out)):
        w = np.zeros(n, np.int64)
        win = out[max(0, pid - 5):pid].astype(np.int64)
        w[win[rowlen[win] < degree]] = 1
        s = ((w @ A) * c) @ A.T
        free = inb & ~placed
        if pid > 0:
            assert s[out[pid]] == s[free].max(), f"pop {pid}"
        placed[out[pid]] = True


def test_pgorder_cpu_phase_leaves_prefix(oracle):
    rng = np.random.default_rng(10)
    n = 2600
    rp, ci = random_csr(rng, n, max_row=5, empty_frac=0.3)
    g, f, perm = oracle.pgorder(n, rp, ci, 64, 5)
    ps, pl, os_, ol = oracle.pgorder_schedule(n, 64)
    first = int(math.log2(64))
    # batches before log2(W) keep the GPU order; later ones are permutations of it
    assert np.array_equal(f[: os_[first]], g[: os_[first]])
    for b in range(first, len(ps)):
        s, l = os_[b], ol[b]
        assert sorted(f[s:s + l].tolist()) == sorted(g[s:s + l].tolist())
    assert oracle.check_perm(perm) and np.array_equal(perm[f], np.arange(n))


def brute_gorder5_gpu(n, A, vsize, width1):
    """gorder5's GPU loop in dense algebra: start scores are column counts,
    batches of vsize, window = the last width1 batches, no hub caps."""
    placed = np.zeros(n, bool)
    placement = np.full(n, -1, np.int64)
    score = A.sum(0).astype(np.int64)
    B = (n + vsize - 1) // vsize
    for b in range(B):
        cand = np.flatnonzero(~placed)
        order = cand[np.lexsort((cand, -score[cand]))]
        start = b * vsize
        take = order[: min(vsize, n - start)]
        placement[start: start + len(take)] = take
        placed[take] = True
        os_ = 0 if b < width1 else (b - width1 + 1) * vsize
        ol = min((b + 1) * vsize if b < width1 else width1 * vsize, n - os_)
        w = np.zeros(n, np.int64)
        w[placement[os_: os_ + ol]] = 1
        score = (w @ A) @ A.T
        score[placed] = 0
    return placement


@pytest.mark.parametrize("seed,n,vsize,width1", [(11, 40, 8, 1), (12, 150, 32, 1), (13, 300, 16, 3), (14, 97, 10, 2),
                                                 (15, 5, 8, 1)])
def test_gorder5_gpu_loop_equals_dense_brute_force(oracle, seed, n, vsize, width1):
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=6, empty_frac=0.3)
    g, f, perm = oracle.gorder5(n, rp, ci, vsize, width1, 5)
    assert np.array_equal(g.astype(np.int64), brute_gorder5_gpu(n, dense(n, rp, ci), vsize, width1))
    # the CPU phase permutes inside every batch and nothing across batches
    for b in range(0, n, vsize):
        assert sorted(f[b:b + vsize].tolist()) == sorted(g[b:b + vsize].tolist())
    assert oracle.check_perm(perm) and np.array_equal(perm[f.astype(np.int64)], np.arange(n))


def test_gorder5_start_scores_are_column_counts(oracle):
    # rows 0..3 of M: [3], [3], [3], [0, 1, 2]: row 3 is the longest row but
    # column 3 has the most entries, so the first placed vertex is 3 only
    # because of the column count; then 0, 1, 2 by id (one entry each)
    rp = np.array([0, 1, 2, 3, 6], np.int32)
    ci = np.array([3, 3, 3, 0, 1, 2], np.int32)
    g, _, _ = oracle.gorder5(4, rp, ci, 1, 1, 5)
    assert g[0] == 3
    rp = np.array([0, 3, 4, 5, 6], np.int32)     # row 0 longest, column counts 1,2,2,1
    ci = np.array([1, 2, 3, 1, 2, 0], np.int32)
    g, _, _ = oracle.gorder5(4, rp, ci, 1, 1, 5)
    assert g[0] == 1


def tile_scores(A, order, pid, width1, width2):
    """Score of every vertex before pop `pid` of the tile-aware order: for
    each x entry of the current tile whose column has fewer than width2
    entries, one for every row holding that column.  The tile is replayed from
    the row lengths of the rows placed so far."""
    n = A.shape[0]
    rowlen, collen = A.sum(1), A.sum(0)
    start = curr = 0
    tile = []
    for q in range(pid):
        v = int(order[q])
        cols = np.flatnonzero(A[v])
        deg = len(cols)
        if curr - start + deg < width1:
            tile += cols.tolist()
            curr += deg
        else:
            start = (curr + deg) // width1 * width1
            tile = [int(c) for i, c in enumerate(cols) if curr + i >= start]
            curr += deg
    w = np.zeros(n, np.int64)
    for c in tile:
        if collen[c] < width2:
            w[c] += 1
    return A @ w


@pytest.mark.parametrize("seed,n,w1,w2", [(21, 60, 16, 100), (22, 200, 64, 6), (23, 150, 7, 1000), (24, 90, 1, 5)])
def test_tile_order_pops_carry_max_score(oracle, seed, n, w1, w2):
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=8, empty_frac=0.2)
    A = dense(n, rp, ci)
    order, perm = oracle.tile_order(n, rp, ci, w1, w2)
    assert oracle.check_perm(perm) and np.array_equal(perm[order.astype(np.int64)], np.arange(n))
    assert order[0] == int(np.argmax(A.sum(1)))        # first longest row
    placed = np.zeros(n, bool)
    placed[order[0]] = True
    for pid in range(1, n):
        s = tile_scores(A, order, pid, w1, w2)
        assert s[order[pid]] == s[~placed].max(), f"pop {pid}"
        placed[order[pid]] = True


@pytest.mark.parametrize("cap", [2, 3, 4])
@pytest.mark.parametrize("seed,n,W", [(11, 120, 32), (12, 300, 64), (13, 257, 64)])
def test_gpu_loop_hub_cap_equals_dense_brute_force(oracle, cap, seed, n, W):
    """The hub cap of the window score (10,000 in the reference) as a
    parameter: with caps of 2-4 the small graphs have hubs on both sides of
    it, and the loop must equal the dense brute force with the same cap."""
    rng = np.random.default_rng(seed)
    rp, ci = random_csr(rng, n, max_row=7, empty_frac=0.2)
    got = oracle.pgorder_gpu_loop(n, rp, ci, W, hub_cap=cap)
    want = brute_gpu_loop(n, dense(n, rp, ci), W, cap=cap)
    assert np.array_equal(got.astype(np.int64), want)


def _csr(n, edges):
    rows = [[] for _ in range(n)]
    for r, c in edges:
        rows[r].append(c)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in sorted(r)], np.int32)
    return rp, ci


def test_gorder1_hand_worked_update_order(oracle):
    """gorder1 (class.cu:1881-1971) on n = 64 (L = floor(8/4) = 2), window 5,
    multiplied matrix M with entries (5,9), (5,10), (7,5); every other vertex
    isolated.  Worked by hand from the reference:
      push 0..63 into bucket 0: the list reads 0, 63, 62, ..., 1 (each push
      goes right after the head); start = first argmax row length = 5 -> bucket 1.
      pop 5.  Rows-of-5 update first (collen(5) = 1 < L): 7 -> bucket 1 [7];
      then cols(5) = {9, 10} ascending: 9 -> [7, 9], 10 -> [7, 10, 9]
      (after the head).  pop 7, then 10, then 9; bucket 1 is empty, so
      bucket 0 follows from its head: 0, 63, 62, ..., 1 without 5, 7, 9, 10.
    Running the column loop before the row loop would pop 9 second."""
    rp, ci = _csr(64, [(5, 9), (5, 10), (7, 5)])
    order, perm = oracle.gorder1(64, rp, ci, 5)
    want = [5, 7, 10, 9, 0] + [v for v in range(63, 0, -1) if v not in (5, 7, 9, 10)]
    assert order.tolist() == want


def test_gorder1_hand_worked_hub_above_cap(oracle):
    """Symmetric star on n = 64 (L = 2): centre 10 with leaves 20, 21, 22 in
    both directions.  collen(10) = 3 >= L, so popping a leaf reaches no
    sibling through the centre.  By hand: start 10 (row length 3); pop 10,
    its columns 20, 21, 22 enter bucket 1 as [20, 22, 21]; the leaves pop in
    that order and add nothing; bucket 0 follows.  Ignoring the cap would give
    popping 20 a +1 for 21 and 22 (both in rows(10)) and pop 21 before 22."""
    edges = [(10, 20), (10, 21), (10, 22), (20, 10), (21, 10), (22, 10)]
    rp, ci = _csr(64, edges)
    order, _ = oracle.gorder1(64, rp, ci, 5)
    want = [10, 20, 22, 21, 0] + [v for v in range(63, 0, -1) if v not in (10, 20, 21, 22)]
    assert order.tolist() == want


def test_order_on_cpu_hand_worked(oracle):
    """order_on_cpu (class.cu:2220-2314) on the slice [3, 1, 4, 0, 2, 5] of
    an 8-vertex M with rows 0:[4] 1:[4,6] 2:[4] 3:[6] 4:[6] 5:[6], window 5,
    degree cap 100.  By hand:
      the slice is pushed in order, so bucket 0 reads 3, 5, 2, 0, 4, 1;
      start = first vertex with the largest column count = 4 (rows 0, 1, 2).
      pop 4: y = 6, z in rows(6) = 1, 3, 5 (4 placed) -> bucket 1 [1, 5, 3].
      pop 1: y = 4 gives 0, 2 -> bucket 1 [5, 2, 0, 3]; y = 6 moves 3 and 5 up
      -> bucket 2 [3, 5], bucket 1 [2, 0].
      pop 3: y = 6 moves 5 up -> bucket 3 [5].   pop 5: nothing left.
      pop 2 (bucket 1 head): y = 4 moves 0 up.    pop 0.
    Traversing rows(6) in descending order would pop 5 second."""
    rp, ci = _csr(8, [(0, 4), (1, 4), (1, 6), (2, 4), (3, 6), (4, 6), (5, 6)])
    got = oracle.order_on_cpu(8, rp, ci, np.array([3, 1, 4, 0, 2, 5]), 7, 5, 100)
    assert got.tolist() == [4, 1, 3, 5, 2, 0]
