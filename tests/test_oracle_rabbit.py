"""Rabbit Order oracle (one deterministic order of rabbitorder/src/rabbit_order.cpp)
against hand-worked graphs and structural invariants."""
from __future__ import annotations

import numpy as np

from conftest import csr_rows, random_csr


def sym(n, pairs):
    rows = [[] for _ in range(n)]
    for a, b in pairs:
        rows[a].append(b)
    return csr_rows([sorted(set(r)) for r in rows], n)


def test_path_of_three(oracle):
    # degrees 1,2,1 -> visit 0, 2, 1; total strength 4.  0 joins 1 (gain
    # 1 - 1*2/4), 2 joins 1 (1 - 1*3/4); 1 stays on top with children 2 -> 0.
    # Depth-first numbering from the last child: 2, 0, then 1.
    rp, ci = sym(3, [(0, 1), (1, 2)])
    ids, ncom = oracle.rabbit(3, rp, ci)
    assert ids.tolist() == [1, 2, 0] and ncom == 1


def test_star(oracle):
    # leaves join the hub one by one (gains 1-4/8, 1-5/8, 1-6/8, 1-7/8 > 0);
    # the hub's child chain is 4 -> 3 -> 2 -> 1 by sibling links
    rp, ci = sym(5, [(0, k) for k in range(1, 5)])
    ids, ncom = oracle.rabbit(5, rp, ci)
    assert ids.tolist() == [4, 3, 2, 1, 0] and ncom == 1


def test_direction_and_self_loops_do_not_matter(oracle):
    rp1, ci1 = sym(6, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (2, 3)])
    rp2, ci2 = sym(6, [(1, 0), (2, 1), (0, 2), (4, 3), (5, 4), (3, 5), (3, 2), (0, 0), (4, 4)])
    assert np.array_equal(oracle.rabbit(6, rp1, ci1)[0], oracle.rabbit(6, rp2, ci2)[0])


def test_two_cliques_one_bridge(oracle):
    k = 6
    pairs = [(a, b) for a in range(k) for b in range(k) if a != b]
    pairs += [(a + k, b + k) for a in range(k) for b in range(k) if a != b]
    pairs += [(0, k)]
    rp, ci = sym(2 * k, pairs)
    ids, ncom = oracle.rabbit(2 * k, rp, ci)
    assert sorted(ids.tolist()) == list(range(2 * k))
    assert ncom == 2
    first = sorted(ids[:k].tolist())
    assert first == list(range(k)) or first == list(range(k, 2 * k))


def test_components_are_contiguous_and_permutation(oracle):
    rng = np.random.default_rng(5)
    # three disjoint random blocks: no merge can cross a block
    blocks = [(0, 300), (300, 700), (700, 1000)]
    pairs = []
    for lo, hi in blocks:
        for _ in range(4 * (hi - lo)):
            a, b = rng.integers(lo, hi, 2)
            pairs.append((int(a), int(b)))
    rp, ci = sym(1000, pairs)
    ids, ncom = oracle.rabbit(1000, rp, ci)
    assert np.array_equal(np.sort(ids), np.arange(1000))
    # every top-level community is one id range; a community never spans blocks,
    # so the ids of each block are a union of whole ranges -- check per block
    # that no id range of another block interleaves inside a community range
    blk = np.zeros(1000, int)
    for i, (lo, hi) in enumerate(blocks):
        blk[lo:hi] = i
    by_id = blk[np.argsort(ids)]
    changes = int(np.count_nonzero(np.diff(by_id)))
    assert changes <= ncom - 1
    rp2, ci2 = random_csr(np.random.default_rng(6), 400, empty_frac=0.2, max_row=6)
    ids2, _ = oracle.rabbit(400, rp2, ci2)
    assert np.array_equal(np.sort(ids2), np.arange(400))
