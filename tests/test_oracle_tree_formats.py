"""Pins the oracle's ancestor-row rebuild (apps/ancestor.py:21-81): closed form
on complete trees, the parent-chain definition on random trees, orphans, and
the out-of-range error; then (GPU) the device rebuild, the parent-link LoD cut
and the checkpoint loader against it."""
import numpy as np
import pytest

from oracle import pyoracle as orc


def complete_tree(k, levels):
    """A complete k-ary tree in breadth-first order: parent(i) = (i - 1) // k."""
    n = sum(k ** d for d in range(levels))
    idx = np.arange(n)
    parent = np.where(idx == 0, -1, (idx - 1) // k)
    depth = np.zeros(n, dtype=np.int64)
    for i in range(1, n):
        depth[i] = depth[parent[i]] + 1
    return parent, depth


def random_forest(n_roots, n, seed, max_depth=9):
    """Well-formed random forest in an arbitrary node order: each node's
    parent sits one level up."""
    rng = np.random.default_rng(seed)
    depth = np.zeros(n, dtype=np.int64)
    parent = np.full(n, -1, dtype=np.int64)
    order = rng.permutation(n)
    roots, rest = order[:n_roots], order[n_roots:]
    placed = list(roots)
    for i in rest:
        p = placed[rng.integers(len(placed))]
        if depth[p] + 1 > max_depth:
            p = roots[rng.integers(n_roots)]
        parent[i] = p
        depth[i] = depth[p] + 1
        placed.append(i)
    return parent, depth


def chain_rows(parent, depth, L):
    """Definition: column d of row i is i's ancestor at depth d."""
    out = np.full((len(depth), L), -1, dtype=np.int32)
    for i in range(len(depth)):
        cur = i
        while parent[cur] >= 0:
            cur = parent[cur]
            out[i, depth[cur]] = cur
    return out


def test_complete_tree_closed_form():
    parent, depth = complete_tree(4, 5)
    rows = orc.rebuild_ancestor_path(parent, depth, 6)
    for i in (0, 1, 4, 5, 20, 21, 85, len(depth) - 1):
        expect, cur = [], i
        while cur > 0:
            cur = (cur - 1) // 4
            expect.append(cur)
        expect = expect[::-1]
        assert list(rows[i, :len(expect)]) == expect
        assert (rows[i, len(expect):] == -1).all()


def test_random_forest_matches_parent_chains():
    parent, depth = random_forest(7, 3000, seed=5)
    L = int(depth.max()) + 2
    assert np.array_equal(orc.rebuild_ancestor_path(parent, depth, L), chain_rows(parent, depth, L))


def test_orphans_and_width():
    parent, depth = complete_tree(2, 4)
    parent = parent.copy()
    parent[2] = -1                      # node 2 (depth 1) loses its parent
    rows = orc.rebuild_ancestor_path(parent, depth, 4)
    assert (rows[2] == -1).all()
    # its children keep only the orphan itself: [-1, 2, -1, -1]
    assert list(rows[5]) == [-1, 2, -1, -1] and list(rows[6]) == [-1, 2, -1, -1]
    with pytest.raises(IndexError):
        orc.rebuild_ancestor_path(parent, depth, 2)
