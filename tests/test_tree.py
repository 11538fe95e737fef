"""Adaptive octree and U-list (stkfmm_tree_*) against a brute-force restatement of the two rules:
split while a box holds more than leaf_cap points; U(B) = leaves whose closed boxes touch B's."""
import numpy as np
import pytest

import stkfmm_amd as S
from stkfmm_amd import workloads as W

BITS = 21


def quantize(x):
    q = np.floor(x * (1 << BITS))
    return np.clip(q, 0, (1 << BITS) - 1).astype(np.int64)


def check_tree(x, cap):
    tr = S.Tree(x, cap)
    n = len(x)
    assert np.array_equal(np.sort(tr.perm), np.arange(n))
    # leaves partition [0, n) in order
    assert tr.leaf_begin[0] == 0 and tr.leaf_end[-1] == n
    assert np.array_equal(tr.leaf_begin[1:], tr.leaf_end[:-1])
    q = quantize(x[tr.perm])
    L = tr.n_leaves
    lo = np.zeros((L, 3), dtype=np.int64)
    size = np.zeros(L, dtype=np.int64)
    for l in range(L):
        b, e, lev = tr.leaf_begin[l], tr.leaf_end[l], tr.leaf_level[l]
        assert e > b
        assert e - b <= cap or lev == BITS
        sh = BITS - lev
        anchors = q[b:e] >> sh
        assert np.all(anchors == anchors[0]), "leaf points outside one box"
        lo[l] = anchors[0] << sh
        size[l] = 1 << sh
        if lev > 0:  # the parent box held more than cap points (else it would be the leaf)
            psh = sh + 1
            inside = np.all((q >> psh) == (anchors[0] >> 1), axis=1)
            assert inside.sum() > cap
    # U-lists against all-pairs touching
    hi = lo + size[:, None]
    touch = np.all((lo[:, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[:, None, :]), axis=2)
    pairs = 0
    for l in range(L):
        pts = np.zeros(n, dtype=bool)
        for r in range(tr.ulist_ptr[l], tr.ulist_ptr[l + 1]):
            assert tr.rng_begin[r] < tr.rng_end[r]
            assert not pts[tr.rng_begin[r]:tr.rng_end[r]].any(), "overlapping ranges"
            pts[tr.rng_begin[r]:tr.rng_end[r]] = True
        expect = np.zeros(n, dtype=bool)
        for a in np.nonzero(touch[l])[0]:
            expect[tr.leaf_begin[a]:tr.leaf_end[a]] = True
        assert np.array_equal(pts, expect), l
        pairs += int(pts.sum()) * int(tr.leaf_end[l] - tr.leaf_begin[l])
    assert pairs == tr.pairs
    return tr


@pytest.mark.parametrize("cap", [1, 7, 50])
def test_uniform(cap):
    x = np.random.default_rng(cap).random((1500, 3))
    check_tree(x, cap)


def test_clustered():
    g = np.random.default_rng(11)
    x = W.wrap(np.concatenate([g.random((800, 3)), 0.3 + 0.01 * g.normal(size=(1200, 3))]))
    tr = check_tree(x, 40)
    assert tr.max_level >= 5


def test_duplicate_points_stop_at_finest_level():
    x = np.full((30, 3), 0.123)
    tr = check_tree(x, 4)
    assert tr.n_leaves == 1 and tr.leaf_level[0] == BITS


def test_single_leaf_and_empty():
    x = np.random.default_rng(3).random((10, 3))
    tr = check_tree(x, 100)
    assert tr.n_leaves == 1 and tr.pairs == 100
    t0 = S.Tree(np.zeros((0, 3)), 10)
    assert t0.n_leaves == 0 and t0.pairs == 0


def test_workload2_statistics():
    tr = S.Tree(W.pvel_points(200_000), 1000)
    sz = tr.leaf_end - tr.leaf_begin
    assert sz.max() <= 1000 and tr.pairs > 0
