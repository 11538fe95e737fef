"""W- and X-lists of the adaptive octree (the M2T / S2L lists of PVFMM's near field; PVFMM is not
vendored, so the kernel-independent FMM's definitions are used) against brute force on small trees:
W(leaf B) = nodes C deeper than B with C not touching B and parent(C) touching B; X = its transpose.
Also checked: the W-list nodes of B cover every leaf that descends from a colleague of B (a same-level
box touching B) and does not touch B exactly once."""
import numpy as np
import pytest

import stkfmm_amd as S

CELL = 2.0 ** -21


def boxes(t):
    lo = np.round(t.node_corner / CELL).astype(np.int64)
    size = (2 ** (21 - t.node_level.astype(np.int64)))
    return lo, size


def touches(lo, size, a, b):
    return bool(np.all(lo[a] <= lo[b] + size[b]) and np.all(lo[b] <= lo[a] + size[a]))


def trees():
    g = np.random.default_rng(11)
    yield S.Tree(g.random((2500, 3)), 40)
    x = np.concatenate([g.random((1500, 3)), 0.7 + 0.02 * g.normal(size=(1500, 3)),
                        0.2 + 0.002 * g.normal(size=(500, 3))]) % 1.0
    yield S.Tree(x, 30)


@pytest.mark.parametrize("t", list(trees()), ids=["uniform", "clustered"])
def test_w_and_x_lists_by_definition(t):
    lo, size = boxes(t)
    node_of_leaf = np.full(t.n_leaves, -1)
    for c in range(t.n_nodes):
        if t.node_leaf[c] >= 0:
            node_of_leaf[t.node_leaf[c]] = c
    for l in range(t.n_leaves):
        b = node_of_leaf[l]
        expect = set()
        for c in range(1, t.n_nodes):
            if t.node_level[c] > t.node_level[b] and not touches(lo, size, c, b) \
                    and touches(lo, size, t.node_parent[c], b):
                expect.add(c)
        got = t.w_node[t.w_ptr[l]:t.w_ptr[l + 1]]
        assert len(got) == len(set(got))
        assert set(got.tolist()) == expect, l
    # X is the transpose of W
    pairs_w = {(l, int(c)) for l in range(t.n_leaves) for c in t.w_node[t.w_ptr[l]:t.w_ptr[l + 1]]}
    pairs_x = {(int(l), c) for c in range(t.n_nodes) for l in t.x_leaf[t.x_ptr[c]:t.x_ptr[c + 1]]}
    assert pairs_w == pairs_x


@pytest.mark.parametrize("t", list(trees()), ids=["uniform", "clustered"])
def test_w_list_covers_separated_descendants_of_colleagues_once(t):
    lo, size = boxes(t)
    node_of_leaf = {int(t.node_leaf[c]): c for c in range(t.n_nodes) if t.node_leaf[c] >= 0}

    def ancestors(c):
        out = [c]
        while t.node_parent[c] >= 0:
            c = t.node_parent[c]
            out.append(c)
        return out

    for l in range(0, t.n_leaves, 3):
        b = node_of_leaf[l]
        wset = set(t.w_node[t.w_ptr[l]:t.w_ptr[l + 1]].tolist())
        for a_leaf, a in node_of_leaf.items():
            if t.node_level[a] <= t.node_level[b] or touches(lo, size, a, b):
                continue
            anc = ancestors(a)
            # the ancestor of A at B's level
            q = [c for c in anc if t.node_level[c] == t.node_level[b]]
            if not q or not touches(lo, size, q[0], b):
                continue   # not below a colleague of B
            assert sum(c in wset for c in anc) == 1, (l, a_leaf)


def test_node_arrays_are_consistent():
    t = next(trees())
    assert t.node_parent[0] == -1 and t.node_level[0] == 0
    for c in range(1, t.n_nodes):
        p = t.node_parent[c]
        assert t.node_level[c] == t.node_level[p] + 1
        assert t.node_begin[p] <= t.node_begin[c] <= t.node_end[c] <= t.node_end[p]
    leaves = t.node_leaf >= 0
    assert leaves.sum() == t.n_leaves
    np.testing.assert_array_equal(t.node_begin[leaves][np.argsort(t.node_leaf[leaves])], t.leaf_begin)


# k_s2l and k_m2t of each KERNEL (BuildKernel lists)
S2L = {(1, 0): 0, (1, 1): 1, (2, 0): 0, (2, 1): 1, (4, 0): 23, (8, 0): 6, (16, 0): 10, (32, 0): 9, (64, 0): 24,
       (128, 0): 7, (128, 1): 8, (256, 0): 7, (256, 1): 8, (512, 0): 7, (512, 1): 8, (1024, 0): 7, (1024, 1): 8}
M2T = {1: 2, 2: 4, 4: 4, 8: 6, 16: 12, 32: 6, 64: 21, 128: 7, 256: 14, 512: 16, 1024: 18}


def test_sub_kernels_of_the_lists():
    for kl, uk in S2L.items():
        assert int(S.get_s2l_uker(*kl)) == uk
    for k, uk in M2T.items():
        assert int(S.get_m2t_uker(k)) == uk
    with pytest.raises(S.PPKernelNotFound):
        S.get_s2l_uker(16, 1)
