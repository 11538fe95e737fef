"""Equivalent / check surfaces (host side): stkfmm_surface against an independent description of the
point set, the reference's point order, and leaf boxes of the tree.

The reference's stkfmm::surface (Lib/include/STKFMM/stkfmm_helpers.hpp:46-85) lists the
6(p-1)^2+2 boundary points of a p x p x p lattice on [-1,1]^3: the corner (-1,-1,-1), three faces
of (p-1)^2 points each, then the point reflection of those, mapped by (u + 1) alpha r + c with
r = 0.5^(depth+1).  Pinned here by: the set of unit points equals the boundary of the p^3 lattice
(an independent characterisation), the reflection symmetry and the first point, the per-face index
order of :53-71 on a small case written out by hand, and the affine map.
"""
import itertools

import numpy as np
import pytest

import stkfmm_amd as S


def unit_points(p):
    # alpha = 2, depth = 0 gives b = alpha * 0.5 = 1, and c = -1 maps (u + 1) * 1 - 1 back to u
    return S.surface(p, np.full(3, -1.0), 2.0, 0)


@pytest.mark.parametrize("p", [2, 3, 4, 6, 8, 10, 12, 16])
def test_point_set_is_the_lattice_boundary(p):
    u = unit_points(p)
    assert u.shape == (6 * (p - 1) ** 2 + 2, 3)
    grid = np.linspace(-1.0, 1.0, p)
    lattice = {tuple(np.round(q, 12)) for q in itertools.product(grid, grid, grid)
               if np.any(np.abs(np.abs(q) - 1.0) < 1e-15)}
    got = [tuple(np.round(q, 12)) for q in u]
    assert len(set(got)) == len(got)          # distinct
    assert set(got) == lattice                # exactly the boundary lattice points


@pytest.mark.parametrize("p", [2, 3, 7, 10])
def test_reflection_and_first_point(p):
    u = unit_points(p)
    n = len(u)
    assert np.array_equal(u[0], [-1.0, -1.0, -1.0])
    # (u + 1) - 1 rounds, so the mirrored halves agree to the last bit of 2, not exactly
    np.testing.assert_allclose(u[n // 2:], -u[: n // 2], rtol=0, atol=4.5e-16)


def test_index_order_small_case():
    # p = 3 written out from the loops of stkfmm_helpers.hpp:53-71 (unit coordinates)
    face_x = [(-1, 0, -1), (-1, 0, 0), (-1, 1, -1), (-1, 1, 0)]
    face_y = [(-1, -1, 0), (-1, -1, 1), (0, -1, 0), (0, -1, 1)]
    face_z = [(0, -1, -1), (0, 0, -1), (1, -1, -1), (1, 0, -1)]
    first = [(-1, -1, -1)] + face_x + face_y + face_z
    expect = np.array(first + [tuple(-v for v in q) for q in first], dtype=float)
    assert np.array_equal(unit_points(3), expect)


@pytest.mark.parametrize("p,alpha,depth", [(6, 1.05, 0), (10, 2.95, 3), (8, 2.95, 11), (4, 1.7, 20)])
def test_affine_map(p, alpha, depth):
    c = np.array([0.125, -0.5, 0.3])
    pts = S.surface(p, c, alpha, depth)
    u = unit_points(p)
    b = alpha * 0.5 * 0.5 ** depth
    np.testing.assert_allclose(pts, (u + 1.0) * b + c, rtol=0, atol=1e-15)
    assert np.all(pts >= c - 1e-15) and np.all(pts <= c + alpha * 0.5 ** depth + 1e-15)


def test_surface_argument_errors():
    with pytest.raises(S.ArgumentError):
        S.surface_count(1)
    with pytest.raises(S.ArgumentError):
        S.surface(1, [0, 0, 0], 1.0, 0)


def test_leaf_boxes_contain_their_points():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random((3000, 3)), 0.3 + 0.01 * rng.normal(size=(3000, 3))]) % 1.0
    t = S.Tree(x, 80)
    xs = x[t.perm]
    h = 0.5 ** t.leaf_level.astype(float)
    for l in range(t.n_leaves):
        pts = xs[t.leaf_begin[l]:t.leaf_end[l]]
        assert np.all(pts >= t.leaf_corner[l] - 1e-12)
        assert np.all(pts < t.leaf_corner[l] + h[l] + 1e-12)
    # corners sit on the level's grid, and leaf boxes do not overlap (empty boxes are not leaves,
    # so the volumes add to at most 1)
    q = t.leaf_corner / h[:, None]
    assert np.array_equal(q, np.round(q))
    assert np.sum(h ** 3) <= 1.0 + 1e-12


# k_s2m of each KERNEL (BuildKernel lists: LaplaceLayerKernel.hpp:315-355, StokesLayerKernel.hpp:38-88,
# RPYKernel.hpp:136-152, StokesRegSingleLayerKernel.hpp:211-232)
S2M = {(1, 0): 0, (1, 1): 1, (2, 0): 0, (2, 1): 1, (4, 0): 23, (8, 0): 6, (16, 0): 10, (32, 0): 9, (64, 0): 24,
       (128, 0): 7, (128, 1): 8, (256, 0): 7, (256, 1): 8, (512, 0): 7, (512, 1): 8, (1024, 0): 7, (1024, 1): 8}


@pytest.mark.parametrize("kl", sorted(S2M))
def test_s2m_sub_kernels(kl):
    assert int(S.get_s2m_uker(*kl)) == S2M[kl]


@pytest.mark.parametrize("k", [4, 8, 16, 32, 64])
def test_s2m_no_double_layer(k):
    with pytest.raises(S.PPKernelNotFound):
        S.get_s2m_uker(k, 1)
