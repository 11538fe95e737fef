"""GPU: equivalent / check surfaces of all leaves built on the device, and the per-leaf S2M-type
(leaf points -> leaf surface) and L2T-type (leaf surface -> leaf points) work lists, against the
host surface generator and the oracle's range-list driver."""
import numpy as np
import pytest
import torch

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.conftest import RADIUS_UKERS
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu


def tree_points(seed, n=6000):
    g = np.random.default_rng(seed)
    return W.wrap(np.concatenate([g.random((n // 2, 3)), 0.55 + 0.02 * g.normal(size=(n // 2, 3))]))


def values(uk, n, rng):
    v = rng.normal(size=(n, O.src_dim(uk)))
    if uk in RADIUS_UKERS:
        v[:, -1] = rng.uniform(1e-4, 5e-4, n)
    return v


@pytest.mark.parametrize("p,alpha", [(6, S.RAD_INNER), (10, S.RAD_OUTER), (2, 1.0)])
def test_device_surfaces_match_host_bitwise(p, alpha):
    say(f"surfaces p={p} alpha={alpha}")
    t = S.Tree(tree_points(1), 100)
    ls = S.LeafSurfaces(t, p, alpha)
    got = host(ls.coord).reshape(t.n_leaves, ls.count, 3)
    for l in range(t.n_leaves):
        r = 0.5 ** (t.leaf_level[l] + 1)
        c = t.leaf_corner[l] - r * (alpha - 1.0)
        assert np.array_equal(got[l], S.surface(p, c, alpha, int(t.leaf_level[l]))), l


def _ranges_oracle(uk, src, v, trg, gb, ge, gp, rb, re):
    return O.ranges(uk, src, v, trg, gb, ge, gp, rb, re)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [S.KERNEL.LapPGrad, S.KERNEL.RPY, S.KERNEL.PVel, S.KERNEL.StokesRegVelOmega,
                                    S.KERNEL.LapQPGradGrad])
def test_s2m_worklist_vs_oracle(kernel, prec):
    uk = S.get_s2m_uker(kernel)
    say(f"S2M kernel={int(kernel)} uker={int(uk)} prec={prec}")
    x = tree_points(10 + int(uk))
    t = S.Tree(x, 120)
    xs = x[t.perm]
    v = values(uk, t.n, np.random.default_rng(int(uk)))
    ls = S.LeafSurfaces(t, 8, S.RAD_OUTER)
    plan = ls.to_surface_plan(uk, prec)
    out = zeros(t.n_leaves * ls.count, O.trg_dim(uk))
    plan.execute(to_dev(xs), to_dev(v), ls.coord, out)
    sb = np.arange(t.n_leaves) * ls.count
    ref = _ranges_oracle(uk, xs, v, host(ls.coord), sb, sb + ls.count, np.arange(t.n_leaves + 1), t.leaf_begin,
                         t.leaf_end)
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [S.KERNEL.LapQPGradGrad, S.KERNEL.RPY, S.KERNEL.PVel, S.KERNEL.StokesRegVelOmega,
                                    S.KERNEL.PVelGrad])
def test_l2t_worklist_vs_oracle(kernel, prec):
    uk = S.get_uker(kernel, S.PPKERNEL.L2T)
    say(f"L2T kernel={int(kernel)} uker={int(uk)} prec={prec}")
    x = tree_points(20 + int(uk))
    t = S.Tree(x, 120)
    xs = x[t.perm]
    ls = S.LeafSurfaces(t, 6, S.RAD_OUTER)
    ns = t.n_leaves * ls.count
    dens = values(uk, ns, np.random.default_rng(100 + int(uk)))
    plan = ls.from_surface_plan(uk, prec)
    out = zeros(t.n, O.trg_dim(uk))
    d_x = to_dev(xs)
    plan.execute(ls.coord, to_dev(dens), d_x, out)
    sb = np.arange(t.n_leaves) * ls.count
    ref = _ranges_oracle(uk, host(ls.coord), dens, xs, t.leaf_begin, t.leaf_end, np.arange(t.n_leaves + 1), sb,
                         sb + ls.count)
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


def test_box_surfaces_argument_checks():
    say("surface args")
    corner = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(S.ArgumentError):
        S.box_surfaces_device(6, 1.05, corner, torch.zeros(4, dtype=torch.int64, device="cuda"))
    with pytest.raises(S.ArgumentError):
        S.box_surfaces_device(1, 1.05, corner, torch.zeros(4, dtype=torch.int32, device="cuda"))
    empty = S.box_surfaces_device(6, 1.05, corner[:0], torch.zeros(0, dtype=torch.int32, device="cuda"))
    assert empty.shape == (0, 3)
