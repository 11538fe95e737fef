"""GPU: W-list (M2T) and X-list (S2L) work lists over all node surfaces against the oracle's range-list
driver, FP64 and mixed."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.conftest import RADIUS_UKERS
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu


def tree(seed):
    g = np.random.default_rng(seed)
    x = W.wrap(np.concatenate([g.random((4000, 3)), 0.45 + 0.015 * g.normal(size=(4000, 3))]))
    return x, S.Tree(x, 60)


def values(uk, n, rng):
    v = rng.normal(size=(n, O.src_dim(uk)))
    if uk in RADIUS_UKERS:
        v[:, -1] = rng.uniform(1e-4, 5e-4, n)
    return v


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [S.KERNEL.PVel, S.KERNEL.RPY, S.KERNEL.LapQPGradGrad, S.KERNEL.StokesRegVelOmega,
                                    S.KERNEL.PVelGrad])
def test_w_list_plan_vs_oracle(kernel, prec):
    uk = S.get_m2t_uker(kernel)
    say(f"W-list kernel={int(kernel)} uker={int(uk)} prec={prec}")
    x, t = tree(30 + int(uk))
    xs = x[t.perm]
    ns = S.NodeSurfaces(t, 6)
    dens = values(uk, t.n_nodes * ns.count, np.random.default_rng(int(uk)))
    out = zeros(t.n, O.trg_dim(uk))
    plan = ns.w_plan(uk, prec)
    plan.execute(ns.coord, to_dev(dens), to_dev(xs), out)
    rb = t.w_node.astype(np.int64) * ns.count
    ref = O.ranges(uk, host(ns.coord), dens, xs, t.leaf_begin, t.leaf_end, t.w_ptr, rb, rb + ns.count)
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [S.KERNEL.PVel, S.KERNEL.RPY, S.KERNEL.LapQPGradGrad, S.KERNEL.StokesRegVelOmega])
def test_x_list_plan_vs_oracle(kernel, prec):
    uk = S.get_s2l_uker(kernel)
    say(f"X-list kernel={int(kernel)} uker={int(uk)} prec={prec}")
    x, t = tree(40 + int(uk))
    xs = x[t.perm]
    ns = S.NodeSurfaces(t, 6)
    v = values(uk, t.n, np.random.default_rng(int(uk) + 1))
    out = zeros(t.n_nodes * ns.count, O.trg_dim(uk))
    plan = ns.x_plan(uk, prec)
    plan.execute(to_dev(xs), to_dev(v), ns.coord, out)
    nodes = np.nonzero(np.diff(t.x_ptr) > 0)[0]
    gb = nodes * ns.count
    ptr = np.concatenate([[0], np.cumsum(np.diff(t.x_ptr)[nodes])])
    ref = O.ranges(uk, xs, v, host(ns.coord), gb, gb + ns.count, ptr, t.leaf_begin[t.x_leaf], t.leaf_end[t.x_leaf])
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)
