"""GPU: mixed precision on near-duplicate points.  Pairs 1e-9 to 1e-15 apart (relative to the unit
box, or to the leaf in a U-list) have 1/r^3 terms beyond the FP32 range; the mixed mode must still
return finite results within its bar of the FP64 oracle, target by target."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from tests.conftest import source_values
from tests.gpu_util import MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu

SEPARATIONS = (1e-9, 1e-11, 1e-13, 1e-14, 1e-15)


def near_points(rng, n_src, n_trg):
    """Sources uniform in [0.25, 0.75)^3; the first len(SEPARATIONS) * 3 targets sit at those
    distances from sources (random directions), the rest are uniform."""
    src = 0.25 + 0.5 * rng.random((n_src, 3))
    trg = rng.random((n_trg, 3))
    k = 0
    for d in SEPARATIONS:
        for _ in range(3):
            u = rng.normal(size=3)
            trg[k] = src[7 * k + 1] + d * u / np.linalg.norm(u)
            k += 1
    return src, trg, k


def per_target_rel(got, ref):
    scale = np.maximum(np.abs(ref).max(axis=1), 1e-300)
    return (np.abs(got - ref).max(axis=1) / scale).max()


@pytest.mark.parametrize("uk", [13, 7, 8, 6, 9, 11, 10, 12, 2, 3, 4, 5, 14, 16, 18, 22])
def test_mixed_direct_sum_near_duplicates(uk):
    say(f"near duplicates uker={uk}")
    rng = np.random.default_rng(900 + uk)
    src, trg, n_near = near_points(rng, 600, 300)
    v = source_values(uk, len(src), rng)
    if uk in (14, 16):
        # the reference's algebra for these two adds and subtracts TrD in sequence
        # (StokesSingleLayerKernel.hpp:76-98, 139-145), which costs the FP64 oracle |TrD| / |f.r|
        # relative accuracy on a close pair (~1e-5 at 1e-11); with TrD = 0 the oracle stays exact
        v[:, 3] = 0.0
    ref = O.direct(uk, src, v, trg)
    out = zeros(len(trg), O.trg_dim(uk))
    S.uker_device(uk, to_dev(src), to_dev(v), to_dev(trg), out, precision=S.PREC.MIXED)
    got = host(out)
    assert np.all(np.isfinite(got)), np.argwhere(~np.isfinite(got))[:5]
    assert np.all(np.isfinite(ref))
    assert per_target_rel(got[:n_near], ref[:n_near]) < MIXED_TOL
    l2, mx = O.max_error(got[n_near:], ref[n_near:])
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)


def test_mixed_ulist_near_duplicates():
    """The same inside a U-list (level-scaled coordinates), fused PVel, with the far-tile kernel."""
    say("near duplicates in a U-list")
    rng = np.random.default_rng(17)
    x = rng.random((20000, 3))
    k = 0
    for d in SEPARATIONS:
        for _ in range(4):
            u = rng.normal(size=3)
            x[100 + k] = x[5000 + 37 * k] + d * u / np.linalg.norm(u)
            k += 1
    tr = S.Tree(x, 300)
    xs = x[tr.perm]
    sl, dl = rng.normal(size=(len(x), 4)), rng.normal(size=(len(x), 9))
    plan = S.Plan.from_tree(tr, 13, S.PREC.MIXED)
    out = zeros(len(x), 4)
    d_x = to_dev(xs)
    plan.execute(d_x, to_dev(sl), d_x, out, src_value2=to_dev(dl))
    got = host(out)
    assert np.all(np.isfinite(got))
    inv = np.empty_like(tr.perm)
    inv[tr.perm] = np.arange(len(x))
    near = np.unique(np.concatenate([inv[100:100 + k], inv[[5000 + 37 * i for i in range(k)]]]))
    leaves = np.unique(np.searchsorted(tr.leaf_begin, near, side="right") - 1)
    ref = O.ranges(13, xs, np.concatenate([sl, dl], axis=1), xs, tr.leaf_begin, tr.leaf_end, tr.ulist_ptr,
                   tr.rng_begin, tr.rng_end, sel=leaves)
    assert per_target_rel(got[near], ref[near]) < MIXED_TOL
    rows = np.concatenate([np.arange(tr.leaf_begin[l], tr.leaf_end[l]) for l in leaves])
    rows = np.setdiff1d(rows, near)
    l2, mx = O.max_error(got[rows], ref[rows])
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)
