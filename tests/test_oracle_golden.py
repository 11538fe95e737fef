"""The oracle (oracle/p2p_oracle.c) against the reference's own scalar oracle Test/SimpleKernel.cpp.

Frozen fixtures (tests/golden/simplekernel_fixtures.npz, written by tests/golden/make_golden.py from
the reference build) are always checked; the live reference build in oracle/_ref is checked too
when it is present.  Micro-kernels SimpleKernel has no entry for are pinned by exact identities
with pinned ones.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import UKER_REF, source_values

TOL = 1e-13


@pytest.mark.parametrize("uk", sorted(UKER_REF))
def test_direct_matches_golden(golden, uk):
    K, layer = UKER_REF[uk]
    tag = f"K{K}_L{layer}"
    out = O.direct(uk, golden[tag + "_src"], golden[tag + "_sv"], golden[tag + "_trg"])
    expect = golden[tag + "_out"]
    kt = O.trg_dim(uk)
    assert out.shape == expect.shape == (expect.shape[0], kt)
    l2, mx = O.max_error(out, expect)
    assert l2 < TOL and mx < TOL, (uk, l2, mx)


@pytest.mark.parametrize("uk", sorted(UKER_REF))
def test_pairs_match_golden(golden, uk):
    K, layer = UKER_REF[uk]
    tag = f"K{K}_L{layer}"
    s, t, f, expect = (golden[tag + k] for k in ("_pair_s", "_pair_t", "_pair_f", "_pair_out"))
    for i in range(len(s)):
        u = O.pair(uk, t[i] - s[i], f[i]) * O.scale(uk)
        np.testing.assert_allclose(u, expect[i], rtol=1e-13, atol=1e-13 * np.abs(expect[i]).max())


def test_coincident_pair_is_zero():
    for uk in range(14):
        f = np.ones(O.src_dim(uk))
        if uk in (9,):  # regularised: the mask is on |d|^2 + eps^2
            f[3] = 0.0
        assert np.all(O.pair(uk, np.zeros(3), f) == 0.0), uk


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("uk", sorted(UKER_REF))
def test_direct_matches_live_reference(uk):
    K, layer = UKER_REF[uk]
    rng = np.random.default_rng(100 + uk)
    ks, kt = O.src_dim(uk), O.trg_dim(uk)
    src = rng.random((97, 3))
    trg = np.concatenate([src[:5], rng.random((60, 3))])
    sv = source_values(uk, 97, rng)
    l2, mx = O.max_error(O.direct(uk, src, sv, trg), O.simplekernel_direct(K, layer, ks, kt, src, sv, trg))
    assert l2 < TOL and mx < TOL


def _rand(rng, n=200):
    src = rng.random((n, 3))
    trg = np.concatenate([src[:7], rng.random((n // 2, 3))])
    return src, trg


def test_identity_laplace_p_is_pgrad_potential():
    rng = np.random.default_rng(1)
    src, trg = _rand(rng)
    q = rng.uniform(-1, 1, (len(src), 1))
    np.testing.assert_allclose(O.direct(0, src, q, trg)[:, 0], O.direct(2, src, q, trg)[:, 0], rtol=1e-13, atol=1e-13)


def test_identity_laplace_dipolep_is_dipolepgrad_potential():
    rng = np.random.default_rng(2)
    src, trg = _rand(rng)
    d = rng.uniform(-1, 1, (len(src), 3))
    np.testing.assert_allclose(O.direct(1, src, d, trg)[:, 0], O.direct(3, src, d, trg)[:, 0], rtol=1e-13, atol=1e-13)


def test_identity_rpy_u_is_rpy_ulapu_velocity():
    rng = np.random.default_rng(3)
    src, trg = _rand(rng)
    v = source_values(11, len(src), rng)
    np.testing.assert_allclose(O.direct(10, src, v, trg), O.direct(11, src, v, trg)[:, :3], rtol=1e-12, atol=1e-13)


def test_identity_stk_ulapu_is_rpy_ulapu_at_zero_radius():
    # stk_ulapu(f) = rpy_ulapu(f, b = 0) (RPYKernel.hpp:58-89 vs :103-128)
    rng = np.random.default_rng(4)
    src, trg = _rand(rng)
    v = source_values(11, len(src), rng)
    v[:, 3] = 0.0
    assert np.array_equal(O.direct(12, src, v[:, :3], trg), O.direct(11, src, v, trg))


def test_identity_stokes_vel_is_stk_ulapu_velocity():
    rng = np.random.default_rng(5)
    src, trg = _rand(rng)
    f = rng.uniform(-1, 1, (len(src), 3))
    np.testing.assert_allclose(O.direct(6, src, f, trg), O.direct(12, src, f, trg)[:, :3], rtol=1e-13, atol=1e-13)


def test_identity_fused_is_sl_plus_dl():
    rng = np.random.default_rng(6)
    src, trg = _rand(rng)
    sl = rng.uniform(-1, 1, (len(src), 4))
    dl = rng.uniform(-1, 1, (len(src), 9))
    fused = O.direct(13, src, np.concatenate([sl, dl], axis=1), trg)
    split = O.direct(7, src, sl, trg) + O.direct(8, src, dl, trg)
    np.testing.assert_allclose(fused, split, rtol=1e-13, atol=1e-13)


def test_error_metrics_match_reference_definition():
    # ComponentError (Test/Test.cpp:113-151) on a hand-checked case
    a = np.array([[1.0], [2.0], [3.0]])
    b = np.array([[1.0], [2.5], [2.0]])
    e = O.error_metrics(a, b)[0]
    assert np.isclose(e["errorL2"], np.sqrt(0.25 + 1.0) / np.sqrt(1 + 6.25 + 4))
    assert np.isclose(e["errorRMS"], np.sqrt(1.25) / np.sqrt(3))
    assert np.isclose(e["errorMaxRel"], 1.0 / 2.5)
    assert np.isclose(e["drift"], (0 - 0.5 + 1.0) / 3)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("uk,parts", [(13, ((128, 0, 0, 4), (128, 1, 4, 13))), (11, ((16, 0, 0, 4),))])
def test_ulist_leaves_match_live_reference(uk, parts):
    """The oracle's U-list evaluation (leaf ranges of an adaptive octree, fused single + double layer
    for PVel) equals the reference's SimpleKernel summed over each leaf's U-list sources.  Each
    target sums ~10^4 sources in another order than the reference loop, so the bound is the FP64
    parity bound of the path, 1e-12."""
    import stkfmm_amd as S
    from stkfmm_amd import workloads as W

    x = W.pvel_points(12000)
    tr = S.Tree(x, 400)
    xs = x[tr.perm]
    rng = np.random.default_rng(7)
    v = source_values(uk, tr.n, rng)
    leaves = np.unique(np.concatenate([rng.choice(tr.n_leaves, 3, replace=False),
                                       [np.argmax(tr.leaf_end - tr.leaf_begin)]]))
    got = O.ranges(uk, xs, v, xs, tr.leaf_begin, tr.leaf_end, tr.ulist_ptr, tr.rng_begin, tr.rng_end, sel=leaves,
                   trg_value=np.zeros((tr.n, O.trg_dim(uk))))
    kt = O.trg_dim(uk)
    for leaf in leaves:
        b, e = tr.leaf_begin[leaf], tr.leaf_end[leaf]
        src = np.concatenate([np.arange(tr.rng_begin[r], tr.rng_end[r])
                              for r in range(tr.ulist_ptr[leaf], tr.ulist_ptr[leaf + 1])])
        ref = sum(O.simplekernel_direct(K, layer, c1 - c0, kt, xs[src], v[src, c0:c1], xs[b:e])
                  for K, layer, c0, c1 in parts)
        l2, mx = O.max_error(got[b:e], ref)
        assert l2 < 1e-12 and mx < 1e-12, (leaf, l2, mx)
