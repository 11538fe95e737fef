"""Box mapping and value scaling (STKFMM::scaleCoord/wrapCoord, FMMData::scaleSrc/scaleTrg) on the host.

Pinned by the identity they exist for: for every KERNEL and layer, the direct sum in the caller's
coordinates equals scale_trg(direct sum over scale_coord'ed points with scale_src'ed values) --
FMMData::evaluateFMM's sequence (Lib/src/FMMData.cpp:196-200) -- which fixes the power of the scale
factor of every source and output component (a wrong exponent or a missed component fails), using
the oracle's pinned kernels.  Plus the exact arithmetic of the mapping and the wrap."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from tests.conftest import RADIUS_UKERS

KERNELS = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]


def cases():
    for k in KERNELS:
        for p2p in (S.PPKERNEL.SLS2T, S.PPKERNEL.DLS2T):
            try:
                uk = S.get_uker(k, p2p)
            except S.PPKernelNotFound:
                continue
            yield k, int(p2p), int(uk)


@pytest.mark.parametrize("k,p2p,uk", list(cases()))
def test_evaluate_in_scaled_box_equals_direct(k, p2p, uk):
    rng = np.random.default_rng(k + p2p)
    origin = np.array([-3.0, 1.5, 10.0])
    length = 7.3
    sF = 1.0 / length
    src = origin + rng.random((60, 3)) * length
    trg = origin + rng.random((40, 3)) * length
    ks = O.src_dim(uk)
    v = rng.uniform(-1, 1, (60, ks))
    if uk in RADIUS_UKERS:
        v[:, -1] = np.abs(v[:, -1]) * 0.05 + 0.01
    direct = O.direct(uk, src, v, trg)

    s2, t2 = src.copy(), trg.copy()
    S.scale_coord(s2, origin, sF)
    S.scale_coord(t2, origin, sF)
    v2 = v.copy()
    if p2p == S.PPKERNEL.SLS2T:
        S.scale_src(k, v2, None, sF)
    else:
        S.scale_src(k, None, v2, sF)
    out = O.direct(uk, s2, v2, t2)
    S.scale_trg(k, out, sF)
    np.testing.assert_allclose(out, direct, rtol=1e-11, atol=1e-12 * np.abs(direct).max())


def test_scale_coord_arithmetic_and_wrap():
    c = np.array([[1.0, 2.0, 3.0], [-1.0, 0.25, 11.0]])
    o = np.array([0.5, 0.0, 1.0])
    a = c.copy()
    S.scale_coord(a, o, 0.25)
    np.testing.assert_array_equal(a, (c - o) * 0.25)
    for pbc, axes in ((S.PAXIS.PX, [0]), (S.PAXIS.PXY, [0, 1]), (S.PAXIS.PXYZ, [0, 1, 2])):
        b = c.copy()
        S.scale_coord(b, o, 0.25, pbc)
        e = (c - o) * 0.25
        for j in axes:
            e[:, j] = e[:, j] - np.floor(e[:, j])
        np.testing.assert_array_equal(b, e)


def test_scale_rules_by_hand():
    # a few multipliers written out from FMMData.cpp:319-470 and :285-317
    sF = 0.3
    t = np.ones((1, 10))
    S.scale_trg(S.KERNEL.LapQPGradGrad, t, sF)
    sf3 = sF * sF * sF
    assert t[0, 0] == sf3 and t[0, 1] == sF * sf3 and t[0, 9] == sF * (sF * sf3)
    t = np.ones((1, 6))
    S.scale_trg(S.KERNEL.RPY, t, sF)
    assert list(t[0]) == [sF] * 3 + [sF * sF * sF] * 3
    v = np.ones((2, 7))
    S.scale_src(S.KERNEL.StokesRegVelOmega, v, None, sF)
    assert list(v[0]) == [1, 1, 1, sF, sF, sF, sF]
    v = np.ones((2, 3))
    S.scale_src(S.KERNEL.Stokes, v, None, sF)
    assert np.all(v == 1)


def test_scaling_argument_errors():
    with pytest.raises(S.KernelNotFound):
        S.scale_trg(2048, np.ones((1, 4)), 0.5)
    with pytest.raises(S.ArgumentError):
        S.scale_coord(np.ones((2, 3)), [0, 0, 0], 1.0, 7)
