"""The library's per-pair micro-kernels (stkfmm_amd/csrc/ukernels*.hpp: stage, eval, post), compiled for
the CPU by tests/cxx/ukernel_host.cpp, against the oracle -- the algebra of every GPU micro-kernel
(its rewrites of the reference expression) checked without a GPU.

FP64: to 1e-12 of each output's largest entry, the parity bar of the FP64 mode (the rewritten
expressions round differently from the reference's order; for outputs that cancel, such as
f_i r^2 - 3 r_i (f.r), that difference reaches a few 1e-13).  FP32: the same pipeline in single precision, to a bound of FP32 rounding amplified by
the kernel's cancellation, on pairs at ordinary and at close range (the unit-direction forms of the
r^-7 / r^-9 kernels must not overflow)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import RADIUS_UKERS, ROOT

LIB = os.path.join(ROOT, "tests", "cxx", "libukernel_host.so")
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    if not os.path.exists(LIB):
        pytest.skip("tests/cxx/libukernel_host.so not built (make -C stkfmm_amd/csrc)")
    L = ctypes.CDLL(LIB)
    L.ukernel_host_pair.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    return L


def host_pair_raw(uk, prec, r, f):
    r = np.ascontiguousarray(r, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.zeros(32)
    kt = lib().ukernel_host_pair(uk, prec, r.ctypes.data_as(_dp), f.ctypes.data_as(_dp), u.ctypes.data_as(_dp))
    return kt, u[:max(kt, 0)]


def host_pair(uk, prec, r, f):
    kt, u = host_pair_raw(uk, prec, r, f)
    assert kt == O.trg_dim(uk)
    return u


def values(uk, rng):
    f = rng.uniform(-1, 1, O.src_dim(uk))
    if uk in RADIUS_UKERS:
        f[-1] = abs(f[-1]) * 0.01 + 1e-3
    return f


@pytest.mark.parametrize("uk", range(25))
def test_fp64_pairs_match_oracle(uk):
    rng = np.random.default_rng(400 + uk)
    for _ in range(200):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-3, 1)
        f = values(uk, rng)
        a, b = host_pair(uk, 0, r, f), O.pair(uk, r, f)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


@pytest.mark.parametrize("uk", range(25))
def test_coincident_pair_is_zero(uk):
    f = np.ones(O.src_dim(uk))
    if uk in (9, 20, 24):
        f[-1] = 0.0
    for prec in (0, 1):
        assert np.all(host_pair(uk, prec, np.zeros(3), f) == 0.0)


@pytest.mark.parametrize("uk", range(25))
def test_fp32_pairs_match_oracle(uk):
    # per-pair bound: each output is a sum of a few terms; FP32 rounding of the inputs and of each
    # operation gives a relative error of a few 1e-7 of the largest term.  Close range included
    # (r down to 1e-4 in the level-scaled frame).
    rng = np.random.default_rng(500 + uk)
    for _ in range(300):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-4, 0.5)
        f = values(uk, rng)
        a, b = host_pair(uk, 1, r, f), O.pair(uk, r, f)
        assert np.all(np.isfinite(a))
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 2e-5 * scale, (r, np.abs(a - b).max() / scale)


@pytest.mark.parametrize("uk", range(25))
def test_fp64_close_pairs_match_reference_rounding(uk):
    # r down to 1e-8: where the reference's in-place updates of its coefficient lose accuracy
    # (e.g. (f.r - TrD) + TrD in the Stokes gradient kernels), the FP64 kernels follow the same
    # sequence and stay within 1e-12 of it
    rng = np.random.default_rng(600 + uk)
    for _ in range(200):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-8, -3)
        f = values(uk, rng)
        a, b = host_pair(uk, 0, r, f), O.pair(uk, r, f)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


@pytest.mark.parametrize("uk", range(25))
def test_unmasked_form_equals_masked_for_separated_pairs(uk):
    # the item kernel runs eval<R, false> on source tiles shown to hold no coincident pair; for
    # r2 > 0 it must give the masked form's result bit for bit, in FP64 and FP32
    rng = np.random.default_rng(700 + uk)
    r0 = np.full(3, 0.5)
    if host_pair_raw(uk, 2, r0, values(uk, rng))[0] == -2:
        pytest.skip("micro-kernel without an unmasked form")
    for _ in range(100):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-6, 1)
        f = values(uk, rng)
        for prec in (0, 1):
            kt, a = host_pair_raw(uk, prec, r, f)
            kt2, b = host_pair_raw(uk, prec | 2, r, f)
            assert kt == kt2 and np.array_equal(a, b), (prec, r)


CORE = range(14)


@pytest.mark.parametrize("uk", CORE)
def test_fp32_close_range_finite(uk):
    # level-scaled pair distances down to 1e-8: no power of 1/r above what FP32 holds there (the
    # literal r^-5 / r^-7 of the Laplace dipole-gradient and gradient-of-gradient kernels overflowed)
    rng = np.random.default_rng(800 + uk)
    for _ in range(200):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-8, -5)
        f = values(uk, rng)
        a, b = host_pair(uk, 1, r, f), O.pair(uk, r, f)
        assert np.all(np.isfinite(a)), r
        # FP32 rounding amplified by cancellation (the double-layer pressure tr D - 3 n.D.n can be
        # 1e-3 of its terms); a literal power that overflowed or lost range would miss by far more
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 2e-4 * scale, (r, np.abs(a - b).max() / scale)


@pytest.mark.parametrize("prec", [0, 1])
def test_nan_displacement_is_masked(prec):
    # the reference's mask r2 > 0 is false for NaN, so 1/r is 0: the Laplace potential of such a pair
    # is 0 (kernels with r_i factors give NaN, as the reference does)
    for r in ([np.nan, 0.1, 0.2], [0.3, np.nan, np.nan]):
        assert np.array_equal(host_pair(0, prec, np.array(r), np.array([1.5])), [0.0])


VJP_UKERS = (0, 2, 6, 7, 8, 9, 10, 11, 12, 13)


def host_vjp(uk, prec, r, f, w):
    L = lib()
    L.vjp_host_pair.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
    r, f, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (r, f, w))
    gr, gf = np.zeros(3), np.zeros(32)
    ks = L.vjp_host_pair(uk, prec, r.ctypes.data_as(_dp), f.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                         gr.ctypes.data_as(_dp), gf.ctypes.data_as(_dp))
    assert ks == O.src_dim(uk)
    return gr, gf[:ks]


@pytest.mark.parametrize("uk", VJP_UKERS)
def test_vjp_pairs_match_oracle(uk):
    # the per-pair adjoints of the device kernels (stkfmm_amd/csrc/vjp_pairs.hpp) against the oracle's
    # (which is pinned against central finite differences of the forward): FP64 to 1e-12 of the
    # gradient's largest entry, FP32 to FP32 rounding amplified by the cancellation of the position
    # gradient's terms; pairs from 1e-3 to 10 apart
    rng = np.random.default_rng(1200 + uk)
    for _ in range(200):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-3, 1)
        f = values(uk, rng)
        w = rng.uniform(-1, 1, O.trg_dim(uk))
        gr_o, gf_o = O.vjp_pair(uk, r, f, w)
        ref = np.concatenate([gr_o, gf_o])
        scale = np.abs(ref).max()
        for prec, tol in ((0, 1e-12), (1, 2e-5)):
            gr, gf = host_vjp(uk, prec, r, f, w)
            err = np.abs(np.concatenate([gr, gf]) - ref).max() / scale
            assert err <= tol, (prec, r, err)


@pytest.mark.parametrize("prec", [0, 1])
def test_vjp_coincident_pair(prec):
    # the forward is masked at r = 0, so the pair's adjoint is exactly zero -- except the regularised
    # Stokeslet (uker 9), finite there (rho^2 = r^2 + eps^2 > 0): its adjoint matches the oracle's
    rng = np.random.default_rng(1300)
    for uk in VJP_UKERS:
        f, w = values(uk, rng), rng.uniform(-1, 1, O.trg_dim(uk))
        gr, gf = host_vjp(uk, prec, np.zeros(3), f, w)
        if uk == 9:
            ref = np.concatenate(O.vjp_pair(uk, np.zeros(3), f, w))
            tol = 1e-12 if prec == 0 else 2e-5
            assert np.abs(np.concatenate([gr, gf]) - ref).max() <= tol * np.abs(ref).max()
        else:
            assert not gr.any() and not gf.any(), uk
