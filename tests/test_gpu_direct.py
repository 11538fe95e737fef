"""GPU parity: direct sums (STKFMM::evaluateKernel path and Ker_t entries) against the oracle."""
import ctypes

import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.conftest import UKER_REF, source_values
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu

UKERS = {(1, 1): 2, (1, 2): 3, (1, 4): 2, (2, 1): 4, (2, 2): 5, (2, 4): 4, (8, 1): 6, (8, 4): 6, (16, 1): 11,
         (16, 4): 12, (32, 1): 9, (32, 4): 6, (128, 1): 7, (128, 2): 8, (128, 4): 7,
         (4, 1): 22, (4, 4): 4, (64, 1): 20, (64, 4): 21, (256, 1): 14, (256, 2): 15, (256, 4): 14,
         (512, 1): 16, (512, 2): 17, (512, 4): 16, (1024, 1): 18, (1024, 2): 19, (1024, 4): 18}


def case(uk, seed, ns=700, nt=333, n_coinc=16):
    rng = np.random.default_rng(seed)
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:n_coinc], rng.random((nt - n_coinc, 3))])
    return src, source_values(uk, ns, rng), trg


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", range(25))
def test_uker_direct_vs_oracle(uk, prec):
    say(f"uker={uk} prec={prec} ns=700 nt=333")
    src, v, trg = case(uk, 1000 + uk)
    ref = O.direct(uk, src, v, trg)
    out = zeros(len(trg), O.trg_dim(uk))
    S.uker_device(uk, to_dev(src), to_dev(v), to_dev(trg), out, precision=prec)
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.parametrize("kp", sorted(UKERS))
def test_evaluate_kernel_host_api_accumulates(kp):
    K, p = kp
    uk = UKERS[kp]
    say(f"evaluate_kernel K={K} p2p={p}")
    src, v, trg = case(uk, 2000 + K + p, ns=517, nt=259)
    prefill = np.random.default_rng(K).normal(size=(len(trg), O.trg_dim(uk)))
    out = prefill.copy()
    S.evaluate_kernel(K, 4, p, src, np.ascontiguousarray(v), trg, out)
    ref = O.direct(uk, src, v, trg)
    l2, mx = O.max_error(out - prefill, ref)
    assert l2 < FP64_TOL and mx < FP64_TOL


@pytest.mark.parametrize("uk", sorted(UKER_REF))
def test_golden_fixtures_through_gpu(golden, uk):
    K, layer = UKER_REF[uk]
    tag = f"K{K}_L{layer}"
    say(f"golden {tag}")
    p = S.PPKERNEL.DLS2T if layer else S.PPKERNEL.SLS2T
    expect = golden[tag + "_out"]
    out = np.zeros_like(expect)
    S.evaluate_kernel(K, 1, p, np.ascontiguousarray(golden[tag + "_src"]), np.ascontiguousarray(golden[tag + "_sv"]),
                      np.ascontiguousarray(golden[tag + "_trg"]), out)
    l2, mx = O.max_error(out, expect)
    assert l2 < FP64_TOL and mx < FP64_TOL


@pytest.mark.parametrize("name", S._capi.KER_T_NAMES)
def test_ker_t_entries(name):
    uk = S._capi.KER_T_UKER[name]
    say(f"Ker_t stkfmm_ker_{name}")
    src, v, trg = case(uk, 3000 + uk, ns=300, nt=200)
    v = np.ascontiguousarray(v)
    out = np.zeros((len(trg), O.trg_dim(uk)))
    dp = ctypes.POINTER(ctypes.c_double)
    f = getattr(S.lib(), "stkfmm_ker_" + name)
    f(src.ctypes.data_as(dp), len(src), v.ctypes.data_as(dp), 1, trg.ctypes.data_as(dp), len(trg),
      out.ctypes.data_as(dp), None)
    l2, mx = O.max_error(out, O.direct(uk, src, v, trg))
    assert l2 < FP64_TOL and mx < FP64_TOL


def test_device_empty_and_sizes_not_multiple_of_tile():
    say("edge sizes")
    for ns, nt in ((1, 1), (1, 130), (65, 1), (63, 257), (129, 127)):
        src, v, trg = case(7, ns * 1000 + nt, ns=ns, nt=nt, n_coinc=min(ns, nt, 1))
        out = zeros(nt, 4)
        S.uker_device(7, to_dev(src), to_dev(v), to_dev(trg), out)
        l2, mx = O.max_error(host(out), O.direct(7, src, v, trg))
        assert mx < FP64_TOL or np.all(O.direct(7, src, v, trg) == 0)
    t = S.p2p_device(S.KERNEL.PVel, S.PPKERNEL.SLS2T, zeros(0, 3), zeros(0, 4), zeros(5, 3), zeros(5, 4))
    assert float(t.abs().sum()) == 0.0


@pytest.mark.slow
def test_workload1_lapgrad_direct_full_size():
    """8192 x 8192 Laplace PGrad direct sum, FP64 (device API) and the host API, vs the oracle."""
    say("workload 1 full size")
    src, q, trg = W.lapgrad_direct()
    ref = O.direct(2, src, q, trg)
    out = zeros(len(trg), 4)
    S.p2p_device(S.KERNEL.LapPGrad, S.PPKERNEL.SLS2T, to_dev(src), to_dev(q), to_dev(trg), out)
    l2, mx = O.max_error(host(out), ref)
    assert l2 < FP64_TOL and mx < FP64_TOL
    h = np.zeros((len(trg), 4))
    S.evaluate_kernel(S.KERNEL.LapPGrad, 0, S.PPKERNEL.SLS2T, src, q, trg, h)
    assert O.max_error(h, ref)[0] < FP64_TOL
    outm = zeros(len(trg), 4)
    S.p2p_device(S.KERNEL.LapPGrad, S.PPKERNEL.SLS2T, to_dev(src), to_dev(q), to_dev(trg), outm,
                 precision=S.PREC.MIXED)
    l2, mx = O.max_error(host(outm), ref)
    assert l2 < MIXED_TOL and mx < MIXED_TOL


@pytest.mark.parametrize("bc", [0, 1], ids=["lds", "scalar"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [0, 6, 7, 13])
def test_nan_coordinates_keep_the_mask(uk, prec, bc):
    # A NaN coordinate makes r2 NaN, which the reference's mask (r2 > 0) zeroes: the Laplace potential
    # stays finite, kernels with r_i factors turn NaN.  Source tiles are only run without the mask
    # when every source is provably separated from the targets; NaN must never pass that test.
    # (Kernels whose rewrite drops a factor r2/r2 -- LapPGrad's p = q/r for the reference's
    # q r^2/r^3 -- give 0 where the reference gives NaN; they are not compared here.)
    say(f"NaN coordinates uker={uk} prec={prec} broadcast={bc}")
    rng = np.random.default_rng(900 + uk)
    ns, nt = 300, 200
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:5], rng.random((nt - 5, 3))])
    v = source_values(uk, ns, rng)
    trg[11, 1] = np.nan
    ref_t = O.direct(uk, src, v, trg)
    src2 = src.copy()
    src2[7, 0] = np.nan
    ref_s = O.direct(uk, src2, v, trg)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    for sc, ref in ((src, ref_t), (src2, ref_s)):
        plan = S.Plan(uk, prec, ns, nt, [0], [nt], None, [0, 1], [0], [ns], tune=dict(broadcast=bc))
        out = zeros(nt, O.trg_dim(uk))
        if uk == S.UKER.STOKES_PVEL_FUSED:
            plan.execute(to_dev(sc), to_dev(v[:, :4]), to_dev(trg), out, src_value2=to_dev(v[:, 4:]))
        else:
            plan.execute(to_dev(sc), to_dev(v), to_dev(trg), out)
        got = host(out)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        rows = ~np.isnan(ref).any(axis=1)
        if rows.any():
            l2, mx = O.max_error(got[rows], ref[rows])
            assert l2 < tol and mx < tol, (l2, mx)
