"""GPU: stkfmm_p2p_eval_device_noalloc (the direct sum that allocates no device memory) against the
plan-based stkfmm_p2p_eval_device and the oracle, through the C ABI."""
import ctypes

import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import _capi
from tests.conftest import source_values
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu

# (KERNEL, PPKERNEL) pairs and their micro-kernels: Laplace PGrad, Stokes PVel single layer, RPY
CASES = [(1, 1, 2), (128, 1, 7), (16, 1, 11)]


def points(seed, ns, nt, n_coinc=16):
    rng = np.random.default_rng(seed)
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:n_coinc], rng.random((nt - n_coinc, 3))])
    return rng, src, trg


def call(entry, K, p, prec, d_src, d_val, d_trg, out):
    rc = entry(K, p, prec, d_src.shape[0], ctypes.c_void_p(d_src.data_ptr()), ctypes.c_void_p(d_val.data_ptr()),
               d_trg.shape[0], ctypes.c_void_p(d_trg.data_ptr()), ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 0, _capi.last_error()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("K,p,uk", CASES)
def test_noalloc_matches_plan_entry_twice(K, p, uk, prec):
    L = _capi.lib()
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    # two calls in a row on the same scratch array, the second smaller than the first
    for ns, nt, seed in ((6000, 3000, 1), (700, 333, 2)):
        say(f"noalloc K={K} p2p={p} prec={prec} ns={ns} nt={nt}")
        rng, src, trg = points(100 * K + seed, ns, nt)
        v = source_values(uk, ns, rng)
        d_src, d_val, d_trg = to_dev(src), to_dev(v), to_dev(trg)
        kt = O.trg_dim(uk)
        prefill = rng.normal(size=(nt, kt))
        a, b = to_dev(prefill), to_dev(prefill)
        call(L.stkfmm_p2p_eval_device, K, p, prec, d_src, d_val, d_trg, a)
        call(L.stkfmm_p2p_eval_device_noalloc, K, p, prec, d_src, d_val, d_trg, b)
        ha, hb = host(a), host(b)
        # same kernels, same chunking at these sizes: bitwise identical
        assert np.array_equal(ha, hb)
        l2, mx = O.max_error(hb - prefill, O.direct(uk, src, v, trg))
        assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_noalloc_long_chunks(prec):
    """65,536 x 65,536: the plan's partial sums (134 MB) exceed the scratch array, so the sources are
    cut into fewer chunks; the result stays within the precision's bar of the plan-based entry's."""
    L = _capi.lib()
    K, p, uk = 128, 1, 7
    n = 65_536
    say(f"noalloc long chunks prec={prec} n={n}")
    rng, src, trg = points(77, n, n)
    v = source_values(uk, n, rng)
    d_src, d_val, d_trg = to_dev(src), to_dev(v), to_dev(trg)
    for _ in range(2):
        a, b = zeros(n, 4), zeros(n, 4)
        call(L.stkfmm_p2p_eval_device, K, p, prec, d_src, d_val, d_trg, a)
        call(L.stkfmm_p2p_eval_device_noalloc, K, p, prec, d_src, d_val, d_trg, b)
        l2, mx = O.max_error(host(b), host(a))
        tol = FP64_TOL if prec == 0 else MIXED_TOL
        assert l2 < tol and mx < tol, (l2, mx)


def test_noalloc_rejects_oversize():
    """Targets whose tiles alone exceed the scratch array: a clear error, nothing launched."""
    K, p = 256, 1  # PVelGrad single layer, 16 outputs per target
    nt = 600_000
    say(f"noalloc oversize nt={nt}")
    src = to_dev(np.random.default_rng(3).random((64, 3)))
    val = to_dev(np.zeros((64, 4)))
    trg = zeros(nt, 3)
    out = zeros(nt, 16)
    with pytest.raises(S.ArgumentError, match="scratch"):
        S.p2p_device(K, p, src, val, trg, out, no_alloc=True)
    assert float(host(out).max()) == 0.0
