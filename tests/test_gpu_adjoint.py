"""GPU parity: adjoint kernels (stkfmm_p2p_vjp_device) and the torch.autograd.Function."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.conftest import source_values
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [0, 2, 6, 7, 8, 9, 10, 11, 12, 13])
def test_vjp_vs_oracle(uk, prec):
    say(f"vjp uker={uk} prec={prec}")
    rng = np.random.default_rng(500 + uk)
    ns, nt = 300, 210
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:9], rng.random((nt - 9, 3))])
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    ref = O.direct_vjp(uk, src, v, trg, w)
    got = S.vjp_device(uk, to_dev(src), to_dev(v), to_dev(trg), to_dev(w), precision=prec)
    tol = 1e-11 if prec == 0 else 1e-5
    for g, r, name in zip(got, ref, ("g_src_value", "g_src_coord", "g_trg_coord")):
        l2, mx = O.max_error(host(g), r)
        assert l2 < tol and mx < tol, (name, l2, mx)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [11, 13])
def test_vjp_multi_tile_waves(uk, prec, monkeypatch):
    # 8200 x 4100 (neither a multiple of the tiles) split into 2 waves per CU: on a 256-CU device each
    # wave of the single-pass adjoint then walks 3 target tiles x 3 source tiles, so the per-wave
    # loops, the partial read-modify-write and the ragged last tiles all run
    monkeypatch.setenv("STKFMM_VJP_WAVES_PER_CU", "2")
    say(f"vjp multi-tile uker={uk} prec={prec} ns=4100 nt=8200 waves/CU=2")
    rng = np.random.default_rng(700 + uk)
    ns, nt = 4100, 8200
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:37], rng.random((nt - 37, 3))])
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    ref = O.direct_vjp(uk, src, v, trg, w)
    got = S.vjp_device(uk, to_dev(src), to_dev(v), to_dev(trg), to_dev(w), precision=prec)
    tol = 1e-11 if prec == 0 else 1e-5
    for g, r, name in zip(got, ref, ("g_src_value", "g_src_coord", "g_trg_coord")):
        l2, mx = O.max_error(host(g), r)
        assert l2 < tol and mx < tol, (name, l2, mx)


@pytest.mark.parametrize("uk", [0, 2, 6, 7, 8, 9, 10, 11, 12, 13])
def test_vjp_mixed_four_targets_per_lane(uk):
    # mixed mode takes 4 targets per lane from 4,096 targets: 4,133 targets (a ragged last tile of 37)
    # against 700 sources (a ragged last source tile), every micro-kernel with an adjoint
    say(f"vjp mixed T=4 uker={uk} ns=700 nt=4133")
    rng = np.random.default_rng(750 + uk)
    ns, nt = 700, 4133
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:11], rng.random((nt - 11, 3))])
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    ref = O.direct_vjp(uk, src, v, trg, w)
    got = S.vjp_device(uk, to_dev(src), to_dev(v), to_dev(trg), to_dev(w), precision=S.PREC.MIXED)
    for g, r, name in zip(got, ref, ("g_src_value", "g_src_coord", "g_trg_coord")):
        l2, mx = O.max_error(host(g), r)
        assert l2 < 1e-5 and mx < 1e-5, (name, l2, mx)


@pytest.mark.parametrize("uk", [0, 2, 6, 7, 9, 10, 11, 12])
def test_vjp_fp64_four_targets_per_lane(uk):
    # FP64 takes 4 targets per lane from 4,096 targets for these micro-kernels: 4,133 targets (a ragged
    # last tile of 37, coincident points included) against 700 sources, to the FP64 bar
    say(f"vjp fp64 T=4 uker={uk} ns=700 nt=4133")
    rng = np.random.default_rng(790 + uk)
    ns, nt = 700, 4133
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:11], rng.random((nt - 11, 3))])
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    ref = O.direct_vjp(uk, src, v, trg, w)
    got = S.vjp_device(uk, to_dev(src), to_dev(v), to_dev(trg), to_dev(w), precision=S.PREC.FP64)
    for g, r, name in zip(got, ref, ("g_src_value", "g_src_coord", "g_trg_coord")):
        l2, mx = O.max_error(host(g), r)
        assert l2 < FP64_TOL and mx < FP64_TOL, (name, l2, mx)


@pytest.mark.parametrize("uk", [0, 2, 6, 7, 8, 9, 10, 11, 12, 13])
def test_vjp_mixed_eight_targets_per_lane(uk, monkeypatch):
    # 8 targets per lane (mixed mode's default for PVel and RPY from 8,192 targets; STKFMM_VJP_T for
    # the rest): 8,300 targets (a ragged last tile of 108) against 650 sources, every micro-kernel
    # with an adjoint
    monkeypatch.setenv("STKFMM_VJP_T", "8")
    say(f"vjp mixed T=8 uker={uk} ns=650 nt=8300")
    rng = np.random.default_rng(770 + uk)
    ns, nt = 650, 8300
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:13], rng.random((nt - 13, 3))])
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    ref = O.direct_vjp(uk, src, v, trg, w)
    got = S.vjp_device(uk, to_dev(src), to_dev(v), to_dev(trg), to_dev(w), precision=S.PREC.MIXED)
    for g, r, name in zip(got, ref, ("g_src_value", "g_src_coord", "g_trg_coord")):
        l2, mx = O.max_error(host(g), r)
        assert l2 < 1e-5 and mx < 1e-5, (name, l2, mx)


def test_vjp_targets_per_lane_override_is_checked(monkeypatch):
    # STKFMM_VJP_T: 1, 2, 4 (FP64: not for the double-layer or fused PVel) and 8 in mixed mode;
    # anything else is refused with an argument error
    rng = np.random.default_rng(760)
    src, trg = to_dev(rng.random((70, 3))), to_dev(rng.random((80, 3)))
    v, w = to_dev(rng.normal(size=(70, 4))), to_dev(rng.normal(size=(80, 4)))
    for bad, prec in (("3", S.PREC.MIXED), ("8", S.PREC.FP64), ("16", S.PREC.MIXED), ("0", S.PREC.MIXED)):
        monkeypatch.setenv("STKFMM_VJP_T", bad)
        with pytest.raises(S.ArgumentError):
            S.vjp_device(S.UKER.STOKES_PVEL, src, v, trg, w, precision=prec)
    v13 = to_dev(rng.normal(size=(70, 13)))
    for uk, vv in ((S.UKER.STOKES_PVEL_FUSED, v13), (S.UKER.STOKES_DOUBLEPVEL, to_dev(rng.normal(size=(70, 9))))):
        monkeypatch.setenv("STKFMM_VJP_T", "4")
        with pytest.raises(S.ArgumentError):
            S.vjp_device(uk, src, vv, trg, w, precision=S.PREC.FP64)
    ref = O.direct_vjp(7, host(src), host(v), host(trg), host(w))
    for t in ("4", "8"):
        monkeypatch.setenv("STKFMM_VJP_T", t)
        got = S.vjp_device(S.UKER.STOKES_PVEL, src, v, trg, w, precision=S.PREC.MIXED)
        for g, r in zip(got, ref):
            assert O.max_error(host(g), r)[1] < 1e-5, t


def test_autograd_function_matches_vjp():
    import torch
    from stkfmm_amd.autograd import p2p
    say("autograd")
    rng = np.random.default_rng(600)
    src = to_dev(rng.random((120, 3))).requires_grad_(True)
    val = to_dev(source_values(13, 120, rng)).requires_grad_(True)
    trg = to_dev(rng.random((90, 3))).requires_grad_(True)
    w = to_dev(rng.normal(size=(90, 4)))
    u = p2p(S.UKER.STOKES_PVEL_FUSED, src, val, trg)
    (u * w).sum().backward()
    ref = O.direct_vjp(13, host(src), host(val), host(trg), host(w))
    for g, r in zip((val.grad, src.grad, trg.grad), ref):
        assert O.max_error(host(g), r)[0] < 1e-11
    # torch's own finite-difference check on a tiny problem
    s2 = to_dev(rng.random((4, 3))).requires_grad_(True)
    v2 = to_dev(source_values(11, 4, rng)).requires_grad_(True)
    t2 = to_dev(rng.random((3, 3))).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b, c: p2p(S.UKER.RPY_ULAPU, a, b, c), (s2, v2, t2), eps=1e-6,
                                    atol=1e-6, rtol=1e-5)


@pytest.mark.parametrize("uk", [0, 2, 6, 9, 10, 12])
def test_autograd_gradcheck_more_kernels(uk):
    # torch's finite-difference check through the autograd Function (FP64, tiny problem; for the
    # regularised Stokeslet the gradient includes d/d eps)
    import torch
    from stkfmm_amd.autograd import p2p
    say(f"gradcheck uker={uk}")
    rng = np.random.default_rng(620 + uk)
    s = to_dev(rng.random((5, 3))).requires_grad_(True)
    v = to_dev(source_values(uk, 5, rng)).requires_grad_(True)
    t = to_dev(rng.random((4, 3))).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b, c: p2p(uk, a, b, c), (s, v, t), eps=1e-6, atol=1e-6, rtol=1e-5)


@pytest.mark.slow
def test_workload4_adjoint_full_size():
    """65,536 x 65,536 PVel: gradients vs central differences (h = 1e-6) on 64 sampled coordinates and
    vs the oracle gradient on sampled entries."""
    import torch
    from stkfmm_amd.autograd import p2p
    say("workload 4 full size")
    src, val, trg, w = W.adjoint_pvel()
    d_src = to_dev(src).requires_grad_(True)
    d_val = to_dev(val).requires_grad_(True)
    d_trg = to_dev(trg).requires_grad_(True)
    d_w = to_dev(w)
    u = p2p(S.UKER.STOKES_PVEL, d_src, d_val, d_trg)
    (u * d_w).sum().backward()
    g_src, g_val, g_trg = host(d_src.grad), host(d_val.grad), host(d_trg.grad)
    rng = np.random.default_rng(4)
    h = 1e-6

    def field(s, v, t):
        out = zeros(t.shape[0], 4)
        S.uker_device(S.UKER.STOKES_PVEL, s, v, t, out)
        return out

    ds, dv, dt = to_dev(src), to_dev(val), to_dev(trg)
    eps = np.finfo(np.float64).eps

    def check(fd_h, fd_2h, g, round_off, what):
        # central differences carry an O(h^2) truncation error; D(2h) - D(h) estimates three times it
        tol = 1e-6 * abs(g) + abs(fd_2h - fd_h) + round_off + 1e-9
        assert abs(fd_h - g) <= tol, (what, fd_h, fd_2h, g)

    # target coordinates: only u_t changes
    for _ in range(22):
        t_i, j = int(rng.integers(len(trg))), int(rng.integers(3))
        fds = []
        for hh in (h, 2 * h):
            tp, tm = trg[t_i:t_i + 1].copy(), trg[t_i:t_i + 1].copy()
            tp[0, j] += hh
            tm[0, j] -= hh
            up, um = field(ds, dv, to_dev(tp)), field(ds, dv, to_dev(tm))
            fds.append(float(((up - um) * d_w[t_i:t_i + 1]).sum()) / (2 * hh))
        mag = float((up.abs() * d_w[t_i:t_i + 1].abs()).sum())
        check(fds[0], fds[1], g_trg[t_i, j], 8 * eps * mag / h, ("trg", t_i, j))
    # source coordinates and forces: every target changes
    for kind in ("src", "val"):
        for _ in range(21):
            s_i, j = int(rng.integers(len(src))), int(rng.integers(3))
            arr = src if kind == "src" else val
            g = g_src[s_i, j] if kind == "src" else g_val[s_i, j]
            fds = []
            for hh in (h, 2 * h):
                p, m = arr.copy(), arr.copy()
                p[s_i, j] += hh
                m[s_i, j] -= hh
                if kind == "src":
                    lp, lm = field(to_dev(p), dv, dt), field(to_dev(m), dv, dt)
                else:
                    lp, lm = field(ds, to_dev(p), dt), field(ds, to_dev(m), dt)
                fds.append(float(((lp - lm) * d_w).sum()) / (2 * hh))
            mag = float((lp.abs() * d_w.abs()).sum())
            # summing 4 n terms: rounding of L is about eps * sum |w u| * log2(4 n)
            check(fds[0], fds[1], g, 20 * eps * mag / h, (kind, s_i, j))
    # oracle gradients on sampled sources / targets
    si = rng.choice(len(src), 48, replace=False)
    ti = rng.choice(len(trg), 48, replace=False)
    gv_ref, gs_ref, _ = O.direct_vjp(7, src[si], val[si], trg, w)
    _, _, gt_ref = O.direct_vjp(7, src, val, trg[ti], w[ti])
    assert O.max_error(g_val[si], gv_ref)[0] < 1e-11
    assert O.max_error(g_src[si], gs_ref)[0] < 1e-11
    assert O.max_error(g_trg[ti], gt_ref)[0] < 1e-11


@pytest.mark.parametrize("uk", [7, 13, 11])
def test_workload4_adjoint_full_size_mixed(uk):
    """65,536 x 65,536 (the workload's PVel points; fused PVel and RPY with values of their own),
    mixed precision at its default targets per lane: every gradient against the oracle's FP64
    gradient on sampled sources and targets (the whole source set against sampled targets and vice
    versa), to the mixed mode's bar, and every entry of the mixed gradient against the FP64 one."""
    say(f"workload 4 full size, mixed, uker={uk}")
    src, val, trg, w = W.adjoint_pvel()
    if uk != 7:
        g = np.random.default_rng(45 + uk)
        val = source_values(uk, len(src), g)
        w = g.normal(size=(len(trg), O.trg_dim(uk)))
    ds, dv, dt, dw = to_dev(src), to_dev(val), to_dev(trg), to_dev(w)
    g_val, g_src, g_trg = (host(x) for x in S.vjp_device(uk, ds, dv, dt, dw, precision=S.PREC.MIXED))
    f_val, f_src, f_trg = (host(x) for x in S.vjp_device(uk, ds, dv, dt, dw))
    rng = np.random.default_rng(44)
    si = rng.choice(len(src), 64, replace=False)
    ti = rng.choice(len(trg), 64, replace=False)
    gv_ref, gs_ref, _ = O.direct_vjp(uk, src[si], val[si], trg, w)
    _, _, gt_ref = O.direct_vjp(uk, src, val, trg[ti], w[ti])
    for got, ref, what in ((g_val[si], gv_ref, "src_value"), (g_src[si], gs_ref, "src_coord"),
                           (g_trg[ti], gt_ref, "trg_coord")):
        l2, mx = O.max_error(got, ref)
        assert l2 < MIXED_TOL and mx < MIXED_TOL, (what, l2, mx)
    for got, ref in ((g_val, f_val), (g_src, f_src), (g_trg, f_trg)):
        assert np.all(np.isfinite(got))
        l2, mx = O.max_error(got, ref)
        assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)
