"""GPU parity: plans that order their points along a Morton curve at every execute (stkfmm_tune
sort_points) -- direct sums and independent instances of unsorted points -- against the oracle, with
the caller's output permutation, graph replay and the work-list shapes the ordering refuses."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.conftest import source_values
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu

SORT = dict(sort_points=1)


def instance_plan(sizes, offsets, prec, uk=6, tune=None):
    n = int(offsets[-1])
    gb, ge = offsets[:-1], offsets[1:]
    return S.Plan(uk, prec, n, n, gb, ge, None, np.arange(len(sizes) + 1), gb, ge, tune)


@pytest.mark.parametrize("bc", [0, 1], ids=["lds", "scalar"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_sorted_instances_vs_oracle(prec, bc):
    say(f"sorted instances prec={prec} broadcast={bc}")
    sizes, offsets, coord, f = W.stokeslet_instances(n_inst=24, n_lo=100, n_hi=5000)
    plan = instance_plan(sizes, offsets, prec, tune=dict(SORT, broadcast=bc))
    out = zeros(len(coord), 3)
    d_c, d_f = to_dev(coord), to_dev(f)
    plan.execute(d_c, d_f, d_c, out)
    got = host(out)
    # targets given as a separate copy of the same points: ordered on their own, same bits
    out2 = zeros(len(coord), 3)
    plan.execute(d_c, d_f, d_c.clone(), out2)
    assert np.array_equal(host(out2), got)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    for i in range(len(sizes)):
        a, b = offsets[i], offsets[i + 1]
        l2, mx = O.max_error(got[a:b], O.direct(6, coord[a:b], f[a:b], coord[a:b]))
        assert l2 < tol and mx < tol, (i, l2, mx)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [13, 11, 2, 9, 4])
def test_sorted_separate_sets_shared_ranges_and_perm(uk, prec):
    """Groups with their own target sets against source ranges (two groups share one range, one
    group has a single target, one no targets), clustered points, the caller's output permutation
    and accumulation into a prefilled output."""
    say(f"sorted general work list uker={uk} prec={prec}")
    g = np.random.default_rng(40 + uk)
    ns_sizes, nt_sizes = [3000, 700, 64, 2500], [1, 2100, 0, 900, 1500]
    src = np.concatenate([g.random((n, 3)) if k % 2 else 0.3 + 0.01 * g.normal(size=(n, 3))
                          for k, n in enumerate(ns_sizes)])
    trg = np.concatenate([g.random((n, 3)) for n in nt_sizes])
    # a few targets of group 1 on its own source points (coincident pairs)
    trg[5:12] = src[3100:3107]
    so = np.concatenate([[0], np.cumsum(ns_sizes)])
    to = np.concatenate([[0], np.cumsum(nt_sizes)])
    grp_src = [0, 1, 2, 0, 3]  # groups 0 and 3 share source range 0
    ks = O.src_dim(uk)
    v = source_values(uk, len(src), g, ks=ks)
    gp = np.arange(len(nt_sizes) + 1)
    rb, re = so[grp_src], so[np.array(grp_src) + 1]
    kt = O.trg_dim(uk)
    perm = g.permutation(len(trg)).astype(np.int32)
    prefill = g.normal(size=(len(trg), kt))
    outs = {}
    for tune in (None, SORT):
        plan = S.Plan(uk, prec, len(src), len(trg), to[:-1], to[1:], None, gp, rb, re, tune)
        out = to_dev(prefill)
        if uk == 13:
            plan.execute(to_dev(src), to_dev(v[:, :4]), to_dev(trg), out, src_value2=to_dev(v[:, 4:]),
                         perm=to_dev(perm))
        else:
            plan.execute(to_dev(src), to_dev(v), to_dev(trg), out, perm=to_dev(perm))
        outs[tune is not None] = host(out)
    ref = np.zeros((len(trg), kt))
    for k in range(len(nt_sizes)):
        if nt_sizes[k]:
            r = O.direct(uk, src[rb[k]:re[k]], v[rb[k]:re[k]], trg[to[k]:to[k + 1]])
            ref[perm[to[k]:to[k + 1]]] = r
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    for key in (False, True):
        l2, mx = O.max_error(outs[key] - prefill, ref)
        assert l2 < tol and mx < tol, (key, l2, mx)


def test_sorted_large_segment_and_graph_replay():
    """One 70,000 x 70,000 direct sum (grid level 4, many sort blocks): eager result against the
    oracle on sampled targets, then graph replays bitwise equal to eager, with eager work between."""
    import torch
    n = 70_000
    say(f"sorted large segment n={n}")
    g = np.random.default_rng(5)
    x = np.concatenate([g.random((n // 2, 3)), W.wrap(0.5 + 0.05 * g.normal(size=(n - n // 2, 3)))])
    g.shuffle(x)
    f = g.normal(size=(n, 3))
    for prec in (1, 0):
        plan = S.Plan(6, prec, n, n, [0], [n], None, [0, 1], [0], [n], SORT)
        d_x, d_f = to_dev(x), to_dev(f)
        out = zeros(n, 3)
        plan.execute(d_x, d_f, d_x, out)
        eager = host(out)
        rows = g.choice(n, 700, replace=False)
        l2, mx = O.max_error(eager[rows], O.direct(6, x, f, x[rows]))
        tol = FP64_TOL if prec == 0 else MIXED_TOL
        assert l2 < tol and mx < tol, (prec, l2, mx)
        stream = torch.cuda.Stream()
        out_g = zeros(n, 3)
        torch.cuda.synchronize()
        plan.capture_graph(d_x, d_f, d_x, out_g, stream=stream)
        for rep in range(3):
            out_g.zero_()
            torch.cuda.synchronize()
            plan.launch_graph(stream=stream)
            stream.synchronize()
            assert np.array_equal(host(out_g), eager), (prec, rep)
            tmp = d_f * 2.0
            del tmp
        # a second eager execute gives the same bits
        out2 = zeros(n, 3)
        plan.execute(d_x, d_f, d_x, out2)
        assert np.array_equal(host(out2), eager)


def test_sort_refuses_unsupported_work_lists():
    say("sort_points refusals")
    # two source ranges in one group
    with pytest.raises(S.Unsupported):
        S.Plan(6, 0, 100, 50, [0], [50], None, [0, 2], [0, 50], [50, 100], SORT)
    # overlapping target ranges
    with pytest.raises(S.Unsupported):
        S.Plan(6, 0, 100, 50, [0, 10], [20, 30], None, [0, 1, 2], [0, 0], [100, 100], SORT)
    # source ranges that overlap without being identical
    with pytest.raises(S.Unsupported):
        S.Plan(6, 0, 100, 50, [0, 20], [20, 50], None, [0, 1, 2], [0, 40], [60, 100], SORT)
    with pytest.raises(S.ArgumentError):
        S.Plan(6, 0, 100, 50, [0], [50], None, [0, 1], [0], [100], dict(sort_points=2))


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("K,p,uk", [(1, 1, 2), (128, 2, 8), (16, 1, 11)])
def test_direct_entry_orders_large_sums(K, p, uk, prec):
    """stkfmm_p2p_eval_device orders the points of sums of at least 16,384 sources and targets:
    20,000 x 24,000 against the oracle on sampled targets, accumulating into a prefilled output."""
    ns, nt = 20_000, 24_000
    say(f"direct entry ordered K={K} p2p={p} prec={prec}")
    g = np.random.default_rng(60 + uk)
    src = g.random((ns, 3))
    trg = np.concatenate([src[:500], g.random((nt - 500, 3))])
    v = source_values(uk, ns, g)
    prefill = g.normal(size=(nt, O.trg_dim(uk)))
    out = to_dev(prefill)
    S.p2p_device(K, p, to_dev(src), to_dev(v), to_dev(trg), out, precision=prec)
    got = host(out) - prefill
    rows = np.concatenate([np.arange(0, 500, 7), g.choice(nt, 400, replace=False)])
    l2, mx = O.max_error(got[rows], O.direct(uk, src, v, trg[rows]))
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.slow
def test_workload5_ordered_graph_replay_full_size():
    """All 512 instances (4.7M points), ordered plans in both precisions: graph replays (ordering
    included) bitwise equal to an eager execute, around unrelated eager work, and the result against
    the oracle on the largest, the smallest and two other instances."""
    import torch
    say("workload 5 full size, ordered, graph replay")
    sizes, offsets, coord, f = W.stokeslet_instances()
    d_c, d_f = to_dev(coord), to_dev(f)
    check = (0, 7, int(np.argmax(sizes)), int(np.argmin(sizes)))
    for prec in (1, 0):
        plan = instance_plan(sizes, offsets, prec, tune=SORT)
        out = zeros(len(coord), 3)
        plan.execute(d_c, d_f, d_c, out)
        eager = host(out)
        tol = FP64_TOL if prec == 0 else MIXED_TOL
        for i in check:
            a, b = offsets[i], offsets[i + 1]
            l2, mx = O.max_error(eager[a:b], O.direct(6, coord[a:b], f[a:b], coord[a:b]))
            assert l2 < tol and mx < tol, (prec, i, l2, mx)
        stream = torch.cuda.Stream()
        out_g = zeros(len(coord), 3)
        torch.cuda.synchronize()
        plan.capture_graph(d_c, d_f, d_c, out_g, stream=stream)
        for rep in range(3):
            out_g.zero_()
            torch.cuda.synchronize()
            plan.launch_graph(stream=stream)
            stream.synchronize()
            assert np.array_equal(host(out_g), eager), (prec, rep)
            tmp = d_f * 3.0
            del tmp
        del plan


@pytest.mark.parametrize("bc", [0, 1], ids=["lds", "scalar"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [13, 6, 11, 0])
def test_sorted_nan_coordinates_keep_the_mask(uk, prec, bc):
    """A NaN target coordinate, then a NaN source coordinate, among ordered points: the ordering still
    places every point (NaN lands in a boundary cell), the NaN pattern of the result equals the
    oracle's and every other entry is within the precision's bar of it.  (A NaN source turns every
    target NaN for kernels with r_i factors, as in the reference.)  Ordered target tiles are compact
    boxes, so a point with one NaN coordinate is often separated from a tile on another axis; the
    Laplace potential (uker 0) is where the unmasked form would show it, as NaN where the reference's
    mask gives 0."""
    say(f"sorted NaN coordinates uker={uk} prec={prec} broadcast={bc}")
    rng = np.random.default_rng(950 + uk)
    ns, nt = 3000, 2000
    src = rng.random((ns, 3))
    trg = np.concatenate([src[:5], rng.random((nt - 5, 3))])
    v = source_values(uk, ns, rng)
    trg[11, 1] = np.nan
    src2 = src.copy()
    src2[7, 0] = np.nan
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    for sc in (src, src2):
        ref = O.direct(uk, sc, v, trg)
        plan = S.Plan(uk, prec, ns, nt, [0], [nt], None, [0, 1], [0], [ns], dict(SORT, broadcast=bc))
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
