"""GPU parity: batched work-list plans (U-list of an adaptive octree, many independent direct sums,
HIP graph replay) against the oracle's range-list driver."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from stkfmm_amd import workloads as W
from tests.gpu_util import FP64_TOL, MIXED_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu


def sorted_problem(x, cap, uk, seed):
    tr = S.Tree(x, cap)
    xs = x[tr.perm]
    rng = np.random.default_rng(seed)
    if uk == S.UKER.STOKES_PVEL_FUSED:
        sl = rng.normal(size=(len(x), 4))
        dl = rng.normal(size=(len(x), 9))
        return tr, xs, sl, dl
    ks = O.src_dim(uk)
    v = rng.normal(size=(len(x), ks))
    if uk in (S.UKER.RPY_ULAPU, S.UKER.STOKES_REGVEL):
        v[:, 3] = rng.uniform(1e-4, 5e-4, len(x))
    return tr, xs, v, None


def run_plan(tr, uk, prec, xs, v, dl=None, tune=None, perm=None, out_rows=None):
    plan = S.Plan.from_tree(tr, uk, prec, tune=tune)
    kt = O.trg_dim(uk)
    out = zeros(out_rows or tr.n, kt)
    d_x = to_dev(xs)
    plan.execute(d_x, to_dev(v), d_x, out, src_value2=None if dl is None else to_dev(dl),
                 perm=None if perm is None else to_dev(perm))
    return host(out), plan


def oracle_leaves(tr, uk, xs, v, dl, sel):
    vals = v if dl is None else np.concatenate([v, dl], axis=1)
    return O.ranges(uk, xs, vals, xs, tr.leaf_begin, tr.leaf_end, tr.ulist_ptr, tr.rng_begin, tr.rng_end, sel=sel)


def leaf_rows(tr, sel):
    return np.concatenate([np.arange(tr.leaf_begin[l], tr.leaf_end[l]) for l in sel])


@pytest.mark.parametrize("bc", [0, 1], ids=["lds", "scalar"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("uk", [13, 2, 11, 6, 9, 7, 8, 4])
def test_small_tree_all_leaves(uk, prec, bc):
    say(f"small tree uker={uk} prec={prec} broadcast={bc}")
    g = np.random.default_rng(uk)
    x = W.wrap(np.concatenate([g.random((6000, 3)), 0.6 + 0.02 * g.normal(size=(6000, 3))]))
    tr, xs, v, dl = sorted_problem(x, 150, uk, 10 + uk)
    out, plan = run_plan(tr, uk, prec, xs, v, dl, tune=dict(broadcast=bc))
    assert plan.stats()["pairs"] == tr.pairs
    ref = oracle_leaves(tr, uk, xs, v, dl, None)
    l2, mx = O.max_error(out, ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)


@pytest.mark.parametrize("tune", [dict(targets_per_lane=1), dict(targets_per_lane=2, persistent=2),
                                  dict(targets_per_lane=4, max_src_per_item=300),
                                  dict(targets_per_lane=2, max_src_per_item=64, waves_per_cu=1),
                                  dict(targets_per_lane=2, split_tiles=-1), dict(targets_per_lane=4, split_tiles=0),
                                  dict(targets_per_lane=1, broadcast=1), dict(targets_per_lane=2, broadcast=1),
                                  dict(targets_per_lane=4, broadcast=1, persistent=2),
                                  dict(targets_per_lane=2, broadcast=1, persistent=3),
                                  dict(queues=1), dict(queues=3, max_src_per_item=256),
                                  dict(queues=64, waves_per_cu=1), dict(queues=8, broadcast=1)])
def test_tuning_knobs_do_not_change_results(tune):
    say(f"tune {tune}")
    g = np.random.default_rng(7)
    x = g.random((9000, 3))
    tr, xs, sl, dl = sorted_problem(x, 200, 13, 77)
    ref = oracle_leaves(tr, 13, xs, sl, dl, None)
    out, _ = run_plan(tr, 13, 0, xs, sl, dl, tune=tune)
    l2, mx = O.max_error(out, ref)
    assert l2 < FP64_TOL and mx < FP64_TOL


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_queue_count_gives_bitwise_identical_results(prec):
    # every item writes its own partial slot and the reduce sums slots in a fixed order, so the
    # item order (one queue or one queue per XCD, with stealing) cannot change a bit
    say(f"queues bitwise prec={prec}")
    g = np.random.default_rng(8)
    x = W.wrap(np.concatenate([g.random((8000, 3)), 0.3 + 0.01 * g.normal(size=(8000, 3))]))
    tr, xs, sl, dl = sorted_problem(x, 150, 13, 78)
    outs = [run_plan(tr, 13, prec, xs, sl, dl, tune=dict(queues=q, max_src_per_item=256))[0] for q in (1, 8, 5)]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_permuted_output_and_accumulation():
    say("perm output")
    g = np.random.default_rng(8)
    x = g.random((5000, 3))
    tr, xs, sl, dl = sorted_problem(x, 100, 13, 88)
    plan = S.Plan.from_tree(tr, 13, 0)
    out = to_dev(np.ones((tr.n, 4)))
    d_x = to_dev(xs)
    plan.execute(d_x, to_dev(sl), d_x, out, src_value2=to_dev(dl), perm=to_dev(tr.perm))
    got = host(out) - 1.0
    ref = oracle_leaves(tr, 13, xs, sl, dl, None)
    unsorted = np.zeros_like(ref)
    unsorted[tr.perm] = ref
    l2, mx = O.max_error(got, unsorted)
    assert l2 < FP64_TOL and mx < FP64_TOL


def test_soa_layouts():
    say("SoA layouts")
    g = np.random.default_rng(9)
    x = g.random((4000, 3))
    tr, xs, v, _ = sorted_problem(x, 100, 11, 99)
    plan = S.Plan.from_tree(tr, 11, 0)
    xs_soa = to_dev(xs.T.copy())          # [3, n]
    v_soa = to_dev(v.T.copy())            # [4, n]
    out_soa = zeros(6, tr.n)
    plan.execute(xs_soa.t(), v_soa.t(), xs_soa.t(), out_soa.t())
    ref = oracle_leaves(tr, 11, xs, v, None, None)
    l2, mx = O.max_error(host(out_soa).T, ref)
    assert l2 < FP64_TOL and mx < FP64_TOL


@pytest.mark.slow
def test_workload2_pvel_ulist_full_size():
    """1M points, leaf cap 1000: sampled leaves vs the oracle (FP64 and mixed); FP64 linearity."""
    say("workload 2 tree")
    x = W.pvel_points()
    tr = S.Tree(x, 1000)
    xs = x[tr.perm]
    sl, dl = W.pvel_values(tr.n, trd=True)
    d_x, d_sl, d_dl = to_dev(xs), to_dev(sl), to_dev(dl)
    outs = {}
    for prec in (0, 1):
        say(f"workload 2 prec={prec}")
        plan = S.Plan.from_tree(tr, 13, prec)
        out = zeros(tr.n, 4)
        plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl)
        outs[prec] = host(out)
    g = np.random.default_rng(2024)
    sz = tr.leaf_end - tr.leaf_begin
    sel = np.unique(np.concatenate([g.choice(tr.n_leaves, 12, replace=False), [np.argmax(sz), np.argmin(sz)]]))
    ref = oracle_leaves(tr, 13, xs, sl, dl, sel)
    rows = leaf_rows(tr, sel)
    l2, mx = O.max_error(outs[0][rows], ref[rows])
    assert l2 < FP64_TOL and mx < FP64_TOL, (l2, mx)
    l2, mx = O.max_error(outs[1][rows], ref[rows])
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)
    # whole-run mixed vs FP64 agreement
    l2, mx = O.max_error(outs[1], outs[0])
    assert l2 < MIXED_TOL, (l2, mx)
    # every target of the 1M run against the oracle's vectorised FP64 U-list (oracle_ranges_simd,
    # pinned against the scalar oracle_ranges by tests/test_oracle_drivers.py; ~1e10 pairs, a few
    # seconds on the host's threads)
    say("workload 2 whole-run oracle")
    ref_all = O.ranges(13, xs, np.concatenate([sl, dl], axis=1), xs, tr.leaf_begin, tr.leaf_end, tr.ulist_ptr,
                       tr.rng_begin, tr.rng_end, simd=True)
    l2, mx = O.max_error(outs[0], ref_all)
    assert l2 < FP64_TOL and mx < FP64_TOL, (l2, mx)
    l2, mx = O.max_error(outs[1], ref_all)
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)
    # the shipped tunings the bench runs (stkfmm_amd/tuned_defaults.json: FP64 scalar-broadcast
    # records, T = 2, 2048 sources per item; mixed LDS staging, far-tile form): the same leaves
    # against the oracle, and the whole run against the default plan's
    for prec, tol in ((0, FP64_TOL), (1, MIXED_TOL)):
        say(f"workload 2 prec={prec} shipped tuning")
        plan = S.Plan.from_tree(tr, 13, prec, tune="auto")
        out = zeros(tr.n, 4)
        plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl)
        got = host(out)
        l2, mx = O.max_error(got[rows], ref[rows])
        assert l2 < tol and mx < tol, (prec, l2, mx)
        l2, mx = O.max_error(got, outs[prec])
        assert l2 < tol, (prec, l2, mx)
    # linearity: doubling every strength doubles the FP64 result exactly
    plan = S.Plan.from_tree(tr, 13, 0)
    out2 = zeros(tr.n, 4)
    plan.execute(d_x, 2.0 * d_sl, d_x, out2, src_value2=2.0 * d_dl)
    assert np.array_equal(host(out2), 2.0 * outs[0])


@pytest.mark.slow
def test_workload3_rpy_full_size():
    """4M clustered points, leaf cap 2000, RPY: sampled leaves vs the oracle, mixed and FP64."""
    say("workload 3 tree")
    x = W.rpy_points()
    tr = S.Tree(x, 2000)
    xs = x[tr.perm]
    v = W.rpy_values(tr.n)
    d_x, d_v = to_dev(xs), to_dev(v)
    outs = {}
    for prec in (1, 0):
        say(f"workload 3 prec={prec} shipped tuning")
        plan = S.Plan.from_tree(tr, 11, prec, tune="auto")  # the tunings the bench runs
        out = zeros(tr.n, 6)
        plan.execute(d_x, d_v, d_x, out)
        outs[prec] = host(out)
    g = np.random.default_rng(3)
    sel = np.unique(g.choice(tr.n_leaves, 8, replace=False))
    ref = oracle_leaves(tr, 11, xs, v, None, sel)
    rows = leaf_rows(tr, sel)
    l2, mx = O.max_error(outs[0][rows], ref[rows])
    assert l2 < FP64_TOL and mx < FP64_TOL, (l2, mx)
    l2, mx = O.max_error(outs[1][rows], ref[rows])
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)


@pytest.mark.slow
def test_workload2_pvel_ulist_4m():
    """The PVel U-list distribution at 4M points (the batched-vs-per-leaf target size): sampled leaves
    vs the oracle, mixed and FP64."""
    say("workload 2 at 4M tree")
    x = W.pvel_points(4_000_000)
    tr = S.Tree(x, 1000)
    xs = x[tr.perm]
    sl, dl = W.pvel_values(tr.n)
    d_x, d_sl, d_dl = to_dev(xs), to_dev(sl), to_dev(dl)
    outs = {}
    for prec in (1, 0):
        say(f"workload 2 at 4M prec={prec} shipped tuning")
        plan = S.Plan.from_tree(tr, 13, prec, tune="auto")  # the tunings the bench runs
        out = zeros(tr.n, 4)
        plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl)
        outs[prec] = host(out)
    g = np.random.default_rng(4)
    sz = tr.leaf_end - tr.leaf_begin
    sel = np.unique(np.concatenate([g.choice(tr.n_leaves, 10, replace=False), [np.argmax(sz)]]))
    ref = oracle_leaves(tr, 13, xs, sl, dl, sel)
    rows = leaf_rows(tr, sel)
    l2, mx = O.max_error(outs[0][rows], ref[rows])
    assert l2 < FP64_TOL and mx < FP64_TOL, (l2, mx)
    l2, mx = O.max_error(outs[1][rows], ref[rows])
    assert l2 < MIXED_TOL and mx < MIXED_TOL, (l2, mx)


def instance_plan(sizes, offsets, prec, uk=6, tune=None):
    n = int(offsets[-1])
    gb, ge = offsets[:-1], offsets[1:]
    gp = np.arange(len(sizes) + 1)
    return S.Plan(uk, prec, n, n, gb, ge, None, gp, gb, ge, tune)


@pytest.mark.parametrize("bc", [0, 1], ids=["lds", "scalar"])
def test_instances_and_graph_replay(bc):
    """Many independent Stokeslet direct sums in one launch; graph replays equal the eager result
    (with the scalar-broadcast staging the graph also holds the record-staging kernel)."""
    import torch
    say(f"instances + graph broadcast={bc}")
    sizes, offsets, coord, f = W.stokeslet_instances(n_inst=24, n_lo=100, n_hi=2000)
    plan = instance_plan(sizes, offsets, 0, tune=dict(broadcast=bc))
    d_c, d_f = to_dev(coord), to_dev(f)
    out = zeros(len(coord), 3)
    plan.execute(d_c, d_f, d_c, out)
    eager = host(out)
    for i in range(len(sizes)):
        a, b = offsets[i], offsets[i + 1]
        ref = O.direct(6, coord[a:b], f[a:b], coord[a:b])
        l2, mx = O.max_error(eager[a:b], ref)
        assert l2 < FP64_TOL and mx < FP64_TOL
    stream = torch.cuda.Stream()
    out_g = zeros(len(coord), 3)
    torch.cuda.synchronize()
    plan.capture_graph(d_c, d_f, d_c, out_g, stream=stream)
    for rep in range(4):
        out_g.zero_()
        torch.cuda.synchronize()
        plan.launch_graph(stream=stream)
        stream.synchronize()
        assert np.array_equal(host(out_g), eager), rep
        tmp = d_f * 1.5  # unrelated eager work between replays
        del tmp


@pytest.mark.slow
def test_workload5_instances_full_size():
    say("workload 5 full size")
    sizes, offsets, coord, f = W.stokeslet_instances()
    d_c, d_f = to_dev(coord), to_dev(f)
    outs = {}
    for prec in (0, 1):
        plan = instance_plan(sizes, offsets, prec)
        out = zeros(len(coord), 3)
        plan.execute(d_c, d_f, d_c, out)
        outs[prec] = host(out)
    for i in (0, 7, int(np.argmax(sizes)), int(np.argmin(sizes))):
        a, b = offsets[i], offsets[i + 1]
        ref = O.direct(6, coord[a:b], f[a:b], coord[a:b])
        l2, mx = O.max_error(outs[0][a:b], ref)
        assert l2 < FP64_TOL and mx < FP64_TOL
        l2, mx = O.max_error(outs[1][a:b], ref)
        assert l2 < MIXED_TOL and mx < MIXED_TOL


def test_autotuner_chooses_a_candidate_and_caches(tmp_path, monkeypatch):
    from stkfmm_amd.autotune import autotune, lookup, signature
    say("autotune")
    monkeypatch.setenv("STKFMM_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    g = np.random.default_rng(31)
    x = g.random((8000, 3))
    tr, xs, sl, dl = sorted_problem(x, 150, 13, 32)
    d_x, d_sl, d_dl = to_dev(xs), to_dev(sl), to_dev(dl)
    scratch = zeros(tr.n, 4)
    cands = [dict(targets_per_lane=1, max_src_per_item=1024, persistent=1),
             dict(targets_per_lane=2, max_src_per_item=2048, persistent=2)]
    run = dict(src_coord=d_x, src_value=d_sl, trg_coord=d_x, trg_value=scratch, src_value2=d_dl)
    tune, timings = autotune(tr, 13, 1, run, candidates=cands, reps=2)
    assert tune in cands and len(timings) == 2
    assert lookup(signature(13, 1, tr.n, tr.pairs, tr.n_leaves)) == tune
    # a plan built with tune="auto" finds the cached choice and gives oracle-parity results
    plan = S.Plan.from_tree(tr, 13, 1, tune="auto")
    out = zeros(tr.n, 4)
    plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl)
    l2, mx = O.max_error(host(out), oracle_leaves(tr, 13, xs, sl, dl, None))
    assert l2 < MIXED_TOL and mx < MIXED_TOL


def piece_tiles(n, tt, pieces=True):
    """Target tiles of a group of n targets as the planner lays them out: whole tiles of tt, then the
    remainder -- cut into power-of-two tiles, largest first, in plans of several groups -- until the
    rest fills its lane slots (tt >> lane split, at least tt/16) to within 1/8 or a cut saves none."""
    def slots(c):
        ls = 0
        while ls < 4 and (c << (ls + 1)) <= tt:
            ls += 1
        return tt >> ls
    out = [tt] * (n // tt)
    n %= tt
    while n > 0:
        q = 1 << (n.bit_length() - 1)
        if not pieces or 8 * (slots(n) - n) <= slots(n) or q == n or q + slots(n - q) >= slots(n):
            out.append(n)
            break
        out.append(q)
        n -= q
    return out


@pytest.mark.parametrize("split", [0, -1], ids=["pieces", "one-lane"])
@pytest.mark.parametrize("T", [1, 2, 4])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_remainder_tiles(prec, T, split):
    """Groups whose sizes leave every kind of remainder (1, just above a power of two, just below a
    whole tile, several whole tiles): the tile count the planner reports, and every target against
    the oracle."""
    say(f"remainder tiles prec={prec} T={T} split={split}")
    sizes = [1, 9, 17, 40, 65, 100, 127, 128, 129, 193, 255, 300, 333, 513]
    g = np.random.default_rng(100 + T)
    uk = 7
    n_trg, n_src = sum(sizes), 5000
    trg = g.random((n_trg, 3))
    src = g.random((n_src, 3))
    v = g.normal(size=(n_src, O.src_dim(uk)))
    gb = np.cumsum([0] + sizes[:-1]).astype(np.int32)
    ge = (gb + np.array(sizes)).astype(np.int32)
    rb, re_, gp = [], [], [0]
    for _ in sizes:
        for _ in range(2):
            a = int(g.integers(0, n_src - 600))
            rb.append(a)
            re_.append(a + int(g.integers(1, 600)))
        gp.append(len(rb))
    work = dict(n_src=n_src, n_trg=n_trg, grp_trg_begin=gb, grp_trg_end=ge, grp_level=None,
                grp_rng_ptr=np.array(gp, np.int32), rng_begin=np.array(rb, np.int32),
                rng_end=np.array(re_, np.int32))
    plan = S.Plan(uk, prec, tune=dict(targets_per_lane=T, split_tiles=split), **work)
    expect = sum(len(piece_tiles(n, 64 * T, split == 0)) for n in sizes)
    assert plan.stats()["n_tiles"] == expect
    out = zeros(n_trg, O.trg_dim(uk))
    plan.execute(to_dev(src), to_dev(v), to_dev(trg), out)
    ref = O.ranges(uk, src, v, trg, gb, ge, gp, rb, re_)
    l2, mx = O.max_error(host(out), ref)
    tol = FP64_TOL if prec == 0 else MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)
