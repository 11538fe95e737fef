"""Benchmarks of the STKFMM P2P path on one MI355X.

Default (the headline metric): Stokes PVel P2P over the U-list of an adaptive octree --
1,000,000 points, half uniform in [0,1)^3 and half in 16 Gaussian clusters (sigma 0.01), leaf cap
1000; single layer f ~ N(0,1)^3 (TrD = 0) and double layer 9 x N(0,1) at every point (SL = DL =
targets; one pair covers both layers).  One step = one full U-list evaluation (every leaf, one
batched persistent launch + the partial-sum reduction) with inputs resident in HBM.  value = Gpair/s
of the mixed-precision batched launch; beside it the FP64 batched launch, the plain FP64
per-leaf-launch version (one launch per leaf), accuracy of each against the FP64 oracle, the
roofline of the dominant kernel (HIP events on its stream) and the CPU baseline (the oracle's
vectorised C restatement of the reference P2P on the host cores).

Other workloads (--workload), each printing its own JSON line:
  pvel-ulist-4m         the default workload's distribution at 4,000,000 points (mixed / FP64 / per-leaf)
  lapgrad-direct        8192 x 8192 Laplace PGrad direct sum (STKFMM::evaluateKernel), FP64
  rpy-ulist             4M log-normal clustered points, leaf cap 2000, RPY, mixed (autotuned) and FP64
  pvel-adjoint          65,536 x 65,536 Stokes PVel forward + adjoint (gradients of sum_t w_t.u_t)
  stokeslet-instances   512 independent Stokeslet direct sums in one HIP-graph replay
  pvel-fmm-lists        the rest of an FMM evaluation's P2P-shaped work on the 1M PVel tree: S2M, L2T,
                        W-list (M2T) and X-list (S2L) surface evaluations, order-8 surfaces
  wall-images           StkWallFMM's no-slip wall image system (Stokes and RPY), 65,536 sources and
                        targets above the wall, by direct sums, FP64
  family-ulist          the rest of the kernel family (PVelGrad, PVelLaplacian, Traction single and
                        double layer, RegFTVelOmega and its sub-kernels, LapQPGradGrad and QuadP) over
                        the U-list of a 262,144-point tree of the default distribution, mixed and FP64

  python bench.py [--workload W] [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X vector FP64, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X vector FP32 (published, packed), FLOP/s
# algorithmic FLOPs per pair and the FP64 share in mixed mode (displacement + accumulation)
FLOPS = {"pvel_fused": (72, 7), "pvel_sl": (35, 7), "rpy": (56, 9), "stokeslet": (29, 6), "lapgrad": (20, 7),
         # adjoint of stokes_pvel per pair: displacement 3, r^2 5, rsqrt 1, powers of 1/r 3, f.r 5, a - T 1,
         # w.r 5, w.f 5, the two coefficients 3 + 8, (a - T)/r^3 1, position gradient 15, value gradient
         # 12 + 2, target and source position sums 3 + 3; mixed keeps only the displacement in FP64
         "pvel_vjp": (75, 3)}
NOTE = ("VALU-bound; peak = MI355X published vector FP64 78.6 / FP32 157.3 TFLOP/s combined by the "
        "FP64/FP32 split of the algorithmic FLOPs; the library is compiled with packed FP32 disabled")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="pvel-ulist",
                    choices=["pvel-ulist", "pvel-ulist-4m", "lapgrad-direct", "rpy-ulist", "pvel-adjoint",
                             "stokeslet-instances", "pvel-fmm-lists", "wall-images", "family-ulist"])
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU work budget of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-leaf", action="store_true")
    ap.add_argument("--no-target-4m", action="store_true",
                    help="default workload: skip the compact 4M-point batched-versus-per-leaf measurement")
    ap.add_argument("--only", choices=["mixed", "fp64", "per-leaf", "mixed-sorted", "fp64-sorted"], default=None,
                    help="run one variant only (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy of known size (PMC calibration) and exit")
    return ap.parse_args()


def peak_for(kind, prec):
    f, f64 = FLOPS[kind]
    if prec == "fp64":
        return PEAK_FP64
    return f / (f64 / PEAK_FP64 + (f - f64) / PEAK_FP32)


def roofline(kind, prec, pairs, kernel_ms, kernel_name):
    f = FLOPS[kind][0]
    achieved = pairs * f / (kernel_ms * 1e-3)
    peak = peak_for(kind, prec)
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None, "kernel": kernel_name, "kernel_ms": kernel_ms,
            "flops_per_pair": f, "pairs_per_launch": pairs, "note": NOTE}


def attach_traffic(rl, workload, variant):
    """roofline.traffic from the committed counter summary of this workload's dominant kernel
    (profiles/<workload>_<variant>_traffic.json: FETCH_SIZE and WRITE_SIZE scaled by the copy
    calibration), or None when there is none."""
    tfile = os.path.join(ROOT, "profiles", f"{workload}_{variant}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        rl["traffic"] = tr.get("hbm_bytes_per_launch")
        rl["traffic_source"] = os.path.relpath(tfile, ROOT)
    return rl


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def threads():
    """Host threads of the CPU baseline: OMP_NUM_THREADS when the environment sets it (the host's
    CPU share allotted to this GPU), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, len(os.sched_getaffinity(0)))


def host_info():
    """What the CPU baseline ran on: model, logical CPUs of the host, CPUs this process may use,
    and where the thread count came from."""
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "threads_from": "OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS", "").isdigit() else "affinity"}


def line(metric, value, unit, args, ms_per_step, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


class Timed:
    """Run a plan for warmup + steps, with wall clock and the library's HIP-event kernel timing."""

    def __init__(self, torch, plan, run_kwargs, stream):
        self.torch, self.plan, self.kw, self.stream = torch, plan, run_kwargs, stream

    def __call__(self, steps, warmup):
        torch = self.torch
        with torch.cuda.stream(self.stream):
            for _ in range(warmup):
                self.plan.execute(stream=self.stream, **self.kw)
            torch.cuda.synchronize()
            self.plan.set_timing(True)
            t = time.perf_counter()
            for _ in range(steps):
                self.plan.execute(stream=self.stream, **self.kw)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t
            item_ms, reduce_ms, count = self.plan.kernel_ms()
            self.plan.set_timing(False)
        return wall, item_ms, reduce_ms, count


# ---- U-list workloads (headline: PVel) ---------------------------------------------------------------

# micro-kernel -> the reference's Test/SimpleKernel.cpp functions that make it up, as (KERNEL value,
# layer, first and last source-value column); the fused PVel is the single plus the double layer
REF_PARTS = {13: ((128, 0, 0, 4), (128, 1, 4, 13)), 11: ((16, 0, 0, 4),)}


def reference_sample(O, uker, tree, xs, allv, kt, leaf_order, budget_s, nth):
    """The reference's own scalar P2P (Test/SimpleKernel.cpp, built into oracle/_ref, OpenMP over
    targets as in Test/Test.cpp:478) on whole leaves' U-lists, in leaf_order until budget_s seconds
    of compute are spent.  Returns (timing dict, target rows, reference values of those rows)."""
    if int(uker) not in REF_PARTS or not O.ref_available():
        return None, None, None
    used = O.simplekernel_threads(nth)
    done, rows, outs, pairs, secs = [], [], [], 0, 0.0
    for leaf in leaf_order:
        b, e = int(tree.leaf_begin[leaf]), int(tree.leaf_end[leaf])
        r0, r1 = int(tree.ulist_ptr[leaf]), int(tree.ulist_ptr[leaf + 1])
        src = np.concatenate([np.arange(tree.rng_begin[r], tree.rng_end[r]) for r in range(r0, r1)])
        sc, tc = np.ascontiguousarray(xs[src]), np.ascontiguousarray(xs[b:e])
        parts = [(K, layer, np.ascontiguousarray(allv[src, c0:c1]), c1 - c0)
                 for K, layer, c0, c1 in REF_PARTS[int(uker)]]
        out = np.zeros((e - b, kt))
        t = time.perf_counter()
        for K, layer, sv, ks in parts:
            out += O.simplekernel_direct(K, layer, ks, kt, sc, sv, tc)
        secs += time.perf_counter() - t
        pairs += (e - b) * src.size
        done.append(leaf)
        rows.append(np.arange(b, e))
        outs.append(out)
        if secs >= budget_s:
            break
    info = {"value": pairs / secs / 1e9, "unit": "Gpair/s", "cores": used, "kind": "reference",
            "sample": (f"{len(done)} of {tree.n_leaves} leaves (whole U-lists), {sum(r.size for r in rows)} targets, "
                       f"{pairs:.3e} pairs, {secs:.1f} s; the reference's scalar Test/SimpleKernel.cpp functions "
                       f"compiled from its sources (g++ -O2), OpenMP over targets as its runSimpleKernel loop")}
    return info, np.concatenate(rows), np.concatenate(outs)


def target_4m(args, torch, S, W):
    """The batched-versus-per-leaf target at 4M points (the default workload's distribution at
    4,000,000 points, leaf cap 1000), compact enough to ride along the default run: mixed and FP64
    batched launches with the shipped tuning, the plain FP64 per-leaf launch, and each one's error
    against the oracle on a random sample of whole leaves."""
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    x = W.pvel_points(4_000_000)
    tree = S.Tree(x, 1000)
    xs = x[tree.perm]
    sl, dl = W.pvel_values(tree.n)
    setup_s = time.perf_counter() - t0
    d_x, d_sl, d_dl = (torch.from_numpy(a).to(dev) for a in (xs, sl, dl))
    out = torch.zeros((tree.n, 4), dtype=torch.float64, device=dev)
    kw = dict(src_coord=d_x, src_value=d_sl, trg_coord=d_x, trg_value=out, src_value2=d_dl)
    stream = torch.cuda.Stream()
    uk = S.UKER.STOKES_PVEL_FUSED
    res, outs = {}, {}
    for name, prec, tune, steps in (("mixed", S.PREC.MIXED, "auto", 5), ("fp64", S.PREC.FP64, "auto", 3),
                                    ("per-leaf", S.PREC.FP64, dict(persistent=3), 1)):
        plan = S.Plan.from_tree(tree, uk, prec, tune=tune)
        wall, item_ms, _, _ = Timed(torch, plan, kw, stream)(steps, 1)
        r = torch.zeros_like(out)
        plan.execute(stream=stream, **dict(kw, trg_value=r))
        torch.cuda.synchronize()
        outs[name] = r.cpu().numpy()
        res[name] = {"gpair_s": tree.pairs * steps / wall / 1e9, "ms_per_step": 1e3 * wall / steps,
                     "item_ms": item_ms}
        del plan, r
        print(f"# target 4M {name}: {res[name]['gpair_s']:.1f} Gpair/s", file=sys.stderr, flush=True)
    acc = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        sel = np.random.default_rng(4).choice(tree.n_leaves, 48, replace=False)
        ref = O.ranges(uk, xs, np.concatenate([sl, dl], axis=1), xs, tree.leaf_begin, tree.leaf_end, tree.ulist_ptr,
                       tree.rng_begin, tree.rng_end, sel=sel, nthreads=threads(), simd=True)
        rows = np.concatenate([np.arange(tree.leaf_begin[l], tree.leaf_end[l]) for l in sel])
        acc = {}
        for name, o in outs.items():
            l2, mx = O.max_error(o[rows], ref[rows])
            acc[name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": int(rows.size), "leaves_checked": 48}
    return {"workload": "stokes-pvel-ulist-4M-clustered-leafcap1000", "points": tree.n, "leaves": tree.n_leaves,
            "pairs": tree.pairs, "setup_s": setup_s, "variants": res, "accuracy_vs_oracle": acc,
            "speedup_mixed_vs_fp64_per_leaf_launch": res["mixed"]["gpair_s"] / res["per-leaf"]["gpair_s"],
            "target": ">= 10x the FP64 per-leaf launch at <= 1e-6 relative error"}


def ulist_bench(args, torch, S, kind, uker, points, cap, values_fn, kt, metric, workload_name, variants_order,
                extra_fields=None):
    from stkfmm_amd.autotune import autotune

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    x = points()
    tree = S.Tree(x, cap)
    xs = x[tree.perm]
    vals = values_fn(tree.n)
    setup_s = time.perf_counter() - t0
    d_x = torch.from_numpy(xs).to(dev)
    d_v = [torch.from_numpy(v).to(dev) for v in vals]
    out = torch.zeros((tree.n, kt), dtype=torch.float64, device=dev)
    pairs = tree.pairs
    stream = torch.cuda.Stream()
    kw = dict(src_coord=d_x, src_value=d_v[0], trg_coord=d_x, trg_value=out)
    if len(d_v) > 1:
        kw["src_value2"] = d_v[1]

    variants, results = {}, {}
    order = variants_order if args.only is None else [args.only]
    for name in order:
        if name == "per-leaf" and args.no_per_leaf:
            continue
        prec = S.PREC.MIXED if name == "mixed" else S.PREC.FP64
        if name == "per-leaf":
            tune = dict(persistent=3)
        else:
            scratch = torch.zeros_like(out)
            tune, _ = autotune(tree, uker, prec, dict(kw, trg_value=scratch, stream=stream))
            del scratch
        plan = S.Plan.from_tree(tree, uker, prec, tune=tune)
        steps = args.steps if name != "per-leaf" else max(1, min(args.steps, 3))
        warm = args.warmup if name != "per-leaf" else 1
        wall, item_ms, reduce_ms, count = Timed(torch, plan, kw, stream)(steps, warm)
        res = torch.zeros_like(out)
        plan.execute(stream=stream, **dict(kw, trg_value=res))
        torch.cuda.synchronize()
        st = plan.stats()
        assert st["pairs"] == pairs
        variants[name] = dict(ms_per_step=1e3 * wall / steps, gpair_s=pairs * steps / wall / 1e9, item_ms=item_ms,
                              reduce_ms=reduce_ms, timed_launches=count, n_items=st["n_items"],
                              n_tiles=st["n_tiles"], grid=st["grid"], tune=tune)
        results[name] = res
        print(f"# {workload_name} {name}: {variants[name]['gpair_s']:.1f} Gpair/s, "
              f"{variants[name]['ms_per_step']:.3f} ms/step, item kernel {item_ms:.3f} ms",
              file=sys.stderr, flush=True)
    head = order[0]
    hv = variants[head]
    T = (hv["tune"] or {}).get("targets_per_lane", 2)
    prec_tag = "MIXED" if head == "mixed" else "FP64"
    # tree-derived mixed plans at T = 2 run the far-tile instantiation of the item kernel
    far = ",far-tile form" if head == "mixed" and T == 2 else ""
    rl = roofline(kind, "mixed" if head == "mixed" else "fp64", pairs, hv["item_ms"],
                  f"p2p_items_kernel<{int(uker)},{prec_tag},T={T}{far}>")
    attach_traffic(rl, workload_name, head)

    accuracy, cpu, ref_cpu = None, None, None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        nth = threads()
        sz = (tree.leaf_end - tree.leaf_begin).astype(np.int64)
        csum = np.concatenate([[0], np.cumsum(tree.rng_end.astype(np.int64) - tree.rng_begin)])
        nsrc = csum[tree.ulist_ptr[1:]] - csum[tree.ulist_ptr[:-1]]
        cost = sz * nsrc
        allv = np.concatenate(vals, axis=1) if len(vals) > 1 else vals[0]
        order_l = np.random.default_rng(12345).permutation(tree.n_leaves)
        # calibrate on ~1% of the pairs, then run the rest of the workload if it fits twice the
        # budget, or a random sample of leaves sized to the budget
        cum = np.cumsum(cost[order_l])
        cal = order_l[: int(np.searchsorted(cum, 0.01 * pairs)) + 1]
        ref = np.zeros((tree.n, kt))
        t = time.perf_counter()
        ref = O.ranges(uker, xs, allv, xs, tree.leaf_begin, tree.leaf_end, tree.ulist_ptr, tree.rng_begin,
                       tree.rng_end, sel=cal, nthreads=nth, simd=True, trg_value=ref)
        cal_s = time.perf_counter() - t
        rate = float(cost[cal].sum()) / max(cal_s, 1e-9)
        rest = order_l[len(cal):]
        if rate * args.cpu_seconds * 2 >= pairs:
            more = rest
        else:
            budget = rate * args.cpu_seconds - float(cost[cal].sum())
            more = rest[: int(np.searchsorted(np.cumsum(cost[rest]), max(budget, 0.0))) + 1]
        t = time.perf_counter()
        ref = O.ranges(uker, xs, allv, xs, tree.leaf_begin, tree.leaf_end, tree.ulist_ptr, tree.rng_begin,
                       tree.rng_end, sel=more, nthreads=nth, simd=True, trg_value=ref)
        main_s = time.perf_counter() - t
        sel = np.concatenate([cal, more])
        sample_pairs = int(cost[sel].sum())
        cpu_s = cal_s + main_s
        rows = np.concatenate([np.arange(tree.leaf_begin[l], tree.leaf_end[l]) for l in sel])
        whole = len(sel) == tree.n_leaves
        cpu = {"value": sample_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"{'all' if whole else len(sel)} of {tree.n_leaves} leaves, {rows.size} targets, "
                          f"{sample_pairs:.3e} pairs, {cpu_s:.1f} s; FP64 C restatement of the reference uKerEval "
                          "algebra, AVX2 lanes over targets, OpenMP over leaves (PVFMM SIMD P2P not buildable "
                          "offline)"),
               **host_info()}
        accuracy = {}
        for name, res in results.items():
            l2, mx = O.max_error(res.cpu().numpy()[rows], ref[rows])
            accuracy[name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": int(rows.size)}
        # the reference's own CPU code on a sample of the same leaves, timed and used as a second checker
        ref_cpu, ref_rows, ref_out = reference_sample(O, uker, tree, xs, allv, kt, sel, args.cpu_seconds / 2, nth)
        if ref_cpu is not None:
            ref_cpu.update(host_info())
            ref_acc = {}
            for name, res in results.items():
                l2, mx = O.max_error(res.cpu().numpy()[ref_rows], ref_out)
                ref_acc[name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": int(ref_rows.size)}

    extra = dict(variants=variants, accuracy_vs_oracle=accuracy, roofline=rl, cpu_baseline=cpu)
    if extra_fields:
        extra.update(extra_fields)
    if not args.no_cpu_baseline and ref_cpu is not None:
        extra.update(cpu_reference=ref_cpu, accuracy_vs_reference=ref_acc)
    if "per-leaf" in variants and head == "mixed":
        extra["speedup_vs_fp64_per_leaf_launch"] = hv["gpair_s"] / variants["per-leaf"]["gpair_s"]
    if "fp64" in variants and head == "mixed":
        extra["speedup_vs_fp64_batched"] = hv["gpair_s"] / variants["fp64"]["gpair_s"]
    config = {"workload": workload_name, "points": tree.n, "leaf_cap": cap, "leaves": tree.n_leaves,
              "max_level": tree.max_level, "pairs": pairs, "precision": head, "setup_s": setup_s}
    line(metric, hv["gpair_s"], "Gpair/s", args, hv["ms_per_step"],
         "f64+f32 (mixed)" if head == "mixed" else "f64", config, **extra)


# ---- other workloads ------------------------------------------------------------------------------

def lapgrad_direct(args, torch, S, W):
    dev = torch.device("cuda:0")
    src, q, trg = W.lapgrad_direct()
    d_src, d_q, d_trg = (torch.from_numpy(a).to(dev) for a in (src, q, trg))
    out = torch.zeros((len(trg), 4), dtype=torch.float64, device=dev)
    uker = S.get_uker(S.KERNEL.LapPGrad, S.PPKERNEL.SLS2T)
    work = dict(n_src=len(src), n_trg=len(trg), grp_trg_begin=[0], grp_trg_end=[len(trg)], grp_level=None,
                grp_rng_ptr=[0, 1], rng_begin=[0], rng_end=[len(src)])
    stream = torch.cuda.Stream()
    kw = dict(src_coord=d_src, src_value=d_q, trg_coord=d_trg, trg_value=out)
    # a single 8192 x 8192 sum is small for 256 CUs: tune the item size down as well
    from stkfmm_amd.autotune import autotune
    cands = [dict(targets_per_lane=t, max_src_per_item=m, persistent=p, broadcast=b)
             for t in (1, 2, 4) for m in (256, 320, 384, 448, 512, 1024, 2048) for p in (1, 2) for b in (0, 1)]
    scratch = torch.zeros_like(out)
    tune, _ = autotune(work, uker, S.PREC.FP64, dict(kw, trg_value=scratch, stream=stream), candidates=cands)
    del scratch
    plan = S.Plan(uker, S.PREC.FP64, tune=tune, **work)
    steps = max(args.steps, 50)
    wall, item_ms, reduce_ms, count = Timed(torch, plan, kw, stream)(steps, args.warmup)
    pairs = len(src) * len(trg)
    # the host-pointer entry (STKFMM::evaluateKernel drop-in): PCIe-inclusive
    h_out = np.zeros((len(trg), 4))
    for _ in range(3):
        S.evaluate_kernel(S.KERNEL.LapPGrad, 0, S.PPKERNEL.SLS2T, src, q, trg, h_out)
    calls = []
    for _ in range(20):
        t = time.perf_counter()
        S.evaluate_kernel(S.KERNEL.LapPGrad, 0, S.PPKERNEL.SLS2T, src, q, trg, h_out)
        calls.append(time.perf_counter() - t)
    host_s = float(np.median(calls))
    cpu = acc = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        nth = threads()
        t = time.perf_counter()
        ref = O.direct(2, src, q, trg, nthreads=nth, chunked=True)
        chunked_s = time.perf_counter() - t
        t = time.perf_counter()
        O.direct(2, src, q, trg, nthreads=nth)
        all_s = time.perf_counter() - t
        res = torch.zeros_like(out)
        plan.execute(**dict(kw, trg_value=res))
        torch.cuda.synchronize()
        l2, mx = O.max_error(res.cpu().numpy(), ref)
        acc = {"fp64": {"errorL2": l2, "errorMaxRel": mx}}
        cpu = {"value": pairs / chunked_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"whole workload; the reference's driver of FMMData.cpp:252-283 (4000-target chunks, "
                          f"static(1): {len(trg) // 4000 + 1} busy threads) {chunked_s:.2f} s; all-core "
                          f"target split {all_s:.2f} s = {pairs / all_s / 1e9:.3f} Gpair/s; scalar FP64 C "
                          "restatement (PVFMM SIMD P2P not buildable offline)"),
               **host_info()}
    line("Gpair-interactions/s Laplace PGrad direct sum (evaluateKernel), FP64", pairs * steps / wall / 1e9,
         "Gpair/s", args, 1e3 * wall / steps, "f64",
         {"workload": "lapgrad-direct-8192x8192", "pairs": pairs, "precision": "fp64"},
         host_api_pcie_inclusive_gpair_s=pairs / host_s / 1e9, tune=tune,
         roofline=attach_traffic(roofline("lapgrad", "fp64", pairs, item_ms,
                                          f"p2p_items_kernel<{int(uker)},FP64,T={tune.get('targets_per_lane', 2)}>"),
                                 "lapgrad-direct-8192x8192", "fp64"),
         cpu_baseline=cpu, accuracy_vs_oracle=acc, kernel_ms={"item": item_ms, "reduce": reduce_ms})


def pvel_adjoint(args, torch, S, W):
    from stkfmm_amd.autograd import p2p
    dev = torch.device("cuda:0")
    src, val, trg, w = W.adjoint_pvel()
    d_src, d_val, d_trg, d_w = (torch.from_numpy(a).to(dev) for a in (src, val, trg, w))
    pairs = len(src) * len(trg)
    uk = S.UKER.STOKES_PVEL
    out = torch.zeros((len(trg), 4), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        S.uker_device(uk, d_src, d_val, d_trg, out)
        S.vjp_device(uk, d_src, d_val, d_trg, d_w)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        S.uker_device(uk, d_src, d_val, d_trg, out)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t) / args.steps
    # HIP events on torch's current stream, the stream the library launches the adjoint on
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        S.vjp_device(uk, d_src, d_val, d_trg, d_w)
    ev1.record()
    torch.cuda.synchronize()
    adj = (time.perf_counter() - t) / args.steps
    adj_kernel_ms = ev0.elapsed_time(ev1) / args.steps
    for _ in range(args.warmup):
        S.vjp_device(uk, d_src, d_val, d_trg, d_w, precision=S.PREC.MIXED)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        S.vjp_device(uk, d_src, d_val, d_trg, d_w, precision=S.PREC.MIXED)
    ev1.record()
    torch.cuda.synchronize()
    adj_mixed = (time.perf_counter() - t) / args.steps
    adj_mixed_kernel_ms = ev0.elapsed_time(ev1) / args.steps
    a = d_src.clone().requires_grad_(True)
    b = d_val.clone().requires_grad_(True)
    c = d_trg.clone().requires_grad_(True)

    def ag_step():
        # the same work every time: fresh gradients (no accumulation into earlier ones)
        a.grad = b.grad = c.grad = None
        (p2p(uk, a, b, c) * d_w).sum().backward()

    for _ in range(2):  # warm-up (first-call allocations and kernel loads)
        ag_step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        ag_step()
    torch.cuda.synchronize()
    ag = (time.perf_counter() - t) / 3
    cpu = acc = None
    if not args.no_cpu_baseline:
        # the oracle's adjoint (FP64 C, OpenMP) on every source against the first m targets: timed as
        # the CPU baseline and used as the checker of both GPU precisions on the same sub-problem
        from oracle import oracle as O
        nth = threads()
        m = min(8192, len(trg))
        t = time.perf_counter()
        ref = O.direct_vjp(int(uk), src, val, trg[:m], w[:m], nthreads=nth)
        cpu_s = time.perf_counter() - t
        acc = {}
        for name, prec in (("fp64", S.PREC.FP64), ("mixed", S.PREC.MIXED)):
            got = S.vjp_device(uk, d_src, d_val, d_trg[:m].contiguous(), d_w[:m].contiguous(), precision=prec)
            acc[name] = {}
            for g, x, r in zip(("src_value", "src_coord", "trg_coord"), got, ref):
                l2, mx = O.max_error(x.cpu().numpy(), r)
                acc[name][g] = {"errorL2": l2, "errorMaxRel": mx}
        cpu = {"value": len(src) * m / cpu_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"{len(src)} sources x the first {m} targets ({len(src) * m:.3e} pairs, {cpu_s:.2f} s): "
                          "the oracle's FP64 adjoint (target pass + source pass), OpenMP"),
               **host_info()}
    line("Gpair-interactions/s Stokes PVel forward + adjoint (direct sum, FP64)", 2 * pairs / (fwd + adj) / 1e9,
         "Gpair/s", args, 1e3 * (fwd + adj), "f64",
         {"workload": "pvel-adjoint-65536x65536", "pairs": pairs, "precision": "fp64",
          "value_counts": "forward pairs + adjoint pairs"},
         forward_ms=1e3 * fwd, adjoint_ms=1e3 * adj, autograd_fwd_bwd_ms=1e3 * ag,
         forward_gpair_s=pairs / fwd / 1e9, adjoint_gpair_s=pairs / adj / 1e9,
         adjoint_mixed_ms=1e3 * adj_mixed, adjoint_mixed_gpair_s=pairs / adj_mixed / 1e9,
         roofline=attach_traffic(
             dict(roofline("pvel_vjp", "fp64", pairs, adj_kernel_ms, "vjp_pair_kernel<stokes_pvel,FP64,T=4>"),
                  note=NOTE + "; kernel_ms = the adjoint's kernels (pair pass + partial-sum reductions) "
                              "between HIP events on their stream"), "pvel-adjoint-65536x65536", "fp64"),
         roofline_mixed=attach_traffic(roofline("pvel_vjp", "mixed", pairs, adj_mixed_kernel_ms,
                                                "vjp_pair_kernel<stokes_pvel,MIXED,T=8>"),
                                       "pvel-adjoint-65536x65536", "mixed"),
         cpu_baseline=cpu, accuracy_vs_oracle=acc)


def pvel_fmm_lists(args, torch, S, W):
    """S2M, L2T, W-list and X-list surface evaluations of PVel (the stokes_pvel sub-kernel of each,
    StokesLayerKernel.hpp:44-50) on the 1M-point tree, order-8 surfaces (296 points): one plan per
    list, mixed and FP64, tune="auto": S2M and L2T (one signature) take the shipped tuning of
    stkfmm_amd/tuned_defaults.json, W and X the default (within 1 % of the best candidate)."""
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    x = W.pvel_points()
    tree = S.Tree(x, 1000)
    xs = x[tree.perm]
    setup_s = time.perf_counter() - t0
    K, p = S.KERNEL.PVel, 8
    uk = S.get_uker(K, S.PPKERNEL.L2T)
    assert uk == S.get_s2m_uker(K) == S.get_m2t_uker(K) == S.get_s2l_uker(K)
    ls, ns = S.LeafSurfaces(tree, p, S.RAD_OUTER), S.NodeSurfaces(tree, p)
    cnt = ls.count
    g = np.random.default_rng(2024)
    v_pts = W.pvel_values(tree.n)[0]
    dens = np.zeros((max(tree.n_leaves, tree.n_nodes) * cnt, 4))
    dens[:, :3] = g.normal(size=(dens.shape[0], 3))
    d_x, d_v, d_dens = (torch.from_numpy(a).to(dev) for a in (xs, v_pts, dens))
    h_ls, h_ns = ls.coord.cpu().numpy(), ns.coord.cpu().numpy()
    nl, nn = tree.n_leaves, tree.n_nodes
    # work lists as the surface classes build them: (sources, values, targets, n_trg, groups, ranges)
    lb = np.arange(nl, dtype=np.int64) * cnt
    xn = np.nonzero(np.diff(tree.x_ptr) > 0)[0]
    xb = xn.astype(np.int64) * cnt
    wb = tree.w_node.astype(np.int64) * cnt
    lists = {
        "S2M": (ls.to_surface_plan, (d_x, d_v, ls.coord), (xs, v_pts, h_ls), nl * cnt,
                (lb, lb + cnt, np.arange(nl + 1), tree.leaf_begin, tree.leaf_end)),
        "L2T": (ls.from_surface_plan, (ls.coord, d_dens, d_x), (h_ls, dens[:nl * cnt], xs), tree.n,
                (tree.leaf_begin, tree.leaf_end, np.arange(nl + 1), lb, lb + cnt)),
        "W": (ns.w_plan, (ns.coord, d_dens, d_x), (h_ns, dens[:nn * cnt], xs), tree.n,
              (tree.leaf_begin, tree.leaf_end, tree.w_ptr, wb, wb + cnt)),
        "X": (ns.x_plan, (d_x, d_v, ns.coord), (xs, v_pts, h_ns), nn * cnt,
              (xb, xb + cnt, np.concatenate([[0], np.cumsum(np.diff(tree.x_ptr)[xn])]),
               tree.leaf_begin[tree.x_leaf], tree.leaf_end[tree.x_leaf])),
    }
    stream = torch.cuda.Stream()
    per, results = {}, {}
    for name, (mk, dev_args, _, n_trg, _) in lists.items():
        for prec_name, prec in (("mixed", S.PREC.MIXED), ("fp64", S.PREC.FP64)):
            plan = mk(uk, prec, tune="auto")
            out = torch.zeros((n_trg, 4), dtype=torch.float64, device=dev)
            kw = dict(src_coord=dev_args[0], src_value=dev_args[1], trg_coord=dev_args[2], trg_value=out)
            wall, item_ms, reduce_ms, _ = Timed(torch, plan, kw, stream)(max(args.steps, 10), args.warmup)
            steps = max(args.steps, 10)
            pairs = plan.stats()["pairs"]
            res = torch.zeros_like(out)
            plan.execute(**dict(kw, trg_value=res))
            torch.cuda.synchronize()
            per.setdefault(name, {"pairs": pairs})[prec_name] = {
                "ms_per_step": 1e3 * wall / steps, "gpair_s": pairs * steps / wall / 1e9, "item_ms": item_ms,
                "reduce_ms": reduce_ms}
            results[(name, prec_name)] = res.cpu().numpy()
            print(f"# pvel-fmm-lists {name} {prec_name}: {pairs:.3e} pairs, {pairs * steps / wall / 1e9:.1f} Gpair/s, "
                  f"item kernel {item_ms:.3f} ms", file=sys.stderr, flush=True)
    total_pairs = sum(v["pairs"] for v in per.values())
    ms_mixed = sum(v["mixed"]["ms_per_step"] for v in per.values())
    ms_fp64 = sum(v["fp64"]["ms_per_step"] for v in per.values())
    big = max(per, key=lambda k: per[k]["pairs"])
    rl = attach_traffic(roofline("pvel_sl", "mixed", per[big]["pairs"], per[big]["mixed"]["item_ms"],
                                 f"p2p_items_kernel<{int(uk)},MIXED,T=2> ({big} list)"),
                        "pvel-fmm-lists-1M-clustered-leafcap1000-p8", f"{big}-mixed")
    cpu = acc = None
    if not args.no_cpu_baseline:
        # the oracle over a sample of each list's groups: timed as the CPU baseline and used as the checker
        from oracle import oracle as O
        nth = threads()
        acc, cpu_pairs, cpu_s = {}, 0, 0.0
        for name, (_, _, host_args, n_trg, wl) in lists.items():
            gb, ge, gp, rb, re = wl
            sizes = np.asarray(ge, dtype=np.int64) - np.asarray(gb, dtype=np.int64)
            csum = np.concatenate([[0], np.cumsum(np.asarray(re, dtype=np.int64) - np.asarray(rb, dtype=np.int64))])
            cost = sizes * (csum[np.asarray(gp)[1:]] - csum[np.asarray(gp)[:-1]])
            order = np.random.default_rng(7).permutation(len(gb))
            sel = order[: int(np.searchsorted(np.cumsum(cost[order]), 1.5e9)) + 1]
            ref = np.zeros((n_trg, 4))
            t = time.perf_counter()
            ref = O.ranges(int(uk), host_args[0], host_args[1], host_args[2], gb, ge, gp, rb, re, sel=sel,
                           nthreads=nth, trg_value=ref)
            cpu_s += time.perf_counter() - t
            cpu_pairs += int(cost[sel].sum())
            rows = np.concatenate([np.arange(gb[i], ge[i]) for i in sel])
            acc[name] = {}
            for prec_name in ("mixed", "fp64"):
                l2, mx = O.max_error(results[(name, prec_name)][rows], ref[rows])
                acc[name][prec_name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": int(rows.size)}
        cpu = {"value": cpu_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"random groups of each list, {cpu_pairs:.3e} pairs in all, {cpu_s:.1f} s; the oracle's "
                          "scalar FP64 restatement of stokes_pvel, OpenMP over groups"),
               **host_info()}
    line("Gpair-interactions/s Stokes PVel FMM surface lists (S2M + L2T + W + X), mixed",
         total_pairs / (ms_mixed * 1e-3) / 1e9, "Gpair/s", args, ms_mixed, "f64+f32 (mixed)",
         {"workload": "pvel-fmm-lists-1M-clustered-leafcap1000-p8", "points": tree.n, "leaves": nl, "nodes": nn,
          "surface_points": cnt, "pairs": total_pairs, "precision": "mixed", "setup_s": setup_s},
         lists=per, fp64_gpair_s=total_pairs / (ms_fp64 * 1e-3) / 1e9, roofline=rl, cpu_baseline=cpu,
         accuracy_vs_oracle=acc)


def wall_images(args, torch, S, W):
    """StkWallFMM's image system (Lib/src/StkWallFMM.cpp:154-292) on 65,536 sources and 65,536
    targets in z in (0, 0.5) above the wall z = 0: Stokes (7 constituent direct sums of n x m pairs:
    the Stokeslet, LapPGrad and LapPGradGrad over originals + images, the dipole sum over images) and
    RPY (10: RPY, LapPGradGrad, LapPGrad and dipole sums over originals + images, the dipole and
    quadrupole gradient sums over images); a pair is one evaluation in one constituent sum."""
    dev = torch.device("cuda:0")
    n = 65_536
    g = np.random.default_rng(31)
    x = g.random((n, 3))
    x[:, 2] = g.uniform(0.0, 0.5, n)
    t = g.random((n, 3))
    t[:, 2] = g.uniform(0.0, 0.5, n)
    res = {}
    for kernel, name, sums, kt in ((S.KERNEL.Stokes, "stokes", 7, 3), (S.KERNEL.RPY, "rpy", 10, 6)):
        f = g.normal(size=(n, 3))
        if kernel == S.KERNEL.RPY:
            f = np.concatenate([f, g.uniform(1e-3, 5e-3, (n, 1))], axis=1)
        d_x, d_f, d_t = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (x, f, t))
        pairs = sums * n * n
        entry = {"pairs": pairs, "constituent_sums": sums}
        outs = {}
        # FP64 only: the library refuses mixed precision for the wall images (the image sums cancel
        # near the wall, which would amplify FP32 rounding beyond 1e-6)
        for prec_name, prec in (("fp64", S.PREC.FP64),):
            out = torch.zeros((n, kt), dtype=torch.float64, device=dev)
            for _ in range(args.warmup):
                S.wall_device(kernel, d_x, d_f, d_t, out, precision=prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                S.wall_device(kernel, d_x, d_f, d_t, out, precision=prec)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) / args.steps
            o = torch.zeros((n, kt), dtype=torch.float64, device=dev)
            S.wall_device(kernel, d_x, d_f, d_t, o, precision=prec)
            torch.cuda.synchronize()
            outs[prec_name] = o.cpu().numpy()
            entry[prec_name] = {"ms_per_step": 1e3 * wall, "gpair_s": pairs / wall / 1e9}
            print(f"# wall-images {name} {prec_name}: {pairs / wall / 1e9:.1f} Gpair/s, {1e3 * wall:.3f} ms",
                  file=sys.stderr, flush=True)
        res[name] = (entry, f, outs)
    cpu = acc = None
    if not args.no_cpu_baseline:
        # the oracle's image system on every source against the first m targets: timed as the CPU
        # baseline and used as the checker
        from oracle import oracle as O
        nth = threads()
        m = 1024
        acc, cpu_pairs, cpu_s = {}, 0, 0.0
        for name, (entry, f, outs) in res.items():
            kernel = 8 if name == "stokes" else 16
            t0 = time.perf_counter()
            ref = O.wall(kernel, x, f, t[:m], nthreads=nth)
            cpu_s += time.perf_counter() - t0
            cpu_pairs += entry["constituent_sums"] * n * m
            acc[name] = {}
            for prec_name, o in outs.items():
                l2, mx = O.max_error(o[:m], ref)
                acc[name][prec_name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": m}
        cpu = {"value": cpu_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"both kernels, all {n} sources x the first {m} targets ({cpu_pairs:.3e} pairs, "
                          f"{cpu_s:.1f} s); the oracle's scalar FP64 image system, OpenMP"),
               **host_info()}
    total = sum(e["pairs"] for e, _, _ in res.values())
    ms = sum(e["fp64"]["ms_per_step"] for e, _, _ in res.values())
    line("Gpair-interactions/s StkWallFMM wall images (Stokes + RPY direct sums), FP64", total / (ms * 1e-3) / 1e9,
         "Gpair/s", args, ms, "f64",
         {"workload": "wall-images-65536x65536", "points": n, "pairs": total, "precision": "fp64"},
         kernels={k: e for k, (e, _, _) in res.items()}, cpu_baseline=cpu, accuracy_vs_oracle=acc)


def stokeslet_instances(args, torch, S, W):
    """512 independent Stokeslet direct sums (src = trg per instance) in one HIP-graph replay.  Variants:
    each precision with the points as given ("unsorted") and with the plan ordering each instance's
    points along a Morton curve inside the same graph ("sorted", stkfmm_tune.sort_points; the ordering
    kernels are part of every replay and of the reported latency)."""
    dev = torch.device("cuda:0")
    sizes, offsets, coord, f = W.stokeslet_instances()
    n = int(offsets[-1])
    pairs = int(np.sum(sizes.astype(np.int64) ** 2))
    d_c, d_f = torch.from_numpy(coord).to(dev), torch.from_numpy(f).to(dev)
    gb, ge = offsets[:-1], offsets[1:]
    res, results = {}, {}
    # FP64: scalar-broadcast staging (unordered: 61.7 -> 59.8 ms item kernel; ordered: 58.2 -> 56.3 ms)
    variants = [("mixed-sorted", S.PREC.MIXED, dict(sort_points=1)),
                ("fp64-sorted", S.PREC.FP64, dict(sort_points=1, broadcast=1)),
                ("mixed", S.PREC.MIXED, None), ("fp64", S.PREC.FP64, dict(broadcast=1))]
    if args.only:
        variants = [v for v in variants if v[0] == args.only]
    stream = torch.cuda.Stream()
    for name, prec, tune in variants:
        plan = S.Plan(S.UKER.STOKES_VEL, prec, n, n, gb, ge, None, np.arange(len(sizes) + 1), gb, ge, tune=tune)
        out = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        # item-kernel time of eager executes (HIP events on the execution stream)
        _, item_ms, reduce_ms, _ = Timed(torch, plan, dict(src_coord=d_c, src_value=d_f, trg_coord=d_c,
                                                          trg_value=out), stream)(3, 1)
        torch.cuda.synchronize()
        plan.capture_graph(d_c, d_f, d_c, out, stream=stream)
        for _ in range(args.warmup):
            plan.launch_graph(stream=stream)
        stream.synchronize()
        lat = []
        for _ in range(args.steps):
            t = time.perf_counter()
            plan.launch_graph(stream=stream)
            stream.synchronize()
            lat.append(time.perf_counter() - t)
        o = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        plan.execute(d_c, d_f, d_c, o, stream=stream)
        stream.synchronize()
        results[name] = o.cpu().numpy()
        res[name] = dict(gpair_s=pairs / float(np.mean(lat)) / 1e9, batch_latency_ms=1e3 * float(np.median(lat)),
                         batch_latency_min_ms=1e3 * float(np.min(lat)), item_ms=item_ms, reduce_ms=reduce_ms,
                         n_items=plan.stats()["n_items"], tune=tune)
        print(f"# stokeslet-instances {name}: {res[name]['gpair_s']:.1f} Gpair/s, batch "
              f"{res[name]['batch_latency_ms']:.3f} ms, item kernel {item_ms:.3f} ms", file=sys.stderr, flush=True)
    head = next(iter(res))
    h = res[head]
    prec_name = "fp64" if head.startswith("fp64") else "mixed"
    kname = {"mixed-sorted": "p2p_items_kernel<6,MIXED,T=2,far-tile form>", "mixed": "p2p_items_kernel<6,MIXED,T=2>",
             "fp64-sorted": "p2p_items_sgpr_kernel<6,FP64,T=2>", "fp64": "p2p_items_sgpr_kernel<6,FP64,T=2>"}[head]
    rl = attach_traffic(roofline("stokeslet", prec_name, pairs, h["item_ms"], kname), "stokeslet-512-instances", head)
    rl["note"] = NOTE + "; kernel_ms = item kernel of an eager execute of the same plan (the replay adds the ordering)"
    cpu = acc = None
    if not args.no_cpu_baseline:
        # the oracle on whole instances in random order until the CPU budget is spent: timed as the CPU
        # baseline and used as the checker of every variant on those instances
        from oracle import oracle as O
        nth = threads()
        order = np.random.default_rng(55).permutation(len(sizes))
        done, cpu_pairs, cpu_s, refs = [], 0, 0.0, {}
        for i in order:
            a, b = int(offsets[i]), int(offsets[i + 1])
            t = time.perf_counter()
            refs[i] = O.direct(6, coord[a:b], f[a:b], coord[a:b], nthreads=nth)
            cpu_s += time.perf_counter() - t
            cpu_pairs += (b - a) ** 2
            done.append(i)
            if cpu_s >= args.cpu_seconds:
                break
        rows = np.concatenate([np.arange(offsets[i], offsets[i + 1]) for i in done])
        ref = np.concatenate([refs[i] for i in done])
        acc = {}
        for name, r in results.items():
            l2, mx = O.max_error(r[rows], ref)
            acc[name] = {"errorL2": l2, "errorMaxRel": mx, "targets_checked": int(rows.size)}
        cpu = {"value": cpu_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": nth, "kind": "port",
               "sample": (f"{len(done)} of {len(sizes)} whole instances in random order, {rows.size} targets, "
                          f"{cpu_pairs:.3e} pairs, {cpu_s:.1f} s; the oracle's scalar FP64 restatement of the "
                          "Stokeslet (PVFMM stokes_vel closed form), OpenMP over targets"),
               **host_info()}
    line("Gpair-interactions/s 512 Stokeslet direct sums in one HIP-graph replay (aggregate)", h["gpair_s"],
         "Gpair/s", args, h["batch_latency_ms"], "f64+f32 (mixed)" if prec_name == "mixed" else "f64",
         {"workload": "stokeslet-512-instances", "instances": len(sizes), "points": n, "pairs": pairs,
          "precision": prec_name, "variant": head,
          "points_order": "sorted on the device inside the graph" if head.endswith("sorted") else "as given"},
         variants=res, roofline=rl, cpu_baseline=cpu, accuracy_vs_oracle=acc)


# micro-kernels of the family workload; the last source value of 20 and 24 is a regularisation length
FAMILY = (14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24)
FAMILY_RADIUS = (20, 24)
# FP64 FLOPs per pair of each family micro-kernel, counted by the VALU counters on its FP64 item kernel
# (64 x (ADD + MUL + TRANS + 2 FMA) / pairs; displacement, masking and accumulation included;
# profiles/family_ulist_262144_fp64_flops_counters.md); the family line's roofline takes them as the
# algorithmic count
FAMILY_FLOPS = {14: 119, 15: 106, 16: 60, 17: 71, 18: 57, 19: 101, 20: 113, 21: 54, 22: 124, 23: 47, 24: 65}


def family_ulist(args, torch, S, W):
    """The rest of the kernel family over one U-list: a 262,144-point tree of the default workload's
    distribution, leaf cap 500, one plan per micro-kernel in each precision (default tuning), every
    plan timed as the headline's (HIP events on the item kernel, wall clock per step), and each
    one's result against the oracle on 8 sampled leaves (the CPU baseline times that oracle)."""
    dev = torch.device("cuda:0")
    x = W.pvel_points(262_144)
    tree = S.Tree(x, 500)
    xs = x[tree.perm]
    d_x = torch.from_numpy(xs).to(dev)
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(31)
    sel = np.sort(rng.choice(tree.n_leaves, 8, replace=False))
    rows = np.concatenate([np.arange(tree.leaf_begin[l], tree.leaf_end[l]) for l in sel])
    sel_pairs = 0
    for l in sel:
        ns = int(np.sum(tree.rng_end[tree.ulist_ptr[l]:tree.ulist_ptr[l + 1]]
                        - tree.rng_begin[tree.ulist_ptr[l]:tree.ulist_ptr[l + 1]]))
        sel_pairs += ns * int(tree.leaf_end[l] - tree.leaf_begin[l])
    order = ["mixed", "fp64"] if args.only not in ("mixed", "fp64") else [args.only]
    kernels, tot_ms = {}, {name: 0.0 for name in order}
    cpu_s = cpu_pairs = 0.0
    for uk in FAMILY:
        ks, kt = S.uker_dims(uk)
        v = rng.uniform(-1.0, 1.0, (tree.n, ks))
        if uk in FAMILY_RADIUS:
            v[:, -1] = np.abs(v[:, -1]) * 0.01
        d_v = torch.from_numpy(v).to(dev)
        out = torch.zeros((tree.n, kt), dtype=torch.float64, device=dev)
        kw = dict(src_coord=d_x, src_value=d_v, trg_coord=d_x, trg_value=out)
        r, res = {}, {}
        for name in order:
            prec = S.PREC.MIXED if name == "mixed" else S.PREC.FP64
            plan = S.Plan.from_tree(tree, uk, prec)
            wall, item_ms, reduce_ms, _ = Timed(torch, plan, kw, stream)(args.steps, args.warmup)
            o = torch.zeros_like(out)
            plan.execute(stream=stream, **dict(kw, trg_value=o))
            torch.cuda.synchronize()
            res[name] = o.cpu().numpy()
            r[name] = dict(gpair_s=tree.pairs * args.steps / wall / 1e9, ms_per_step=1e3 * wall / args.steps,
                           item_ms=item_ms, reduce_ms=reduce_ms)
            tot_ms[name] += 1e3 * wall / args.steps
            del plan
        if not args.no_cpu_baseline:
            from oracle import oracle as O
            t = time.perf_counter()
            ref = O.ranges(uk, xs, v, xs, tree.leaf_begin, tree.leaf_end, tree.ulist_ptr, tree.rng_begin,
                           tree.rng_end, sel=sel, nthreads=threads())
            cpu_s += time.perf_counter() - t
            cpu_pairs += sel_pairs
            for name in order:
                l2, mx = O.max_error(res[name][rows], ref[rows])
                r[name]["errorL2"], r[name]["errorMaxRel"] = l2, mx
        kernels[S.UKER(uk).name.lower()] = dict(uker=uk, **r)
        print(f"# family-ulist {S.UKER(uk).name.lower()}: " +
              ", ".join(f"{n} {r[n]['gpair_s']:.1f} Gpair/s ({r[n]['item_ms']:.3f} ms)" for n in order),
              file=sys.stderr, flush=True)
    head = order[0]
    # the dominant kernel: the family member with the longest item kernel in the head precision
    dom = max(kernels.values(), key=lambda k: k[head]["item_ms"])
    uk = dom["uker"]
    rl = None
    if uk in FAMILY_FLOPS:
        f = FAMILY_FLOPS[uk]
        FLOPS["family"] = (f, 3 + S.uker_dims(uk)[1])
        # the library's default targets per lane: 1 in mixed mode for the family members with many
        # outputs (the plain item kernel), else 2 (mixed: the far-tile instantiation)
        t1 = head == "mixed" and uk in (14, 15, 17, 19, 20, 22)
        rl = roofline("family", head, tree.pairs, dom[head]["item_ms"],
                      f"p2p_items_kernel<{uk},{'MIXED' if head == 'mixed' else 'FP64'},T={1 if t1 else 2}"
                      f"{',far-tile form' if head == 'mixed' and not t1 else ''}>")
        rl["note"] = NOTE + "; FLOPs per pair counted by the VALU counters on the FP64 kernel (FMA = 2)"
        attach_traffic(rl, "family-ulist-262144-clustered-leafcap500", f"uk{uk}-{head}")
    cpu = None
    if cpu_s > 0:
        cpu = {"value": cpu_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": threads(), "kind": "port",
               "sample": (f"8 sampled leaves ({sel_pairs:.3e} pairs) per micro-kernel, all {len(FAMILY)}; the "
                          "oracle's scalar FP64 restatement of each reference uKerEval, OpenMP over targets"),
               **host_info()}
    n_k = len(FAMILY)
    line(f"Gpair-interactions/s rest of the kernel family over an adaptive-octree U-list ({head})",
         n_k * tree.pairs / (tot_ms[head] * 1e-3) / 1e9, "Gpair/s", args, tot_ms[head] / n_k,
         "f64+f32 (mixed)" if head == "mixed" else "f64",
         {"workload": "family-ulist-262144-clustered-leafcap500", "points": tree.n, "leaves": tree.n_leaves,
          "pairs_per_kernel": tree.pairs, "kernels": n_k, "precision": head},
         kernels=kernels, roofline=rl, cpu_baseline=cpu,
         aggregate_gpair_s={n: n_k * tree.pairs / (tot_ms[n] * 1e-3) / 1e9 for n in order})


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "bench.py measures one GPU only: the P2P path runs on a single MI355X "
                                   "(multi-GPU sharding is a design note, not a measured mode)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import stkfmm_amd as S
    from stkfmm_amd import workloads as W

    if args.copy_calibration:
        n = 1 << 27  # 1 GiB of doubles
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))
        return
    if args.workload == "pvel-ulist":
        t4m = None if args.no_target_4m else {"target_4m": target_4m(args, torch, S, W)}
        torch.cuda.empty_cache()
        ulist_bench(args, torch, S, "pvel_fused", S.UKER.STOKES_PVEL_FUSED, W.pvel_points, 1000, W.pvel_values, 4,
                    "Gpair-interactions/s StokesPVel P2P: FP64 vs mixed at ≤1e-6 rel err",
                    "stokes-pvel-ulist-1M-clustered-leafcap1000", ["mixed", "fp64", "per-leaf"], extra_fields=t4m)
    elif args.workload == "pvel-ulist-4m":
        # the same point distribution at 4M points (the size of the batched-vs-per-leaf target)
        ulist_bench(args, torch, S, "pvel_fused", S.UKER.STOKES_PVEL_FUSED, lambda: W.pvel_points(4_000_000), 1000,
                    W.pvel_values, 4, "Gpair-interactions/s StokesPVel P2P: FP64 vs mixed at ≤1e-6 rel err",
                    "stokes-pvel-ulist-4M-clustered-leafcap1000", ["mixed", "fp64", "per-leaf"])
    elif args.workload == "lapgrad-direct":
        lapgrad_direct(args, torch, S, W)
    elif args.workload == "rpy-ulist":
        ulist_bench(args, torch, S, "rpy", S.UKER.RPY_ULAPU, W.rpy_points, 2000, lambda n: (W.rpy_values(n),), 6,
                    "Gpair-interactions/s RPY P2P over an adaptive-octree U-list, mixed (autotuned) vs FP64",
                    "rpy-ulist-4M-lognormal-leafcap2000", ["mixed", "fp64"])
    elif args.workload == "pvel-adjoint":
        pvel_adjoint(args, torch, S, W)
    elif args.workload == "pvel-fmm-lists":
        pvel_fmm_lists(args, torch, S, W)
    elif args.workload == "wall-images":
        wall_images(args, torch, S, W)
    elif args.workload == "family-ulist":
        family_ulist(args, torch, S, W)
    else:
        stokeslet_instances(args, torch, S, W)


if __name__ == "__main__":
    main()
