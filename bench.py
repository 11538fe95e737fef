"""Benchmark: Stokes PVel P2P over the U-list of an adaptive octree on one MI355X.

Workload (the headline metric's configuration): 1,000,000 points, half uniform in [0,1)^3 and half
in 16 Gaussian clusters (sigma 0.01), leaf cap 1000; single layer f ~ N(0,1)^3 (TrD = 0) and double
layer 9 x N(0,1) at every point (SL = DL = targets; one pair = both layers).  One step = one full
U-list evaluation (all leaves, one batched persistent launch + the partial-sum reduction), inputs
resident in HBM.

Reported: value = Gpair/s of the mixed-precision batched launch (whole job, 1 GPU); beside it the
FP64 batched launch and the plain FP64 per-leaf-launch version (one launch per leaf), accuracy of
both modes against the FP64 oracle on sampled leaves, the roofline of the dominant kernel (timed with
HIP events on its own stream) and the CPU baseline (the oracle's vectorised C restatement of the
reference P2P on the host cores, on a bounded sample of leaves).

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpair-interactions/s StokesPVel P2P: FP64 vs mixed at ≤1e-6 rel err"
FLOPS_PER_PAIR = 72            # stokes_pvel + stokes_doublepvel fused, algorithmic count
FLOPS_FP64_MIXED = 7           # FP64 part in mixed mode (3 displacement subtractions + 4 accumulations)
PEAK_FP64 = 78.6e12            # MI355X vector FP64, FLOP/s
PEAK_FP32 = 157.3e12           # MI355X vector FP32 (packed), FLOP/s
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--leaf-cap", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work for the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-leaf", action="store_true")
    ap.add_argument("--only", choices=["mixed", "fp64", "per-leaf"], default=None,
                    help="run one variant only (for profiling); prints its own line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy of known size (PMC calibration) and exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU only: the P2P path runs on a single "
                                   "MI355X (multi-GPU sharding is a design note, not a measured mode)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import stkfmm_amd as S
    from stkfmm_amd import workloads as W

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        n = 1 << 27  # 1 GiB of doubles
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))
        return

    t0 = time.perf_counter()
    x = W.pvel_points(args.points)
    tree = S.Tree(x, args.leaf_cap)
    xs = x[tree.perm]
    sl, dl = W.pvel_values(tree.n)
    setup_s = time.perf_counter() - t0
    d_x = torch.from_numpy(xs).to(dev)
    d_sl = torch.from_numpy(sl).to(dev)
    d_dl = torch.from_numpy(dl).to(dev)
    out = torch.zeros((tree.n, 4), dtype=torch.float64, device=dev)
    pairs = tree.pairs
    stream = torch.cuda.Stream()

    from stkfmm_amd.autotune import autotune

    def run_variant(prec, tune, steps, warmup):
        if tune == "autotune":
            scratch = torch.zeros((tree.n, 4), dtype=torch.float64, device=dev)
            tune, _ = autotune(tree, S.UKER.STOKES_PVEL_FUSED, prec,
                               dict(src_coord=d_x, src_value=d_sl, trg_coord=d_x, trg_value=scratch,
                                    src_value2=d_dl, stream=stream))
            del scratch
        plan = S.Plan.from_tree(tree, S.UKER.STOKES_PVEL_FUSED, prec, tune=tune)
        res = torch.zeros((tree.n, 4), dtype=torch.float64, device=dev)
        with torch.cuda.stream(stream):
            for _ in range(warmup):
                plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl, stream=stream)
            torch.cuda.synchronize()
            plan.set_timing(True)
            t = time.perf_counter()
            for _ in range(steps):
                plan.execute(d_x, d_sl, d_x, out, src_value2=d_dl, stream=stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t
            item_ms, reduce_ms, count = plan.kernel_ms()
            plan.set_timing(False)
            plan.execute(d_x, d_sl, d_x, res, src_value2=d_dl, stream=stream)
            torch.cuda.synchronize()
        st = plan.stats()
        assert st["pairs"] == pairs
        return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, gpair_s=pairs * steps / wall / 1e9,
                    item_ms=item_ms, reduce_ms=reduce_ms, timed_launches=count, n_items=st["n_items"],
                    n_tiles=st["n_tiles"], grid=st["grid"], tune=tune), res

    variants = {}
    results = {}
    plan_order = ["mixed", "fp64", "per-leaf"] if args.only is None else [args.only]
    for name in plan_order:
        if name == "per-leaf" and args.no_per_leaf:
            continue
        prec = S.PREC.MIXED if name == "mixed" else S.PREC.FP64
        # batched variants use the autotuner's choice (cached per workload class); the per-leaf
        # baseline launches every leaf separately with the library's default tiles
        tune = dict(persistent=3) if name == "per-leaf" else "autotune"
        steps = args.steps if name != "per-leaf" else max(1, min(args.steps, 3))
        warm = args.warmup if name != "per-leaf" else 1
        variants[name], results[name] = run_variant(prec, tune, steps, warm)
        print(f"# {name}: {variants[name]['gpair_s']:.1f} Gpair/s, {variants[name]['ms_per_step']:.3f} ms/step, "
              f"item kernel {variants[name]['item_ms']:.3f} ms", file=sys.stderr, flush=True)

    head = "mixed" if "mixed" in variants else plan_order[0]
    hv = variants[head]

    # roofline of the dominant kernel (the item kernel of the headline variant)
    item_s = hv["item_ms"] * 1e-3
    achieved = pairs * FLOPS_PER_PAIR / item_s
    if head == "mixed":
        peak = FLOPS_PER_PAIR / (FLOPS_FP64_MIXED / PEAK_FP64 + (FLOPS_PER_PAIR - FLOPS_FP64_MIXED) / PEAK_FP32)
    else:
        peak = PEAK_FP64
    roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "kernel": f"p2p_items_kernel<STOKES_PVEL_FUSED,{'MIXED' if head == 'mixed' else 'FP64'},"
                          f"T={(hv['tune'] or {}).get('targets_per_lane', 2)}>",
                "kernel_ms": hv["item_ms"], "flops_per_pair": FLOPS_PER_PAIR, "pairs_per_launch": pairs,
                "note": "VALU-bound; peak uses MI355X published vector FP64 78.6 and packed FP32 157.3 TFLOP/s; "
                        "this build keeps packed FP32 disabled, so FP32 runs at the FP64 VALU rate"}
    traffic_file = os.path.join(ROOT, "profiles", "pvel_ulist_1M_mixed_traffic.json")
    if head == "mixed" and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        roofline["traffic"] = tr.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = "profiles/pvel_ulist_1M_mixed_traffic.json (rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE)"

    # accuracy and CPU baseline on sampled leaves (oracle = checker and timed CPU restatement)
    accuracy = None
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = max(1, min(16, os.cpu_count() or 1))
        g = np.random.default_rng(12345)
        order = g.permutation(tree.n_leaves)
        sz = (tree.leaf_end - tree.leaf_begin).astype(np.int64)
        nsrc = np.array([int((tree.rng_end[tree.ulist_ptr[l]:tree.ulist_ptr[l + 1]] -
                              tree.rng_begin[tree.ulist_ptr[l]:tree.ulist_ptr[l + 1]]).sum())
                         for l in range(tree.n_leaves)], dtype=np.int64)
        vals = np.concatenate([sl, dl], axis=1)
        # evaluate random leaves in chunks until about cpu-seconds of CPU work has been done
        ref = np.zeros((tree.n, 4))
        done = []
        cpu_s = 0.0
        pos = 0
        chunk = 4 * threads
        while pos < tree.n_leaves and cpu_s < args.cpu_seconds:
            sel_c = order[pos:pos + chunk]
            pos += len(sel_c)
            t = time.perf_counter()
            ref = O.ranges(13, xs, vals, xs, tree.leaf_begin, tree.leaf_end, tree.ulist_ptr, tree.rng_begin,
                           tree.rng_end, sel=sel_c, nthreads=threads, simd=True, trg_value=ref)
            cpu_s += time.perf_counter() - t
            done.append(sel_c)
        sel = np.concatenate(done)
        sample_pairs = int((sz[sel] * nsrc[sel]).sum())
        rows = np.concatenate([np.arange(tree.leaf_begin[l], tree.leaf_end[l]) for l in sel])
        cpu = {"value": sample_pairs / cpu_s / 1e9, "unit": "Gpair/s", "cores": threads, "kind": "port",
               "sample": f"{len(sel)} of {tree.n_leaves} leaves (random), {rows.size} targets, "
                         f"{sample_pairs:.3e} pairs, {cpu_s:.1f} s; FP64 C restatement of the reference "
                         "uKerEval algebra, AVX2 lanes over targets, OpenMP over leaves "
                         "(PVFMM SIMD P2P not buildable offline)",
               "cpu_model": _cpu_model()}
        accuracy = {}
        for name, res in results.items():
            got = res.cpu().numpy()[rows]
            l2, mx = O.max_error(got, ref[rows])
            accuracy[name] = {"errorL2": l2, "errorMaxRel": mx}

    value = hv["gpair_s"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Gpair/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hv["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (mixed)" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "stokes-pvel-ulist-1M-clustered-leafcap1000", "points": tree.n,
                   "leaf_cap": args.leaf_cap, "leaves": tree.n_leaves, "pairs": pairs,
                   "layers": "single+double layer fused (SL = DL = targets)", "precision": head,
                   "setup_s": setup_s},
        "variants": variants,
        "accuracy_vs_oracle": accuracy,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if "per-leaf" in variants and head == "mixed":
        line["speedup_vs_fp64_per_leaf_launch"] = value / variants["per-leaf"]["gpair_s"]
    if "fp64" in variants and head == "mixed":
        line["speedup_vs_fp64_batched"] = value / variants["fp64"]["gpair_s"]
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
