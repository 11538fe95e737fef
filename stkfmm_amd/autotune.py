"""Autotuner for batched P2P plans.

For a work list (an octree U-list or a batch of direct sums) and a micro-kernel/precision pair it
times candidate tunings of ``stkfmm_tune`` -- targets per lane (1/2/4), sources per work item, LDS or
scalar-broadcast source staging, the scheduling mode and one or per-XCD work queues -- on the real
data, with HIP events on the execution stream, and keeps the fastest.  Results are cached in JSON
under a key made of the GPU architecture, micro-kernel, precision and a coarse signature of the work list (target count, pair count and group count,
bucketed by powers of two), so a later plan of the same class reuses the choice without timing.

The cache is read from ``tuned_defaults.json`` next to this file (choices measured on MI355X for
the benchmark workloads, shipped with the package) and from the user cache
(``$STKFMM_AUTOTUNE_CACHE`` or ``~/.cache/stkfmm_amd/autotune.json``), which ``autotune`` updates.
"""
import itertools
import json
import math
import os

from . import Plan, device_info

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(HERE, "tuned_defaults.json")

# LDS-staged tiles over the full grid; scalar-broadcast records (wins for FP64 on MI355X) over the
# sizes where it is competitive
CANDIDATES = ([dict(targets_per_lane=t, max_src_per_item=m, persistent=p, broadcast=0)
               for t, m, p in itertools.product((1, 2, 4), (1024, 2048, 4096), (1, 2))]
              + [dict(targets_per_lane=t, max_src_per_item=m, persistent=p, broadcast=1)
                 for t, m, p in itertools.product((1, 2, 4), (1024, 2048), (1, 2))]
              + [dict(targets_per_lane=t, max_src_per_item=m, persistent=1, queues=8)
                 for t, m in itertools.product((1, 2), (1024, 2048))])


def cache_path():
    return os.environ.get("STKFMM_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "stkfmm_amd", "autotune.json"))


def _bucket(x):
    return int(round(math.log2(max(1, x))))


def signature(uker, precision, n_trg, pairs, n_groups, arch=None):
    arch = arch or device_info()["arch"].split(":")[0]
    return f"{arch}|uk{int(uker)}|p{int(precision)}|n{_bucket(n_trg)}|q{_bucket(pairs)}|g{_bucket(n_groups)}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(key):
    """Cached tuning for a signature, user cache first, then the shipped defaults; None if absent."""
    for path in (cache_path(), DEFAULTS_PATH):
        entry = _load(path).get(key)
        if entry:
            return entry["tune"]
    return None


def _store(key, entry):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    data = _load(path)
    data[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(work, uker, precision, run_args, candidates=None, reps=3, warmup=1, use_cache=True, verbose=False):
    """Pick the fastest tuning for ``work`` and return (tune dict, timings list).

    ``work`` is a dict with the Plan work-list arguments (n_src, n_trg, grp_trg_begin, grp_trg_end,
    grp_level, grp_rng_ptr, rng_begin, rng_end) or a ``Tree``.  ``run_args`` is a dict of the
    ``Plan.execute`` keyword arguments (device tensors); the output tensor is accumulated into, so
    pass a scratch output.
    """
    import torch

    if hasattr(work, "leaf_begin"):
        tree = work
        work = dict(n_src=tree.n, n_trg=tree.n, grp_trg_begin=tree.leaf_begin, grp_trg_end=tree.leaf_end,
                    grp_level=tree.leaf_level, grp_rng_ptr=tree.ulist_ptr, rng_begin=tree.rng_begin,
                    rng_end=tree.rng_end)
    first = Plan(uker, precision, tune=None, **work)
    pairs = first.stats()["pairs"]
    key = signature(uker, precision, work["n_trg"], pairs, len(work["grp_trg_begin"]))
    if use_cache:
        hit = lookup(key)
        if hit is not None:
            return hit, []
    del first
    timings = []
    best = None
    for cand in candidates or CANDIDATES:
        plan = Plan(uker, precision, tune=cand, **work)
        for _ in range(warmup):
            plan.execute(**run_args)
        torch.cuda.synchronize()
        plan.set_timing(True)
        for _ in range(reps):
            plan.execute(**run_args)
        torch.cuda.synchronize()
        item_ms, reduce_ms, _ = plan.kernel_ms()
        ms = item_ms + reduce_ms
        timings.append(dict(tune=cand, ms=ms))
        if verbose:
            print(f"autotune {key} {cand} {ms:.3f} ms", flush=True)
        if best is None or ms < best[1]:
            best = (cand, ms)
        del plan
    entry = dict(tune=best[0], ms=best[1], gpair_s=pairs / best[1] / 1e6, pairs=pairs)
    if use_cache:
        _store(key, entry)
    return best[0], timings
