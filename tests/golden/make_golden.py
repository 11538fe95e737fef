"""Write tests/golden/simplekernel_fixtures.npz from the reference's own scalar oracle.

The reference's Test/SimpleKernel.cpp is compiled where it lies (oracle/Makefile `ref` target, output
in oracle/_ref/) and driven through oracle/ref_shim.cpp with the accumulation of runSimpleKernel
(Test/Test.cpp:478-510).  Inputs: the reference's fixed source points (Test/Util/PointDistribution.cpp
:15-48, box 1, shift 0) plus seeded random points; targets include points coincident with sources
(the reference's own direct tests use SL = DL = target points, Test/Test.cpp:183,191).

Run in a container that has the reference tree:  python tests/golden/make_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# (KERNEL value, layer, kdim_src, kdim_trg, value rule)
CASES = [
    (1, 0, 1, 4, "uniform"), (1, 1, 3, 4, "uniform"),
    (2, 0, 1, 10, "uniform"), (2, 1, 3, 10, "uniform"),
    (8, 0, 3, 3, "uniform"),
    (16, 0, 4, 6, "radius"),
    (32, 0, 4, 3, "radius"),
    (128, 0, 4, 4, "uniform"), (128, 1, 9, 4, "uniform"),
    # appended after the first set (the random stream of the cases above is unchanged)
    (256, 0, 4, 16, "uniform"), (256, 1, 9, 16, "uniform"),
    (512, 0, 4, 7, "uniform"), (512, 1, 9, 7, "uniform"),
    (1024, 0, 4, 9, "uniform"), (1024, 1, 9, 9, "uniform"),
    (64, 0, 7, 6, "radius"),
    (4, 0, 9, 10, "uniform"),
]

FIXED = np.array([[0.3, 0.2, 0.1], [0.2, 0.4, 0.3],                        # fixedPoints(2)
                  [0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.3, 0.3, 0.3], [0.4, 0.4, 0.4]])  # fixedPoints(4)


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    if not O.ref_available():
        raise SystemExit("reference oracle could not be built (reference tree absent?)")
    rng = np.random.default_rng(20191223)
    out = {}
    for K, layer, ks, kt, rule in CASES:
        ns, nt = 40, 30
        src = np.concatenate([FIXED, rng.random((ns - len(FIXED), 3))])
        trg = np.concatenate([src[:8], rng.random((nt - 8, 3))])   # 8 coincident targets
        sv = rng.uniform(-1.0, 1.0, (ns, ks))                          # Test/Test.cpp:318-319
        if rule == "radius":                                          # Test/Test.cpp:321-328, eps = 0.01
            sv[:, -1] = np.abs(sv[:, -1]) * 0.01
        expect = O.simplekernel_direct(K, layer, ks, kt, src, sv, trg)
        tag = f"K{K}_L{layer}"
        out[tag + "_src"] = src
        out[tag + "_sv"] = sv
        out[tag + "_trg"] = trg
        out[tag + "_out"] = expect
        # single pairs, including a coincident one, for per-pair checks
        pr = rng.random((16, 3)) * 0.5 + 0.25
        ps = np.concatenate([pr[:1], rng.random((15, 3))])
        pf = rng.uniform(-1.0, 1.0, (16, ks))
        if rule == "radius":
            pf[:, -1] = np.abs(pf[:, -1]) * 0.01
        pout = np.array([O.simplekernel_pair(K, layer, ps[i], pr[i], pf[i])[:kt] for i in range(16)])
        out[tag + "_pair_s"] = ps
        out[tag + "_pair_t"] = pr
        out[tag + "_pair_f"] = pf
        out[tag + "_pair_out"] = pout
    path = os.path.join(HERE, "simplekernel_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
