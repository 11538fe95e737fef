"""GPU: the wall image system (StkWallFMM's evalStokes / evalRPY with direct sums) against the oracle,
through the host and the device entries, FP64 and mixed; and no slip on the wall from the GPU."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from tests.gpu_util import FP64_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu

# Mixed mode: the image sums cancel the direct sums near the wall (u = 0 on it), so the FP32 rounding
# of each constituent sum (~1e-7 of its magnitude) is amplified by the ratio of those magnitudes to
# the assembled velocity: measured maxrel ~4e-6 on these inputs (targets on and near the wall).
WALL_MIXED_TOL = 2e-5


def problem(kernel, seed, ns=700, nt=500):
    rng = np.random.default_rng(seed)
    x = rng.random((ns, 3))
    x[:, 2] = rng.uniform(0.0, 0.5, ns)
    f = rng.normal(size=(ns, 3))
    if kernel == S.KERNEL.RPY:
        f = np.concatenate([f, rng.uniform(1e-3, 5e-3, (ns, 1))], axis=1)
    t = rng.random((nt, 3))
    t[:, 2] = rng.uniform(0.0, 0.5, nt)
    k = min(8, ns, nt)
    t[:k] = x[:k]          # targets at source points
    t[k:min(40, nt), 2] = 0.0   # targets on the wall
    return x, f, t


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kernel", [S.KERNEL.Stokes, S.KERNEL.RPY])
def test_wall_device_vs_oracle(kernel, prec):
    say(f"wall device kernel={int(kernel)} prec={prec}")
    x, f, t = problem(kernel, 10 + int(kernel))
    kt = 3 if kernel == S.KERNEL.Stokes else 6
    ref = O.wall(int(kernel), x, f, t)
    out = zeros(len(t), kt)
    S.wall_device(kernel, to_dev(x), to_dev(np.ascontiguousarray(f)), to_dev(t), out, precision=prec)
    got = host(out)
    l2, mx = O.max_error(got, ref)
    tol = FP64_TOL if prec == 0 else WALL_MIXED_TOL
    assert l2 < tol and mx < tol, (l2, mx)
    if prec == 0:   # no slip on the wall, from the GPU's own sums
        scale = np.abs(got[:, :3]).max()
        assert np.abs(got[8:40, :3]).max() < 1e-13 * scale


@pytest.mark.parametrize("kernel", [S.KERNEL.Stokes, S.KERNEL.RPY])
def test_wall_host_api_accumulates(kernel):
    say(f"wall host kernel={int(kernel)}")
    x, f, t = problem(kernel, 20 + int(kernel), ns=333, nt=211)
    kt = 3 if kernel == S.KERNEL.Stokes else 6
    pre = np.random.default_rng(1).normal(size=(len(t), kt))
    out = pre.copy()
    S.wall_evaluate(kernel, x, np.ascontiguousarray(f), t, out)
    l2, mx = O.max_error(out - pre, O.wall(int(kernel), x, f, t))
    assert l2 < FP64_TOL and mx < FP64_TOL


def test_wall_repeated_calls_with_other_sizes():
    # the workspace grows and is reused; results stay right across shapes
    say("wall sizes")
    for ns, nt in ((50, 20), (1000, 900), (10, 5), (1, 1)):
        x, f, t = problem(S.KERNEL.RPY, ns + nt, ns=ns, nt=nt)
        out = zeros(nt, 6)
        S.wall_device(S.KERNEL.RPY, to_dev(x), to_dev(np.ascontiguousarray(f)), to_dev(t), out)
        l2, mx = O.max_error(host(out), O.wall(16, x, f, t))
        assert mx < FP64_TOL or np.all(O.wall(16, x, f, t) == 0)
