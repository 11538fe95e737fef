"""GPU: the device box mapping and value scaling equal the host versions bit for bit, and a scaled-box
U-list plan (FMMData::evaluateFMM's sequence around the plan) equals the direct evaluation."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O
from tests.gpu_util import FP64_TOL, host, say, to_dev, zeros

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024])
def test_device_scaling_bitwise_equals_host(k):
    say(f"scaling kernel={k}")
    rng = np.random.default_rng(k)
    ks, kd, kt = S.get_kernel_dimension(k)
    sl, dl, tv = rng.normal(size=(300, ks)), rng.normal(size=(200, kd)), rng.normal(size=(250, kt))
    c = rng.normal(size=(400, 3)) * 5
    h_sl, h_dl, h_tv, h_c = sl.copy(), dl.copy(), tv.copy(), c.copy()
    S.scale_src(k, h_sl, h_dl, 0.37)
    S.scale_trg(k, h_tv, 0.37)
    S.scale_coord(h_c, [1.0, -2.0, 0.5], 0.37, S.PAXIS.PXY)
    d_sl, d_dl, d_tv, d_c = to_dev(sl), to_dev(dl), to_dev(tv), to_dev(c)
    S.scale_src(k, d_sl, d_dl, 0.37)
    S.scale_trg(k, d_tv, 0.37)
    S.scale_coord(d_c, [1.0, -2.0, 0.5], 0.37, S.PAXIS.PXY)
    for a, b in ((d_sl, h_sl), (d_dl, h_dl), (d_tv, h_tv), (d_c, h_c)):
        assert np.array_equal(host(a), b)


def test_scaled_box_ulist_plan_equals_direct():
    say("scaled box U-list")
    rng = np.random.default_rng(9)
    origin, length = np.array([10.0, -4.0, 2.0]), 3.5
    x = origin + rng.random((5000, 3)) * length
    v = rng.normal(size=(5000, 4))
    xs = x.copy()
    S.scale_coord(xs, origin, 1 / length)
    tree = S.Tree(xs, 200)
    d_x = to_dev(xs[tree.perm])
    d_v = to_dev(v[tree.perm])
    S.scale_src(S.KERNEL.PVel, d_v, None, 1 / length)
    out = zeros(len(x), 4)
    plan = S.Plan.from_tree(tree, 7, S.PREC.FP64)
    plan.execute(d_x, d_v, d_x, out, perm=to_dev(tree.perm.astype(np.int32)))
    S.scale_trg(S.KERNEL.PVel, out, 1 / length)
    # the U-list is the near field only: compare on the near field computed by the oracle directly
    near = O.ranges(7, x[tree.perm], v[tree.perm], x[tree.perm], tree.leaf_begin, tree.leaf_end, tree.ulist_ptr,
                    tree.rng_begin, tree.rng_end)
    expect = np.zeros_like(near)
    expect[tree.perm] = near
    l2, mx = O.max_error(host(out), expect)
    assert l2 < FP64_TOL and mx < FP64_TOL
