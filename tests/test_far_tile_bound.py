"""The mixed mode's far-tile displacement (stkfmm_amd/csrc/p2p_items.hpp, STKP2P_FAR_F32): for a source s
with |h| + |s - c| <= R dist(s, box) (c the centre and |h| the half-diagonal of the box of an item's
targets), the displacement is formed as fl32(fl32(t - c) - fl32(s - c)) instead of fl32(t - s).  The
claim is that each component then errs by at most (R + 1) 2^-24 r, r = |t - s|.

Checked here by emulating that arithmetic in IEEE FP32 (numpy rounds to nearest even, as the device's
v_sub_f32 and v_cvt_f32_f64 do) on random target boxes of several sizes and positions, against the
exact FP64 displacement; and the test itself is checked to be sharp enough to fail when the criterion is
dropped (sources next to the box)."""
import numpy as np
import pytest

R = 4.0          # STKP2P_FAR_RATIO
EPS32 = 2.0 ** -24


def far_displacement(t, s, c):
    tf = (t - c).astype(np.float32)
    sf = (s - c).astype(np.float32)
    return (tf[:, None, :] - sf[None, :, :]).astype(np.float64)


def box_of(t):
    lo, hi = t.min(0), t.max(0)
    return 0.5 * (lo + hi), 0.5 * (hi - lo)


def passes(s, c, h):
    gap = np.maximum(np.abs(s - c) - h, 0.0)
    return np.linalg.norm(h) + np.linalg.norm(s - c, axis=1) <= R * np.linalg.norm(gap, axis=1)


def worst_ratio(t, s, c):
    exact = t[:, None, :] - s[None, :, :]
    r = np.linalg.norm(exact, axis=2)
    err = np.abs(far_displacement(t, s, c) - exact).max(axis=2)
    return float((err / (EPS32 * r)).max())


@pytest.mark.parametrize("seed", range(6))
def test_far_tile_error_within_bound(seed):
    g = np.random.default_rng(seed)
    for _ in range(20):
        # a target box of 128 points: random size (leaf-like, down to deep levels) and position, in
        # the level-scaled frame (coordinates up to ~2^10)
        size = 10.0 ** g.uniform(-4, 0)
        origin = g.uniform(0, 1024.0, 3)
        t = origin + size * g.random((128, 3))
        c, h = box_of(t)
        # candidate sources around the box, at distances from touching to a few box sizes
        s = c + size * g.uniform(-6, 6, (4000, 3))
        s = s[passes(s, c, h)]
        if len(s) == 0:
            continue
        assert worst_ratio(t, s, c) <= R + 1.0


def test_bound_is_sharp_enough_to_catch_near_sources():
    # without the criterion (sources close to the box, cancellation in fl32(t - c) - fl32(s - c)), the
    # same check fails: the test can tell the far-tile form from an unconditional one
    g = np.random.default_rng(11)
    t = 512.0 + 0.01 * g.random((128, 3))
    c, h = box_of(t)
    s = t[:64] + 1e-7 * g.normal(size=(64, 3))
    assert not np.all(passes(s, c, h))
    assert worst_ratio(t, s, c) > R + 1.0
