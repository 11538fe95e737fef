"""The oracle's restatement of StkWallFMM's wall image system (Lib/src/StkWallFMM.cpp:154-292) with
direct sums, pinned by what the physics fixes:

  * no slip: the Stokes and RPY velocities vanish on the wall z = 0 (a wrong sign, image position or
    assembly index breaks this);
  * the RPY system at b = 0 is the Stokes system (its b-dependent sums vanish);
  * incompressibility of the Stokes wall flow, and the RPY Laplacian output at b = 0 equals the
    finite-difference Laplacian of the Stokes wall velocity;
  * far from the wall the Stokes system tends to the free-space Stokeslet (image terms ~ d / h);
  * linearity in the forces.
and the C ABI's argument checks (no compute call: there is no GPU here)."""
import numpy as np
import pytest

import stkfmm_amd as S
from oracle import oracle as O


def sources(rng, n, zlo=0.05, zhi=0.5, b=None):
    x = rng.random((n, 3))
    x[:, 2] = rng.uniform(zlo, zhi, n)
    f = rng.normal(size=(n, 3))
    if b is not None:
        f = np.concatenate([f, np.full((n, 1), b) if np.isscalar(b) else b], axis=1)
    return x, f


@pytest.mark.parametrize("kernel,b", [(8, None), (16, 0.02), (16, 0.0)])
def test_no_slip_on_the_wall(kernel, b):
    rng = np.random.default_rng(kernel)
    x, f = sources(rng, 40, b=b)
    trg = rng.random((50, 3)) * 2 - 0.5
    trg[:, 2] = 0.0
    u = O.wall(kernel, x, f, trg)
    free = O.direct(6, x, f[:, :3], trg)
    assert np.abs(u[:, :3]).max() < 1e-14 * np.abs(free).max()


def test_rpy_at_zero_radius_is_stokes():
    rng = np.random.default_rng(3)
    x, f = sources(rng, 30, b=0.0)
    trg = rng.random((40, 3))
    a = O.wall(16, x, f, trg)[:, :3]
    s = O.wall(8, x, f[:, :3], trg)
    np.testing.assert_allclose(a, s, rtol=1e-13, atol=1e-14 * np.abs(s).max())


def test_stokes_wall_flow_is_incompressible_and_rpy_laplacian_matches():
    rng = np.random.default_rng(4)
    x, f = sources(rng, 6, 0.2, 0.5)
    f0 = np.concatenate([f, np.zeros((6, 1))], axis=1)
    vel = lambda t: O.wall(8, x, f, t[None])[0]  # noqa: E731
    for t in rng.random((5, 3)) * np.array([1, 1, 0.15]) + np.array([0, 0, 0.02]):
        J = np.zeros((3, 3))
        for j in range(3):
            e = np.zeros(3)
            e[j] = 1e-6
            J[:, j] = (vel(t + e) - vel(t - e)) / 2e-6
        assert abs(np.trace(J)) < 1e-7 * np.abs(J).max()
        H = 2e-4
        lap = -6 * vel(t)
        for j in range(3):
            e = np.zeros(3)
            e[j] = H
            lap = lap + vel(t + e) + vel(t - e)
        lap /= H * H
        out = O.wall(16, x, f0, t[None])[0, 3:]
        np.testing.assert_allclose(out, lap, rtol=2e-4, atol=2e-4 * np.abs(lap).max())


def test_far_from_the_wall_is_free_space():
    rng = np.random.default_rng(5)
    h = 1e3
    x = np.array([[0.2, 0.3, h]])
    f = rng.normal(size=(1, 3))
    trg = x + rng.normal(size=(10, 3)) * 0.05
    u = O.wall(8, x, f, trg)
    free = O.direct(6, x, f, trg)
    d = np.linalg.norm(trg - x, axis=1).max()
    assert np.abs(u - free).max() < 10 * (d / h) * np.abs(free).max()


@pytest.mark.parametrize("kernel", [8, 16])
def test_linear_in_forces(kernel):
    rng = np.random.default_rng(6 + kernel)
    x, f = sources(rng, 20, b=0.01 if kernel == 16 else None)
    _, g = sources(rng, 20, b=0.01 if kernel == 16 else None)
    trg = rng.random((15, 3))
    k = 3
    lhs = O.wall(kernel, x, np.concatenate([2 * f[:, :k] - g[:, :k], f[:, k:]], axis=1), trg)
    rhs = 2 * O.wall(kernel, x, f, trg) - O.wall(kernel, x, np.concatenate([g[:, :k], f[:, k:]], axis=1), trg)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-11, atol=1e-12 * np.abs(rhs).max())


def test_wall_api_rejects_other_kernels_and_bad_shapes():
    rng = np.random.default_rng(7)
    x, f = sources(rng, 4)
    trg = rng.random((3, 3))
    with pytest.raises(S.KernelNotFound):
        S.wall_evaluate(S.KERNEL.PVel, x, f, trg, np.zeros((3, 4)))
    with pytest.raises(S.ArgumentError):
        S.wall_evaluate(S.KERNEL.RPY, x, f, trg, np.zeros((3, 6)))   # RPY needs 4 values per source
    # empty inputs are a no-op (no device touched)
    out = np.ones((3, 3))
    S.wall_evaluate(S.KERNEL.Stokes, np.zeros((0, 3)), np.zeros((0, 3)), trg, out)
    assert np.all(out == 1.0)
