"""Pin the oracle's restatements of the rest of the kernel family beyond the golden fixtures.

PVelGrad, PVelLaplacian, Traction (single and double layer), StokesRegVelOmega and LapQPGradGrad
are checked against the reference's own SimpleKernel in test_oracle_golden.py.  This file adds what
the mathematics fixes (so that a slip shared by a restatement and a fixture would still show), and
pins the three sub-kernels SimpleKernel has no entry for (stokes_velomega, stokes_regftvel,
laplace_quadp) by identities with pinned kernels:

  * PVelGrad: first four outputs = PVel; grad p and grad u = finite differences of p and u;
    trace(grad u) = 0;
  * PVelLaplacian: first four outputs = PVel; lap u = FD Laplacian of u = grad p (Stokes);
  * Traction = -p I + grad u + grad u^T from the PVelGrad outputs (the Newtonian stress);
  * StokesRegVelOmega: omega = curl(u)/2 by FD; at t = 0 the velocity is the regularised
    Stokeslet stokes_regvel; at eps = 0, t = 0 it is stokes_velomega;
  * stokes_velomega: velocity = Stokeslet, omega = curl(u)/2;
  * stokes_regftvel = velocity of stokes_regftvelomega (same expression, bit for bit);
  * LapQPGradGrad: potential = laplace_quadp = sum_ij Q_ij d_i d_j (1/r) (the Hessian of the
    pinned laplace_pgradgrad); gradient and Hessian = FD; Hessian trace = 0.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_invariants import fd_grad, fd_laplacian, field, random_pair

UK = O.UK


def _curl(J):
    return np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])


@pytest.mark.parametrize("uk,base,ks", [(UK["STOKES_PVELGRAD"], UK["STOKES_PVEL"], 4),
                                        (UK["STOKES_DOUBLEPVELGRAD"], UK["STOKES_DOUBLEPVEL"], 9)])
def test_pvelgrad_is_fd_gradient_of_pvel(uk, base, ks):
    rng = np.random.default_rng(300 + uk)
    for _ in range(8):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, ks)
        u = field(uk, s, f, t)
        np.testing.assert_allclose(u[:4], field(base, s, f, t), rtol=1e-13, atol=1e-14 * np.abs(u[:4]).max())
        gp = fd_grad(lambda x: field(base, s, f, x)[:1], t, 1e-6)[0]
        np.testing.assert_allclose(u[4:7], gp, rtol=1e-6, atol=1e-7 * np.abs(gp).max())
        J = fd_grad(lambda x: field(base, s, f, x)[1:4], t, 1e-6)   # J[i, j] = du_i/dx_j
        np.testing.assert_allclose(u[7:16], J.ravel(), rtol=1e-6, atol=1e-7 * np.abs(J).max())
        assert abs(u[7] + u[11] + u[15]) < 1e-13 * np.abs(u[7:16]).max()


@pytest.mark.parametrize("uk,grad,base,ks", [
    (UK["STOKES_PVELLAPLACIAN"], UK["STOKES_PVELGRAD"], UK["STOKES_PVEL"], 4),
    (UK["STOKES_DOUBLELAPLACIAN"], UK["STOKES_DOUBLEPVELGRAD"], UK["STOKES_DOUBLEPVEL"], 9)])
def test_pvellaplacian_is_fd_laplacian_and_grad_p(uk, grad, base, ks):
    rng = np.random.default_rng(310 + uk)
    for _ in range(8):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, ks)
        u = field(uk, s, f, t)
        np.testing.assert_allclose(u[:4], field(base, s, f, t), rtol=1e-13, atol=1e-14 * np.abs(u[:4]).max())
        lap = fd_laplacian(lambda x: field(base, s, f, x)[1:4], t, 1e-4)
        np.testing.assert_allclose(u[4:7], lap, rtol=2e-5, atol=2e-5 * np.abs(lap).max())
        # Stokes' equation: lap u = grad p, the latter from the gradient kernel
        gp = field(grad, s, f, t)[4:7]
        np.testing.assert_allclose(u[4:7], gp, rtol=1e-13, atol=1e-13 * np.abs(gp).max())


@pytest.mark.parametrize("uk,grad,ks", [(UK["STOKES_TRACTION"], UK["STOKES_PVELGRAD"], 4),
                                        (UK["STOKES_DOUBLETRACTION"], UK["STOKES_DOUBLEPVELGRAD"], 9)])
def test_traction_is_newtonian_stress_of_pvelgrad(uk, grad, ks):
    rng = np.random.default_rng(320 + uk)
    for _ in range(20):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, ks)
        g = field(grad, s, f, t)
        J = g[7:16].reshape(3, 3)
        sigma = -g[0] * np.eye(3) + J + J.T
        tr = field(uk, s, f, t)
        np.testing.assert_allclose(tr, sigma.ravel(), rtol=1e-12, atol=1e-13 * np.abs(sigma).max())


def test_regftvelomega_omega_is_half_curl_of_velocity():
    uk = UK["STOKES_REGFTVELOMEGA"]
    rng = np.random.default_rng(330)
    for eps in (0.0, 0.02, 0.3):
        for _ in range(6):
            s, t = random_pair(rng)
            f = np.concatenate([rng.uniform(-1, 1, 6), [eps]])
            u = field(uk, s, f, t)
            J = fd_grad(lambda x: field(uk, s, f, x)[:3], t, 1e-6)
            np.testing.assert_allclose(u[3:], 0.5 * _curl(J), rtol=1e-6, atol=1e-7 * np.abs(u[3:]).max())


def test_regftvelomega_reduces_to_regvel_and_velomega():
    uk = UK["STOKES_REGFTVELOMEGA"]
    rng = np.random.default_rng(331)
    for _ in range(20):
        s, t = random_pair(rng)
        fv = rng.uniform(-1, 1, 3)
        eps = abs(rng.uniform(-1, 1)) * 0.05
        u = field(uk, s, np.concatenate([fv, [0.0, 0.0, 0.0, eps]]), t)
        reg = field(UK["STOKES_REGVEL"], s, np.concatenate([fv, [eps]]), t)
        np.testing.assert_allclose(u[:3], reg, rtol=1e-13, atol=1e-14 * np.abs(reg).max())
        u0 = field(uk, s, np.concatenate([fv, [0.0, 0.0, 0.0, 0.0]]), t)
        vo = field(UK["STOKES_VELOMEGA"], s, fv, t)
        np.testing.assert_allclose(u0, vo, rtol=1e-13, atol=1e-14 * np.abs(vo).max())


def test_velomega_is_stokeslet_and_half_curl():
    uk = UK["STOKES_VELOMEGA"]
    rng = np.random.default_rng(332)
    for _ in range(8):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, 3)
        u = field(uk, s, f, t)
        np.testing.assert_allclose(u[:3], field(UK["STOKES_VEL"], s, f, t), rtol=1e-13, atol=1e-15)
        J = fd_grad(lambda x: field(uk, s, f, x)[:3], t, 1e-6)
        np.testing.assert_allclose(u[3:], 0.5 * _curl(J), rtol=1e-6, atol=1e-7 * np.abs(u[3:]).max())


def test_regftvel_is_regftvelomega_velocity_bitwise():
    rng = np.random.default_rng(333)
    src = rng.random((50, 3))
    trg = np.concatenate([src[:5], rng.random((40, 3))])
    v = rng.uniform(-1, 1, (50, 7))
    v[:, 6] = np.abs(v[:, 6]) * 0.01
    a = O.direct(UK["STOKES_REGFTVEL"], src, v, trg)
    b = O.direct(UK["STOKES_REGFTVELOMEGA"], src, v, trg)[:, :3]
    assert np.array_equal(a, b)


def test_quadp_is_quadrupole_contraction_of_laplace_hessian():
    rng = np.random.default_rng(334)
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for _ in range(20):
        s, t = random_pair(rng)
        Q = rng.uniform(-1, 1, 9)
        H1 = field(UK["LAPLACE_PGRADGRAD"], s, [1.0], t)
        H = np.zeros((3, 3))
        for (i, j), k in idx.items():
            H[i, j] = H[j, i] = H1[k]
        expect = float((Q.reshape(3, 3) * H).sum())
        p = field(UK["LAPLACE_QUADP"], s, Q, t)[0]
        assert abs(p - expect) < 1e-13 * np.abs(Q.reshape(3, 3) * H).sum()
        np.testing.assert_allclose(field(UK["LAPLACE_QUADPGRADGRAD"], s, Q, t)[0], p, rtol=1e-13, atol=1e-15)


def test_quadpgradgrad_gradient_hessian_are_fd_and_traceless():
    uk = UK["LAPLACE_QUADPGRADGRAD"]
    rng = np.random.default_rng(335)
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for _ in range(8):
        s, t = random_pair(rng)
        Q = rng.uniform(-1, 1, 9)
        u = field(uk, s, Q, t)
        g = fd_grad(lambda x: field(uk, s, Q, x)[:1], t, 1e-6)[0]
        np.testing.assert_allclose(u[1:4], g, rtol=1e-6, atol=1e-7 * np.abs(g).max())
        H = fd_grad(lambda x: field(uk, s, Q, x)[1:4], t, 1e-6)
        scale = np.abs(u[4:]).max()
        for (i, j), k in idx.items():
            assert abs(u[k] - H[i, j]) < 1e-6 * scale
        assert abs(u[4] + u[7] + u[9]) < 1e-12 * scale


@pytest.mark.parametrize("uk", [UK["STOKES_PVELGRAD"], UK["STOKES_DOUBLEPVELGRAD"], UK["STOKES_PVELLAPLACIAN"],
                                UK["STOKES_DOUBLELAPLACIAN"], UK["STOKES_TRACTION"], UK["STOKES_DOUBLETRACTION"],
                                UK["STOKES_VELOMEGA"], UK["LAPLACE_QUADPGRADGRAD"], UK["LAPLACE_QUADP"]])
def test_new_kernels_linear_and_translation_invariant(uk):
    rng = np.random.default_rng(340 + uk)
    ks = O.src_dim(uk)
    src = rng.random((30, 3))
    trg = rng.random((20, 3))
    a = rng.uniform(-1, 1, (30, ks))
    b = rng.uniform(-1, 1, (30, ks))
    lhs = O.direct(uk, src, 2.0 * a - 0.5 * b, trg)
    rhs = 2.0 * O.direct(uk, src, a, trg) - 0.5 * O.direct(uk, src, b, trg)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-11, atol=1e-12 * np.abs(rhs).max())
    off = np.array([0.25, -0.5, 0.125])
    np.testing.assert_allclose(O.direct(uk, src + off, a, trg + off), O.direct(uk, src, a, trg), rtol=1e-10,
                               atol=1e-11 * np.abs(rhs).max())


@pytest.mark.parametrize("uk", range(25))
def test_coincident_pair_is_zero_all_kernels(uk):
    f = np.ones(O.src_dim(uk))
    if uk in (9, 20, 24):   # regularised: the mask is on |d|^2 + eps^2
        f[-1] = 0.0
    assert np.all(O.pair(uk, np.zeros(3), f) == 0.0)
