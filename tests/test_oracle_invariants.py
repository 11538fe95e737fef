"""Pin the oracle with what the mathematics fixes, independently of its own formulas.

Each check would fail for a plausible slip in the restatement (a dropped term, a wrong sign, a wrong
index or a transposed double-layer tensor):
  * long-double closed forms on single pairs;
  * Laplace: gradient = finite-difference gradient of the potential, Hessian = FD of the gradient,
    Hessian trace = 0 (harmonic);
  * Stokes (Stokeslet, PVel single and double layer): div u = 0 and Stokes' equation lap u = grad p
    by finite differences;
  * RPY: lap u output = FD Laplacian of u, and u + (a^2/6) lap u equals the Rotne-Prager-Yamakawa
    mobility tensor of Python/rpy_test.py:10-56 (an independent formula) for r > 2a;
  * regularised Stokeslet at eps = 0 equals the Stokeslet;
  * linearity in strengths, translation invariance;
  * adjoints (hand-derived VJPs) against central finite differences of the forward.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import source_values


def field(uk, s, f, t):
    return O.direct(uk, np.asarray(s)[None], np.asarray(f)[None], np.asarray(t)[None])[0]


def fd_grad(fun, t, h):
    """d fun / d t_j, central differences: returns [len(fun), 3]."""
    cols = []
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        cols.append((fun(t + e) - fun(t - e)) / (2 * h))
    return np.stack(cols, axis=1)


def fd_laplacian(fun, t, h):
    acc = -6.0 * fun(t)
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        acc = acc + fun(t + e) + fun(t - e)
    return acc / (h * h)


def random_pair(rng):
    s = rng.random(3)
    t = s + (rng.random(3) - 0.5) * 0.6 + 0.1
    return s, t


@pytest.mark.parametrize("uk", [2, 7, 8, 11])
def test_longdouble_closed_form(uk):
    rng = np.random.default_rng(10 + uk)
    for _ in range(50):
        r = rng.uniform(-1, 1, 3) * 10 ** rng.uniform(-3, 0)
        f = source_values(uk, 1, rng)[0]
        a = O.pair(uk, r, f)
        b = O.pair_longdouble(uk, r, f)
        np.testing.assert_allclose(a, b, rtol=2e-14, atol=2e-14 * np.abs(b).max())


@pytest.mark.parametrize("uk,layer_dim", [(2, 1), (3, 3)])
def test_laplace_gradient_is_fd_of_potential(uk, layer_dim):
    rng = np.random.default_rng(20 + uk)
    for _ in range(10):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, layer_dim)
        u = field(uk, s, f, t)
        g = fd_grad(lambda x: field(uk, s, f, x)[:1], t, 1e-6)[0]
        np.testing.assert_allclose(u[1:4], g, rtol=1e-6, atol=1e-7 * np.abs(u[1:4]).max())


@pytest.mark.parametrize("uk,layer_dim", [(4, 1), (5, 3)])
def test_laplace_hessian_is_fd_of_gradient_and_traceless(uk, layer_dim):
    rng = np.random.default_rng(30 + uk)
    idx = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for _ in range(10):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, layer_dim)
        u = field(uk, s, f, t)
        # first four entries agree with the PGrad kernel of the same layer
        np.testing.assert_allclose(u[:4], field(uk - 2, s, f, t), rtol=1e-13, atol=1e-15)
        H = fd_grad(lambda x: field(uk, s, f, x)[1:4], t, 1e-6)
        scale = np.abs(u[4:]).max()
        for (i, j), k in idx.items():
            assert abs(u[k] - H[i, j]) < 1e-6 * scale, (i, j)
            assert abs(u[k] - H[j, i]) < 1e-6 * scale, (j, i)
        assert abs(u[4] + u[7] + u[9]) < 1e-13 * scale


def _velocity_fn(uk, s, f, vel_slice):
    return lambda x: field(uk, s, f, x)[vel_slice]


@pytest.mark.parametrize("uk,ks,vel,pre", [(6, 3, slice(0, 3), None), (7, 4, slice(1, 4), 0), (8, 9, slice(1, 4), 0),
                                          (12, 3, slice(0, 3), None)])
def test_stokes_incompressible_and_stokes_equation(uk, ks, vel, pre):
    rng = np.random.default_rng(40 + uk)
    for _ in range(8):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, ks)
        u = field(uk, s, f, t)
        J = fd_grad(_velocity_fn(uk, s, f, vel), t, 1e-6)
        scale = np.abs(J).max()
        assert abs(np.trace(J)) < 1e-7 * scale
        if pre is not None:
            lap = fd_laplacian(_velocity_fn(uk, s, f, vel), t, 1e-4)
            gp = fd_grad(lambda x: field(uk, s, f, x)[pre:pre + 1], t, 1e-6)[0]
            np.testing.assert_allclose(lap, gp, rtol=2e-5, atol=2e-5 * np.abs(gp).max())
        del u


def test_stokeslet_laplacian_output_is_fd_laplacian():
    # stk_ulapu: outputs 3..5 are the Laplacian of outputs 0..2 (RPYKernel.hpp:97-129)
    rng = np.random.default_rng(50)
    for _ in range(8):
        s, t = random_pair(rng)
        f = rng.uniform(-1, 1, 3)
        u = field(12, s, f, t)
        lap = fd_laplacian(lambda x: field(12, s, f, x)[:3], t, 1e-4)
        np.testing.assert_allclose(u[3:], lap, rtol=2e-5, atol=2e-5 * np.abs(u[3:]).max())


def test_rpy_laplacian_output_is_fd_laplacian():
    rng = np.random.default_rng(51)
    for _ in range(8):
        s, t = random_pair(rng)
        f = np.concatenate([rng.uniform(-1, 1, 3), [0.05]])
        u = field(11, s, f, t)
        lap = fd_laplacian(lambda x: field(11, s, f, x)[:3], t, 1e-4)
        np.testing.assert_allclose(u[3:], lap, rtol=2e-5, atol=2e-5 * np.abs(u[3:]).max())


def test_rpy_matches_rotne_prager_mobility():
    # Python/rpy_test.py:10-56,124: for r > 2a, u(b=a) + a^2/6 lap u = M_RPY(r) F with eta = 1
    rng = np.random.default_rng(52)
    a = 0.02
    n = 0
    while n < 40:
        s, t = random_pair(rng)
        r = t - s
        d = np.linalg.norm(r)
        if d <= 2 * a:
            continue
        n += 1
        F = rng.normal(size=3)
        u = field(11, s, np.concatenate([F, [a]]), t)
        mf = u[:3] + a * a * u[3:] / 6.0
        factor = 1.0 / (6.0 * np.pi)
        fr = factor * (0.75 / d + a * a / (2.0 * d ** 3))
        gr = factor * (0.75 / d ** 3 - 1.5 * a * a / d ** 5)
        M = fr * np.eye(3) + gr * np.outer(r, r)
        np.testing.assert_allclose(mf, M @ F, rtol=1e-12, atol=1e-14)


def test_regvel_zero_eps_is_stokeslet():
    rng = np.random.default_rng(53)
    src = rng.random((60, 3))
    trg = rng.random((40, 3))
    f = rng.uniform(-1, 1, (60, 3))
    v = np.concatenate([f, np.zeros((60, 1))], axis=1)
    np.testing.assert_allclose(O.direct(9, src, v, trg), O.direct(6, src, f, trg), rtol=1e-13, atol=1e-14)


def test_regvel_regularised_self_term_is_finite():
    # the mask is on |d|^2 + eps^2 (StokesRegSingleLayerKernel.hpp:30-32): eps > 0 gives a finite self term
    f = np.array([1.0, -2.0, 0.5, 0.1])
    u = O.pair(9, np.zeros(3), f)
    np.testing.assert_allclose(u, f[:3] * 2 * 0.01 / 0.1 ** 3, rtol=1e-14)


@pytest.mark.parametrize("uk", [0, 1, 2, 3, 4, 5, 6, 7, 8, 12, 13])
def test_linearity_in_strengths(uk):
    rng = np.random.default_rng(60 + uk)
    src = rng.random((80, 3))
    trg = np.concatenate([src[:4], rng.random((50, 3))])
    ks = O.src_dim(uk)
    v1, v2 = rng.uniform(-1, 1, (80, ks)), rng.uniform(-1, 1, (80, ks))
    a, b = 0.7, -1.3
    lhs = O.direct(uk, src, a * v1 + b * v2, trg)
    rhs = a * O.direct(uk, src, v1, trg) + b * O.direct(uk, src, v2, trg)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-12 * np.abs(rhs).max())


@pytest.mark.parametrize("uk", [2, 7, 8, 9, 11, 13])
def test_translation_invariance(uk):
    rng = np.random.default_rng(70 + uk)
    src = rng.random((80, 3))
    trg = rng.random((50, 3))
    v = source_values(uk, 80, rng)
    c = np.array([0.25, -0.5, 0.125])
    ref = O.direct(uk, src, v, trg)
    np.testing.assert_allclose(O.direct(uk, src + c, v, trg + c), ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


ADJOINT_UKERS = [0, 2, 6, 7, 8, 9, 10, 11, 12, 13]


@pytest.mark.parametrize("uk", ADJOINT_UKERS)
def test_adjoint_matches_finite_differences(uk):
    rng = np.random.default_rng(80 + uk)
    ns, nt = 5, 4
    src = rng.random((ns, 3))
    trg = rng.random((nt, 3))  # no coincident pair: finite differences across r = 0 are meaningless
    v = source_values(uk, ns, rng)
    w = rng.normal(size=(nt, O.trg_dim(uk)))
    gsv, gsc, gtc = O.direct_vjp(uk, src, v, trg, w)

    def L(s, vv, t):
        return float(np.sum(w * O.direct(uk, s, vv, t)))

    h = 1e-6
    for arr, grad, name in ((src, gsc, "src"), (v, gsv, "val"), (trg, gtc, "trg")):
        fd = np.zeros_like(arr)
        for idx in np.ndindex(arr.shape):
            p, m = arr.copy(), arr.copy()
            p[idx] += h
            m[idx] -= h
            args_p = {"src": (p, v, trg), "val": (src, p, trg), "trg": (src, v, p)}[name]
            args_m = {"src": (m, v, trg), "val": (src, m, trg), "trg": (src, v, m)}[name]
            fd[idx] = (L(*args_p) - L(*args_m)) / (2 * h)
        np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-5 * np.abs(fd).max(), err_msg=name)


@pytest.mark.parametrize("uk", [u for u in ADJOINT_UKERS if u != 9])
def test_adjoint_of_coincident_pair_is_zero(uk):
    # (the regularised Stokeslet is finite at r = 0: its eps > 0 keeps rho > 0)
    rng = np.random.default_rng(90 + uk)
    gr, gf = O.vjp_pair(uk, np.zeros(3), source_values(uk, 1, rng)[0], rng.normal(size=O.trg_dim(uk)))
    assert np.all(gr == 0.0) and np.all(gf == 0.0)
