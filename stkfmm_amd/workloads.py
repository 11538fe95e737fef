"""Synthetic point sets and source values of the five benchmark workloads (seeded, numpy PCG64).

Seeds are ``1000*workload + stream``.  Coordinates are wrapped into [0,1)^3 with ``v - floor(v)``
as the reference's point generator does (Test/Util/PointDistribution.cpp:143-146).

1. Laplace PGrad direct sum: 8192 sources and 8192 targets i.i.d. U[0,1)^3, charges U[-1,1].
2. Stokes PVel U-list: 1,000,000 points, half U[0,1)^3 and half in 16 Gaussian clusters (sigma 0.01,
   centres U[0,1)^3); single layer f ~ N(0,1)^3 with TrD = 0, double layer 9 x N(0,1); SL = DL =
   target points; leaf cap 1000.
3. RPY: 4,000,000 points in 64 clusters, centre U[0,1)^3, cluster radius R_c log-uniform in
   [1e-3, 5e-2], offset R_c * rho * n (rho ~ LogNormal(0, 0.5), n uniform on the sphere); radius
   b ~ U[1e-4, 5e-4], f ~ N(0,1)^3; leaf cap 2000.
4. PVel adjoint: 65,536 sources and targets i.i.d. U[0,1)^3, f ~ N(0,1)^3 (TrD = 0), weights
   w ~ N(0,1)^4.
5. Stokeslet instances: 512 direct sums, N_i = floor(exp(U(ln 1024, ln 32768))), sources = targets,
   points U[0,1)^3 per instance, f ~ N(0,1)^3.
"""
import numpy as np


def _rng(workload, stream):
    return np.random.default_rng(1000 * workload + stream)


def wrap(x):
    return x - np.floor(x)


def lapgrad_direct(n_src=8192, n_trg=8192):
    """Workload 1: (src_coord [ns,3], charges [ns,1], trg_coord [nt,3])."""
    src = _rng(1, 0).random((n_src, 3))
    trg = _rng(1, 1).random((n_trg, 3))
    q = _rng(1, 2).uniform(-1.0, 1.0, (n_src, 1))
    return src, q, trg


def pvel_points(n=1_000_000, n_clusters=16, sigma=0.01):
    """Workload 2 points: half uniform, half in Gaussian clusters; [n,3] in [0,1)."""
    n_uni = n // 2
    n_cl = n - n_uni
    uni = _rng(2, 0).random((n_uni, 3))
    centres = _rng(2, 1).random((n_clusters, 3))
    which = np.arange(n_cl) % n_clusters
    cl = centres[which] + _rng(2, 2).normal(0.0, sigma, (n_cl, 3))
    return wrap(np.concatenate([uni, cl], axis=0))


def pvel_values(n, trd=False):
    """Workload 2 sources: single layer [n,4] (f ~ N(0,1)^3, TrD = 0 unless trd) and double layer [n,9]."""
    sl = np.zeros((n, 4))
    sl[:, :3] = _rng(2, 3).normal(size=(n, 3))
    if trd:
        sl[:, 3] = _rng(2, 5).normal(size=n)
    dl = _rng(2, 4).normal(size=(n, 9))
    return sl, dl


def rpy_points(n=4_000_000, n_clusters=64, r_lo=1e-3, r_hi=5e-2):
    """Workload 3 points: log-normal clusters; [n,3] in [0,1)."""
    g = _rng(3, 0)
    centres = g.random((n_clusters, 3))
    radii = np.exp(g.uniform(np.log(r_lo), np.log(r_hi), n_clusters))
    which = np.arange(n) % n_clusters
    h = _rng(3, 1)
    rho = h.lognormal(0.0, 0.5, n)
    v = h.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return wrap(centres[which] + (radii[which] * rho)[:, None] * v)


def rpy_values(n):
    """Workload 3 sources [n,4]: f ~ N(0,1)^3, b ~ U[1e-4, 5e-4]."""
    val = np.empty((n, 4))
    val[:, :3] = _rng(3, 2).normal(size=(n, 3))
    val[:, 3] = _rng(3, 3).uniform(1e-4, 5e-4, n)
    return val


def adjoint_pvel(n=65536):
    """Workload 4: (src_coord, src_value [n,4] TrD=0, trg_coord, w [n,4])."""
    src = _rng(4, 0).random((n, 3))
    trg = _rng(4, 1).random((n, 3))
    val = np.zeros((n, 4))
    val[:, :3] = _rng(4, 2).normal(size=(n, 3))
    w = _rng(4, 3).normal(size=(n, 4))
    return src, val, trg, w


def stokeslet_instances(n_inst=512, n_lo=1024, n_hi=32768):
    """Workload 5: (sizes [n_inst], offsets [n_inst+1], coord [N,3], forces [N,3]); instance i owns
    points offsets[i]:offsets[i+1], in its own unit cube."""
    sizes = np.floor(np.exp(_rng(5, 0).uniform(np.log(n_lo), np.log(n_hi), n_inst))).astype(np.int64)
    offsets = np.zeros(n_inst + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    coord = _rng(5, 1).random((int(offsets[-1]), 3))
    f = _rng(5, 2).normal(size=(int(offsets[-1]), 3))
    return sizes, offsets, coord, f
