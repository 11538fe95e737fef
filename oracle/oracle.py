"""ctypes wrapper over the CPU oracle (oracle/liboracle.so) and the reference build (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s cpu_baseline leg.  The product package ``stkfmm_amd`` never imports this module.

The micro-kernel ids follow the numbering of ``include/stkfmm_p2p.h`` (STKFMM_UKER_*); the
reference file:line each one restates is cited in ``oracle/p2p_oracle.c``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libstkfmm_simplekernel.so")

UK = dict(
    LAPLACE_P=0, LAPLACE_DIPOLEP=1, LAPLACE_PGRAD=2, LAPLACE_DIPOLEPGRAD=3, LAPLACE_PGRADGRAD=4,
    LAPLACE_DIPOLEPGRADGRAD=5, STOKES_VEL=6, STOKES_PVEL=7, STOKES_DOUBLEPVEL=8, STOKES_REGVEL=9,
    RPY_U=10, RPY_ULAPU=11, STK_ULAPU=12, STOKES_PVEL_FUSED=13, STOKES_PVELGRAD=14,
    STOKES_DOUBLEPVELGRAD=15, STOKES_PVELLAPLACIAN=16, STOKES_DOUBLELAPLACIAN=17, STOKES_TRACTION=18,
    STOKES_DOUBLETRACTION=19, STOKES_REGFTVELOMEGA=20, STOKES_VELOMEGA=21, LAPLACE_QUADPGRADGRAD=22,
    LAPLACE_QUADP=23, STOKES_REGFTVEL=24,
)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_uker_scale.restype = ctypes.c_double
        L.oracle_uker_scale.argtypes = [ctypes.c_int]
        L.oracle_uker_src_dim.argtypes = [ctypes.c_int]
        L.oracle_uker_trg_dim.argtypes = [ctypes.c_int]
        L.oracle_uker_pair.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_direct.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.oracle_direct_chunked.argtypes = L.oracle_direct.argtypes
        rargs = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, ctypes.c_int, _ip,
                 ctypes.c_int]
        L.oracle_ranges.argtypes = rargs
        L.oracle_ranges_simd.argtypes = rargs
        L.oracle_vjp_pair.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_direct_vjp.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                        _dp, ctypes.c_int]
        L.oracle_pair_longdouble.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_wall.argtypes = [ctypes.c_uint, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_LIB_PATH)


def ref():
    """The reference's own scalar oracle (Test/SimpleKernel.cpp) built into oracle/_ref."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference oracle not built: {REF_LIB_PATH}")
        L = ctypes.CDLL(REF_LIB_PATH)
        L.simplekernel_available.argtypes = [ctypes.c_uint, ctypes.c_int]
        L.simplekernel_threads.argtypes = [ctypes.c_int]
        L.simplekernel_pair.argtypes = [ctypes.c_uint, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.simplekernel_direct.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _dp, _dp, ctypes.c_int, _dp, _dp]
        _ref = L
    return _ref


def src_dim(uk):
    return lib().oracle_uker_src_dim(uk)


def trg_dim(uk):
    return lib().oracle_uker_trg_dim(uk)


def scale(uk):
    return lib().oracle_uker_scale(uk)


def pair(uk, r, f):
    """One pair without the scale factor (reference-order restatement)."""
    r = _f64(r)
    f = _f64(f)
    u = np.zeros(trg_dim(uk))
    assert lib().oracle_uker_pair(uk, _d(r), _d(f), _d(u)) == 0
    return u


def pair_longdouble(uk, r, f):
    r = _f64(r)
    f = _f64(f)
    u = np.zeros(trg_dim(uk))
    assert lib().oracle_pair_longdouble(uk, _d(r), _d(f), _d(u)) == 0
    return u


def direct(uk, src_coord, src_value, trg_coord, trg_value=None, nthreads=0, chunked=False):
    """trg_value += direct sum (AoS arrays); returns trg_value."""
    sc = _f64(src_coord).reshape(-1)
    sv = _f64(src_value).reshape(-1)
    tc = _f64(trg_coord).reshape(-1)
    ns, nt = sc.size // 3, tc.size // 3
    kt = trg_dim(uk)
    out = np.zeros(nt * kt) if trg_value is None else _f64(trg_value).reshape(-1).copy()
    fn = lib().oracle_direct_chunked if chunked else lib().oracle_direct
    assert fn(uk, ns, _d(sc), _d(sv), nt, _d(tc), _d(out), nthreads) == 0
    return out.reshape(nt, kt)


def wall(kernel, src_coord, src_value, trg_coord, trg_value=None, nthreads=0):
    """StkWallFMM's image system with direct sums (kernel 8 Stokes -> 3, 16 RPY -> 6); trg_value +=."""
    sc = _f64(src_coord).reshape(-1)
    sv = _f64(src_value).reshape(-1)
    tc = _f64(trg_coord).reshape(-1)
    ns, nt = sc.size // 3, tc.size // 3
    kt = 3 if int(kernel) == 8 else 6
    out = np.zeros(nt * kt) if trg_value is None else _f64(trg_value).reshape(-1).copy()
    assert lib().oracle_wall(int(kernel), ns, _d(sc), _d(sv), nt, _d(tc), _d(out), nthreads) == 0
    return out.reshape(nt, kt)


def ranges(uk, src_coord, src_value, trg_coord, grp_trg_begin, grp_trg_end, grp_rng_ptr, rng_src_begin,
           rng_src_end, sel=None, nthreads=0, simd=False, trg_value=None):
    """Range-list (U-list / batch) evaluation over leaf-sorted AoS arrays; returns trg_value (+=)."""
    sc = _f64(src_coord).reshape(-1)
    sv = _f64(src_value).reshape(-1)
    tc = _f64(trg_coord).reshape(-1)
    nt = tc.size // 3
    kt = trg_dim(uk)
    out = np.zeros(nt * kt) if trg_value is None else _f64(trg_value).reshape(-1).copy()
    gb, ge, gp, rb, re = map(_i32, (grp_trg_begin, grp_trg_end, grp_rng_ptr, rng_src_begin, rng_src_end))
    s = _i32(sel)
    fn = lib().oracle_ranges_simd if simd else lib().oracle_ranges
    rc = fn(uk, _d(sc), _d(sv), _d(tc), _d(out), len(gb), _i(gb), _i(ge), _i(gp), _i(rb), _i(re),
            0 if s is None else len(s), _i(s), nthreads)
    if rc != 0:
        raise ValueError(f"oracle ranges rc={rc}")
    return out.reshape(nt, kt)


def vjp_pair(uk, r, f, w):
    r, f, w = _f64(r), _f64(f), _f64(w)
    gr = np.zeros(3)
    gf = np.zeros(src_dim(uk))
    assert lib().oracle_vjp_pair(uk, _d(r), _d(f), _d(w), _d(gr), _d(gf)) == 0
    return gr, gf


def direct_vjp(uk, src_coord, src_value, trg_coord, w, nthreads=0):
    """Gradients of L = sum_t w_t . u_t: (g_src_value, g_src_coord, g_trg_coord)."""
    sc = _f64(src_coord).reshape(-1)
    sv = _f64(src_value).reshape(-1)
    tc = _f64(trg_coord).reshape(-1)
    w = _f64(w).reshape(-1)
    ns, nt = sc.size // 3, tc.size // 3
    ks = src_dim(uk)
    gsv = np.zeros(ns * ks)
    gsc = np.zeros(ns * 3)
    gtc = np.zeros(nt * 3)
    rc = lib().oracle_direct_vjp(uk, ns, _d(sc), _d(sv), nt, _d(tc), _d(w), _d(gsv), _d(gsc), _d(gtc), nthreads)
    if rc != 0:
        raise ValueError(f"oracle vjp rc={rc}")
    return gsv.reshape(ns, ks), gsc.reshape(ns, 3), gtc.reshape(nt, 3)


def simplekernel_direct(kernel, layer, ks, kt, src_coord, src_value, trg_coord):
    """Reference Test/SimpleKernel.cpp direct sum (runSimpleKernel accumulation), fresh output."""
    sc = _f64(src_coord).reshape(-1)
    sv = _f64(src_value).reshape(-1)
    tc = _f64(trg_coord).reshape(-1)
    ns, nt = sc.size // 3, tc.size // 3
    out = np.zeros(nt * kt)
    rc = ref().simplekernel_direct(kernel, layer, ks, kt, ns, _d(sc), _d(sv), nt, _d(tc), _d(out))
    if rc != 0:
        raise ValueError("simplekernel unavailable for this kernel/layer")
    return out.reshape(nt, kt)


def simplekernel_threads(n=0):
    """Set (n > 0) and return the OpenMP thread count of the reference direct sum."""
    return ref().simplekernel_threads(int(n))


def simplekernel_pair(kernel, layer, s, t, f):
    s, t = _f64(s), _f64(t)
    ff = np.zeros(16)
    f = _f64(f).reshape(-1)
    ff[: f.size] = f
    out = np.zeros(20)
    assert ref().simplekernel_pair(kernel, layer, _d(s), _d(t), _d(ff), _d(out)) == 0
    return out


def error_metrics(value, true):
    """Per-component error as the reference's ComponentError (Test/Test.cpp:113-151).

    value, true: arrays [n, k].  Returns a list of dicts with errorL2, errorRMS, errorMaxRel, drift.
    """
    value = np.asarray(value, dtype=np.float64).reshape(len(true), -1)
    true = np.asarray(true, dtype=np.float64).reshape(len(value), -1)
    out = []
    n = value.shape[0]
    for k in range(value.shape[1]):
        a, b = value[:, k], true[:, k]
        l2t = float(np.sum(b * b))
        maxt = float(np.max(np.abs(b))) if n else 0.0
        err = a - b
        l2e = float(np.sum(err * err))
        out.append(dict(
            errorRMS=float(np.sqrt(l2e) / np.sqrt(n)) if n else 0.0,
            errorL2=float(np.sqrt(l2e) / np.sqrt(l2t)) if l2t > 0 else float(np.sqrt(l2e)),
            errorMaxRel=float(np.max(np.abs(err)) / maxt) if maxt > 0 else float(np.max(np.abs(err), initial=0)),
            drift=float(np.sum(err) / n) if n else 0.0,
        ))
    return out


def max_error(value, true):
    """(max over components of errorL2, max over components of errorMaxRel)."""
    m = error_metrics(value, true)
    return max(e["errorL2"] for e in m), max(e["errorMaxRel"] for e in m)
