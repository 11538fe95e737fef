"""stkfmm_amd -- MI355X-native (gfx950 HIP) P2P near-field kernels of STKFMM.

Host-side mirror of the reference's P2P interface over the C ABI of ``libstkfmm_p2p.so``
(``include/stkfmm_p2p.h``):

* enums ``PAXIS``, ``KERNEL`` (as ``Python/PySTKFMM.py:11-31``), ``PPKERNEL``
  (``Lib/include/STKFMM/STKFMM_common.hpp:37-41``), plus ``PREC`` and ``UKER``;
* ``get_kernel_dimension`` / ``get_kernel_name`` (``Lib/src/STKFMM.cpp:33-46``);
* ``evaluate_kernel`` and ``Stk3DFMM.evaluate_kernel`` -- ``STKFMM::evaluateKernel``
  (``Lib/include/STKFMM/STKFMM.hpp:108-110``): host numpy AoS arrays, results added to ``trg_value``;
* device entry points on torch tensors (``p2p_device``, ``uker_device``, ``vjp_device``), the
  adaptive octree / U-list builder (``Tree``) and batched work-list plans (``Plan``).

Every compute call goes to the HIP library; there is no CPU fallback.  torch is only used for device
memory and streams.
"""
import ctypes
import enum

import numpy as np

from . import _capi
from ._capi import StkArray, StkOutArray, StkTune


class PAXIS(enum.IntEnum):
    NONE = 0
    PX = 1
    PXY = 2
    PXYZ = 3


class KERNEL(enum.IntFlag):
    LapPGrad = 1
    LapPGradGrad = 2
    LapQPGradGrad = 4
    Stokes = 8
    RPY = 16
    StokesRegVel = 32
    StokesRegVelOmega = 64
    PVel = 128
    PVelGrad = 256
    PVelLaplacian = 512
    Traction = 1024


class PPKERNEL(enum.IntEnum):
    SLS2T = 1
    DLS2T = 2
    L2T = 4


class PREC(enum.IntEnum):
    FP64 = 0
    MIXED = 1


class UKER(enum.IntEnum):
    LAPLACE_P = 0
    LAPLACE_DIPOLEP = 1
    LAPLACE_PGRAD = 2
    LAPLACE_DIPOLEPGRAD = 3
    LAPLACE_PGRADGRAD = 4
    LAPLACE_DIPOLEPGRADGRAD = 5
    STOKES_VEL = 6
    STOKES_PVEL = 7
    STOKES_DOUBLEPVEL = 8
    STOKES_REGVEL = 9
    RPY_U = 10
    RPY_ULAPU = 11
    STK_ULAPU = 12
    STOKES_PVEL_FUSED = 13
    STOKES_PVELGRAD = 14
    STOKES_DOUBLEPVELGRAD = 15
    STOKES_PVELLAPLACIAN = 16
    STOKES_DOUBLELAPLACIAN = 17
    STOKES_TRACTION = 18
    STOKES_DOUBLETRACTION = 19
    STOKES_REGFTVELOMEGA = 20
    STOKES_VELOMEGA = 21
    LAPLACE_QUADPGRADGRAD = 22
    LAPLACE_QUADP = 23
    STOKES_REGFTVEL = 24


# ---- errors ---------------------------------------------------------------------------------------

class STKFMMError(RuntimeError):
    code = None


class KernelNotFound(STKFMMError):
    """Unknown KERNEL (the reference prints and calls exit(1), Lib/src/STKFMM.cpp:53-55)."""
    code = -1


class PPKernelNotFound(STKFMMError):
    """Sub-kernel absent (the reference prints and returns, Lib/src/FMMData.cpp:270-273)."""
    code = -2


class ArgumentError(STKFMMError, ValueError):
    code = -3


class HIPError(STKFMMError):
    code = -4


class Unsupported(STKFMMError):
    code = -5


class KernelNotActive(STKFMMError):
    """Kernel not active in the object (the reference prints and calls exit(1), STKFMM.cpp:92-95)."""
    code = -6


_ERRORS = {c.code: c for c in (KernelNotFound, PPKernelNotFound, ArgumentError, HIPError, Unsupported,
                               KernelNotActive)}


def _check(rc):
    if rc != 0:
        raise _ERRORS.get(rc, STKFMMError)(f"[{rc}] {_capi.last_error()}")


def lib():
    return _capi.lib()


# ---- kernel table ---------------------------------------------------------------------------------

def get_kernel_dimension(kernel):
    """[kdimSL, kdimDL, kdimTrg] (Lib/src/STKFMM.cpp:33-40)."""
    dims = np.zeros(3, dtype=np.int32)
    _check(lib().stkfmm_get_kernel_dimension(int(kernel), dims.ctypes.data_as(_capi.iptr)))
    return dims


def get_kernel_name(kernel):
    buf = ctypes.create_string_buffer(64)
    _check(lib().stkfmm_get_kernel_name(int(kernel), buf, 64))
    return buf.value.decode()


def get_uker(kernel, p2p):
    """Micro-kernel evaluated for (KERNEL, PPKERNEL) (Lib/src/FMMData.cpp:261-268)."""
    uk = ctypes.c_int(-1)
    _check(lib().stkfmm_get_uker(int(kernel), int(p2p), ctypes.byref(uk)))
    return UKER(uk.value)


def get_s2m_uker(kernel, double_layer=False):
    """Micro-kernel k_s2m of a KERNEL: sources -> upward check surface (the BuildKernel lists of
    LaplaceLayerKernel.hpp:315-355, StokesLayerKernel.hpp:38-88, RPYKernel.hpp:136-152,
    StokesRegSingleLayerKernel.hpp:211-232)."""
    uk = ctypes.c_int(-1)
    _check(lib().stkfmm_get_s2m_uker(int(kernel), int(bool(double_layer)), ctypes.byref(uk)))
    return UKER(uk.value)


def get_s2l_uker(kernel, double_layer=False):
    """Micro-kernel k_s2l of a KERNEL: sources -> downward check surface (the X-list kernel)."""
    uk = ctypes.c_int(-1)
    _check(lib().stkfmm_get_s2l_uker(int(kernel), int(bool(double_layer)), ctypes.byref(uk)))
    return UKER(uk.value)


def get_m2t_uker(kernel):
    """Micro-kernel k_m2t of a KERNEL: upward equivalent densities -> targets (the W-list kernel)."""
    uk = ctypes.c_int(-1)
    _check(lib().stkfmm_get_m2t_uker(int(kernel), ctypes.byref(uk)))
    return UKER(uk.value)


def uker_dims(uker):
    """(source dim, target dim) of a micro-kernel."""
    ks, kt = ctypes.c_int(), ctypes.c_int()
    _check(lib().stkfmm_uker_dims(int(uker), ctypes.byref(ks), ctypes.byref(kt)))
    return ks.value, kt.value


def device_info():
    n_cu, ws = ctypes.c_int(), ctypes.c_int()
    arch = ctypes.create_string_buffer(64)
    _check(lib().stkfmm_device_info(ctypes.byref(n_cu), ctypes.byref(ws), arch, 64))
    return dict(n_cu=n_cu.value, warp_size=ws.value, arch=arch.value.decode())


# ---- host (numpy) entry: STKFMM::evaluateKernel ------------------------------------------------------

def _host_arr(a, n, k, name, writable=False):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ArgumentError(f"{name}: expected a C-contiguous float64 numpy array")
    if writable and not a.flags.writeable:
        raise ArgumentError(f"{name}: array is read-only")
    if a.size != n * k:
        raise ArgumentError(f"{name}: expected {n}x{k} values, got {a.size}")
    return a.ctypes.data_as(_capi.dptr)


def evaluate_kernel(kernel, n_threads, p2p, src_coord, src_value, trg_coord, trg_value, precision=PREC.FP64):
    """Direct O(N^2) P2P: ``trg_value += K(trg, src) src_value`` (STKFMM::evaluateKernel).

    Arrays are host numpy float64, C-contiguous, AoS ([n,3] coords, [n,kdim] values).  The number
    of points is taken from the coordinate arrays.  ``n_threads`` is ignored (GPU).
    """
    uk = get_uker(kernel, p2p)
    ks, kt = uker_dims(uk)
    ns = src_coord.size // 3
    nt = trg_coord.size // 3
    a = _host_arr(src_coord, ns, 3, "src_coord")
    b = _host_arr(src_value, ns, ks, "src_value")
    c = _host_arr(trg_coord, nt, 3, "trg_coord")
    d = _host_arr(trg_value, nt, kt, "trg_value", writable=True)
    _check(lib().stkfmm_p2p_eval(int(kernel), int(p2p), int(precision), ns, a, b, nt, c, d))
    return trg_value


class Stk3DFMM:
    """P2P subset of ``Stk3DFMM`` (Python/PySTKFMM.py:34-80, Lib/include/STKFMM/STKFMM.hpp:30-110).

    The FMM parts (set_box/set_points/setup_tree/evaluate_fmm) are the CPU PVFMM path and are not
    part of this library; ``evaluate_kernel`` runs on the GPU.
    """

    def __init__(self, mult_order, max_pts, pbc, kernels):
        self.mult_order = int(mult_order)
        self.max_pts = int(max_pts)
        self.pbc = PAXIS(int(pbc))
        self.kernels = int(kernels)

    def get_kernel_dimension(self, kernel):
        return get_kernel_dimension(kernel)

    def is_kernel_active(self, kernel):
        return bool(self.kernels & int(kernel))

    def show_active_kernels(self):
        names = [get_kernel_name(k) for k in KERNEL if self.kernels & int(k)]
        print("active kernels: " + "".join("\t" + n for n in names))

    def evaluate_kernel(self, kernel, n_threads, p2p, src_coord, src_value, trg_coord, trg_value,
                        precision=PREC.FP64):
        if not self.is_kernel_active(kernel):
            raise KernelNotActive(f"Error: no such FMMData exists for kernel {get_kernel_name(kernel)}")
        return evaluate_kernel(kernel, n_threads, p2p, src_coord, src_value, trg_coord, trg_value, precision)


def ker_eval(uker, r_src, v_src, r_trg, v_trg, dof=1):
    """Ker_t-shaped micro-kernel entry on host arrays: v_trg += ScaleFactor * sum (FP64)."""
    ks, kt = uker_dims(uker)
    ns, nt = r_src.size // 3, r_trg.size // 3
    a = _host_arr(r_src, ns, 3, "r_src")
    b = _host_arr(v_src, ns, ks, "v_src")
    c = _host_arr(r_trg, nt, 3, "r_trg")
    d = _host_arr(v_trg, nt, kt, "v_trg", writable=True)
    _check(lib().stkfmm_ker_eval(int(uker), a, ns, b, int(dof), c, nt, d))
    return v_trg


# ---- box mapping and value scaling (STKFMM::scaleCoord/wrapCoord, FMMData::scaleSrc/scaleTrg) ----

def _np_or_torch(a):
    torch = None
    try:
        import torch  # noqa: F811
    except ImportError:  # pragma: no cover
        pass
    return torch is not None and isinstance(a, torch.Tensor)


def scale_coord(coord, origin, scale_factor, pbc=PAXIS.NONE, stream=None):
    """In place: coord <- (coord - origin) * scale_factor, wrapped to [0,1) on the periodic axes
    (Lib/src/STKFMM.cpp:113-150).  coord: [n, 3] float64 numpy array or contiguous CUDA tensor."""
    o = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
    if _np_or_torch(coord):
        _dev2d(coord, coord.shape[0], 3, "coord")
        if not coord.is_contiguous():
            raise ArgumentError("coord: expected a contiguous tensor")
        _check(lib().stkfmm_scale_coord_device(int(coord.shape[0]), ctypes.c_void_p(coord.data_ptr()),
                                               o.ctypes.data_as(_capi.dptr), float(scale_factor), int(pbc),
                                               _stream_handle(stream)))
        return coord
    n = np.asarray(coord).reshape(-1, 3).shape[0]
    p = _host_arr(coord, n, 3, "coord", writable=True)
    _check(lib().stkfmm_scale_coord(n, p, o.ctypes.data_as(_capi.dptr), float(scale_factor), int(pbc)))
    return coord


def scale_src(kernel, src_sl_value, src_dl_value, scale_factor, stream=None):
    """In place: FMMData::scaleSrc (Lib/src/FMMData.cpp:285-317).  Arrays [n, kdim] (numpy or CUDA);
    either may be None."""
    ks, kd, _ = get_kernel_dimension(kernel)
    sl, dl = src_sl_value, src_dl_value
    if (sl is not None and _np_or_torch(sl)) or (dl is not None and _np_or_torch(dl)):
        nsl = 0 if sl is None else int(sl.shape[0])
        ndl = 0 if dl is None else int(dl.shape[0])
        for t, n, k, name in ((sl, nsl, ks, "src_sl_value"), (dl, ndl, kd, "src_dl_value")):
            if t is not None:
                _dev2d(t, n, k, name)
                if not t.is_contiguous() or t.shape[1] != k:
                    raise ArgumentError(f"{name}: expected a contiguous [{n}, {k}] tensor")
        _check(lib().stkfmm_scale_src_device(int(kernel), nsl, ctypes.c_void_p(0 if sl is None else sl.data_ptr()),
                                             ndl, ctypes.c_void_p(0 if dl is None else dl.data_ptr()),
                                             float(scale_factor), _stream_handle(stream)))
        return sl, dl
    nsl = 0 if sl is None else np.asarray(sl).size // ks
    ndl = 0 if dl is None else np.asarray(dl).size // kd
    ps = None if sl is None else _host_arr(sl, nsl, ks, "src_sl_value", writable=True)
    pd = None if dl is None else _host_arr(dl, ndl, kd, "src_dl_value", writable=True)
    _check(lib().stkfmm_scale_src(int(kernel), nsl, ps, ndl, pd, float(scale_factor)))
    return sl, dl


def scale_trg(kernel, trg_value, scale_factor, stream=None):
    """In place: FMMData::scaleTrg (Lib/src/FMMData.cpp:319-470).  trg_value [n, kdimTrg]."""
    kt = get_kernel_dimension(kernel)[2]
    if _np_or_torch(trg_value):
        n = int(trg_value.shape[0])
        _dev2d(trg_value, n, kt, "trg_value")
        if not trg_value.is_contiguous() or trg_value.shape[1] != kt:
            raise ArgumentError(f"trg_value: expected a contiguous [{n}, {kt}] tensor")
        _check(lib().stkfmm_scale_trg_device(int(kernel), n, ctypes.c_void_p(trg_value.data_ptr()),
                                             float(scale_factor), _stream_handle(stream)))
        return trg_value
    n = np.asarray(trg_value).size // kt
    _check(lib().stkfmm_scale_trg(int(kernel), n, _host_arr(trg_value, n, kt, "trg_value", writable=True),
                                  float(scale_factor)))
    return trg_value


def _wall_dims(kernel):
    if int(kernel) == KERNEL.Stokes:
        return 3, 3
    if int(kernel) == KERNEL.RPY:
        return 4, 6
    raise KernelNotFound("Kernel not supported")


def wall_evaluate(kernel, src_coord, src_value, trg_coord, trg_value, precision=PREC.FP64):
    """StkWallFMM's no-slip wall image system (Lib/src/StkWallFMM.cpp:154-292) by direct sums, host
    arrays: trg_value += u.  The wall is z = 0; KERNEL.Stokes (f -> u) or KERNEL.RPY ([f, b] ->
    [u, lap u]).  FP64 only: PREC.MIXED raises ``Unsupported`` (the image sums cancel near the wall,
    which would amplify FP32 rounding beyond the mixed mode's 1e-6 bound)."""
    ks, kt = _wall_dims(kernel)
    ns = np.asarray(src_coord).reshape(-1, 3).shape[0]
    nt = np.asarray(trg_coord).reshape(-1, 3).shape[0]
    sc = _host_arr(src_coord, ns, 3, "src_coord")
    sv = _host_arr(src_value, ns, ks, "src_value")
    tc = _host_arr(trg_coord, nt, 3, "trg_coord")
    tv = _host_arr(trg_value, nt, kt, "trg_value", writable=True)
    _check(lib().stkfmm_wall_evaluate(int(kernel), int(precision), ns, sc, sv, nt, tc, tv))
    return trg_value


# ---- device (torch) entries ----------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _stream_handle(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _dev2d(t, n, k, name, writable=False):
    """Validate a float64 CUDA tensor with at least n rows and k columns; returns the tensor."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise ArgumentError(f"{name}: expected a float64 CUDA tensor")
    if t.dim() != 2 or t.shape[0] < n or t.shape[1] < k:
        raise ArgumentError(f"{name}: expected shape >= ({n}, {k}), got {tuple(t.shape)}")
    if n > 0 and (t.stride(0) < 0 or t.stride(1) < 0):
        raise ArgumentError(f"{name}: negative strides")
    return t


def _sarr(t):
    return StkArray(t.data_ptr(), t.stride(0), t.stride(1))


def _oarr(t):
    return StkOutArray(t.data_ptr(), t.stride(0), t.stride(1))


def p2p_device(kernel, p2p, src_coord, src_value, trg_coord, trg_value, precision=PREC.FP64, stream=None,
               no_alloc=False):
    """Direct sum on device tensors (contiguous AoS [n,3], [n,ks], [nt,3], [nt,kt]); trg_value +=.

    no_alloc=True calls the entry that allocates no device memory (its work list and partial sums sit
    in a fixed per-device array; sizes above it raise ArgumentError)."""
    uk = get_uker(kernel, p2p)
    ks, kt = uker_dims(uk)
    ns, nt = src_coord.shape[0], trg_coord.shape[0]
    for t, n, k, nm in ((src_coord, ns, 3, "src_coord"), (src_value, ns, ks, "src_value"),
                        (trg_coord, nt, 3, "trg_coord"), (trg_value, nt, kt, "trg_value")):
        _dev2d(t, n, k, nm)
        if not t.is_contiguous() or t.shape[1] != k or t.shape[0] != n:
            raise ArgumentError(f"{nm}: expected a contiguous ({n}, {k}) tensor")
    entry = lib().stkfmm_p2p_eval_device_noalloc if no_alloc else lib().stkfmm_p2p_eval_device
    _check(entry(int(kernel), int(p2p), int(precision), ns, src_coord.data_ptr(), src_value.data_ptr(), nt,
                 trg_coord.data_ptr(), trg_value.data_ptr(), _stream_handle(stream)))
    return trg_value


_DIRECT_PLANS = {}


def uker_device(uker, src_coord, src_value, trg_coord, trg_value, precision=PREC.FP64, stream=None):
    """Direct sum of one micro-kernel on device tensors via a single-group plan; trg_value +=.

    For ``UKER.STOKES_PVEL_FUSED`` src_value is [n, 13] = [f(3), TrD, D(9)].
    """
    ns, nt = src_coord.shape[0], trg_coord.shape[0]
    key = (int(uker), int(precision), ns, nt, src_coord.device.index)
    plan = _DIRECT_PLANS.get(key)
    if plan is None:
        if len(_DIRECT_PLANS) >= 16:
            _DIRECT_PLANS.pop(next(iter(_DIRECT_PLANS)))
        plan = _DIRECT_PLANS[key] = Plan(uker, precision, ns, nt, [0], [nt], None, [0, 1], [0], [ns])
    if int(uker) == UKER.STOKES_PVEL_FUSED:
        _dev2d(src_value, ns, 13, "src_value")
        plan.execute(src_coord, src_value[:, :4], trg_coord, trg_value, src_value2=src_value[:, 4:],
                     stream=stream)
    else:
        plan.execute(src_coord, src_value, trg_coord, trg_value, stream=stream)
    return trg_value


def vjp_device(uker, src_coord, src_value, trg_coord, w, precision=PREC.FP64, need=(True, True, True),
               stream=None):
    """Gradients of L = sum_t w_t . u_t for the direct sum u of ``uker``.

    Returns (g_src_value [ns,ks], g_src_coord [ns,3], g_trg_coord [nt,3]); entries not asked for in
    ``need`` are None.  Device tensors, contiguous AoS.
    """
    torch = _torch()
    ks, kt = uker_dims(uker)
    ns, nt = src_coord.shape[0], trg_coord.shape[0]
    for t, n, k, nm in ((src_coord, ns, 3, "src_coord"), (src_value, ns, ks, "src_value"),
                        (trg_coord, nt, 3, "trg_coord"), (w, nt, kt, "w")):
        _dev2d(t, n, k, nm)
        if not t.is_contiguous() or t.shape[1] != k or t.shape[0] != n:
            raise ArgumentError(f"{nm}: expected a contiguous ({n}, {k}) tensor")
    dev = src_coord.device
    gsv = torch.empty((ns, ks), dtype=torch.float64, device=dev) if need[0] else None
    gsc = torch.empty((ns, 3), dtype=torch.float64, device=dev) if need[1] else None
    gtc = torch.empty((nt, 3), dtype=torch.float64, device=dev) if need[2] else None
    p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    _check(lib().stkfmm_p2p_vjp_device(int(uker), int(precision), ns, src_coord.data_ptr(), src_value.data_ptr(),
                                       nt, trg_coord.data_ptr(), w.data_ptr(), p(gsv), p(gsc), p(gtc),
                                       _stream_handle(stream)))
    return gsv, gsc, gtc


# ---- octree / U-list ---------------------------------------------------------------------------

def wall_device(kernel, src_coord, src_value, trg_coord, trg_value, precision=PREC.FP64, stream=None):
    """``wall_evaluate`` on contiguous AoS float64 CUDA tensors; trg_value += u (asynchronous)."""
    ks, kt = _wall_dims(kernel)
    ns, nt = int(src_coord.shape[0]), int(trg_coord.shape[0])
    for t, n, k, name in ((src_coord, ns, 3, "src_coord"), (src_value, ns, ks, "src_value"),
                          (trg_coord, nt, 3, "trg_coord"), (trg_value, nt, kt, "trg_value")):
        _dev2d(t, n, k, name)
        if not t.is_contiguous() or t.shape[1] != k:
            raise ArgumentError(f"{name}: expected a contiguous [{n}, {k}] tensor")
    _check(lib().stkfmm_wall_eval_device(int(kernel), int(precision), ns, ctypes.c_void_p(src_coord.data_ptr()),
                                         ctypes.c_void_p(src_value.data_ptr()), nt,
                                         ctypes.c_void_p(trg_coord.data_ptr()), ctypes.c_void_p(trg_value.data_ptr()),
                                         _stream_handle(stream)))
    return trg_value


class Tree:
    """Adaptive octree + U-list over points in [0,1)^3 (host).

    Attributes (numpy int32): perm (sorted -> original index), leaf_begin, leaf_end, leaf_level,
    ulist_ptr, rng_begin, rng_end; plus n_leaves, n_ranges, max_level, pairs; leaf_corner [n_leaves, 3].
    All nodes (non-empty boxes, root 0): node_corner, node_level, node_begin, node_end, node_leaf
    (-1 internal), node_parent; W-lists w_ptr/w_node (per leaf, node indices) and X-lists
    x_ptr/x_leaf (per node, leaf ids).
    """

    def __init__(self, coord, leaf_cap):
        coord = np.ascontiguousarray(coord, dtype=np.float64).reshape(-1, 3)
        n = coord.shape[0]
        h = ctypes.c_void_p()
        _check(lib().stkfmm_tree_build(ctypes.byref(h), n, coord.ctypes.data_as(_capi.dptr), int(leaf_cap)))
        try:
            nl, nr, ml = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            pairs = ctypes.c_longlong()
            _check(lib().stkfmm_tree_info(h, ctypes.byref(nl), ctypes.byref(nr), ctypes.byref(ml),
                                          ctypes.byref(pairs)))
            self.n = n
            self.leaf_cap = int(leaf_cap)
            self.n_leaves, self.n_ranges, self.max_level, self.pairs = nl.value, nr.value, ml.value, pairs.value
            self.perm = np.zeros(n, dtype=np.int32)
            self.leaf_begin = np.zeros(self.n_leaves, dtype=np.int32)
            self.leaf_end = np.zeros(self.n_leaves, dtype=np.int32)
            self.leaf_level = np.zeros(self.n_leaves, dtype=np.int32)
            self.ulist_ptr = np.zeros(self.n_leaves + 1, dtype=np.int32)
            self.rng_begin = np.zeros(max(self.n_ranges, 1), dtype=np.int32)
            self.rng_end = np.zeros(max(self.n_ranges, 1), dtype=np.int32)
            ip = lambda a: a.ctypes.data_as(_capi.iptr)  # noqa: E731
            _check(lib().stkfmm_tree_get(h, ip(self.perm), ip(self.leaf_begin), ip(self.leaf_end),
                                         ip(self.leaf_level), ip(self.ulist_ptr), ip(self.rng_begin),
                                         ip(self.rng_end)))
            self.rng_begin = self.rng_begin[: self.n_ranges]
            self.rng_end = self.rng_end[: self.n_ranges]
            self.leaf_corner = np.zeros((self.n_leaves, 3))
            _check(lib().stkfmm_tree_leaf_corners(h, self.leaf_corner.ctypes.data_as(_capi.dptr)))
            nn, nw = ctypes.c_int(), ctypes.c_int()
            _check(lib().stkfmm_tree_node_info(h, ctypes.byref(nn), ctypes.byref(nw)))
            self.n_nodes = nn.value
            self.node_corner = np.zeros((self.n_nodes, 3))
            self.node_level, self.node_begin, self.node_end, self.node_leaf, self.node_parent = (
                np.zeros(self.n_nodes, dtype=np.int32) for _ in range(5))
            _check(lib().stkfmm_tree_get_nodes(h, self.node_corner.ctypes.data_as(_capi.dptr), ip(self.node_level),
                                               ip(self.node_begin), ip(self.node_end), ip(self.node_leaf),
                                               ip(self.node_parent)))
            self.w_ptr = np.zeros(self.n_leaves + 1, dtype=np.int32)
            self.w_node = np.zeros(max(nw.value, 1), dtype=np.int32)
            self.x_ptr = np.zeros(self.n_nodes + 1, dtype=np.int32)
            self.x_leaf = np.zeros(max(nw.value, 1), dtype=np.int32)
            _check(lib().stkfmm_tree_get_wx(h, ip(self.w_ptr), ip(self.w_node), ip(self.x_ptr), ip(self.x_leaf)))
            self.w_node = self.w_node[: nw.value]
            self.x_leaf = self.x_leaf[: nw.value]
        finally:
            lib().stkfmm_tree_destroy(h)


# ---- equivalent / check surfaces -----------------------------------------------------------------

# PVFMM's surface radii relative to the box (M2L/SVD_pvfmm.hpp:37-38): upward equivalent and
# downward check surfaces at 1.05, upward check and downward equivalent surfaces at 2.95
RAD_INNER = 1.05
RAD_OUTER = 2.95


def surface_count(p):
    """Points of an order-p surface: 6 (p-1)^2 + 2."""
    n = lib().stkfmm_surface_count(int(p))
    if n < 0:
        raise ArgumentError("surface order p must be >= 2")
    return n


def surface(p, c, alpha, depth):
    """stkfmm::surface (Lib/include/STKFMM/stkfmm_helpers.hpp:46-85) as a [count, 3] array."""
    c = np.ascontiguousarray(c, dtype=np.float64).reshape(3)
    out = np.zeros((surface_count(p), 3))
    _check(lib().stkfmm_surface(int(p), c.ctypes.data_as(_capi.dptr), float(alpha), int(depth),
                                out.ctypes.data_as(_capi.dptr)))
    return out


def box_surfaces_device(p, alpha, corner, level, stream=None):
    """Surfaces of many octree boxes on the device: corner [n, 3] (float64) and level [n] (int32) CUDA
    tensors -> [n * count, 3] tensor; box b's points are rows [b*count, (b+1)*count), centred on the
    box with half-width alpha * 0.5^(level+1)."""
    torch = _torch()
    n = int(corner.shape[0])
    _dev2d(corner, n, 3, "corner")
    if not corner.is_contiguous() or corner.shape[1] != 3:
        raise ArgumentError("corner: expected a contiguous [n, 3] tensor")
    if level.dtype != torch.int32 or not level.is_cuda or not level.is_contiguous() or level.shape[0] != n:
        raise ArgumentError("level: expected a contiguous int32 CUDA tensor of length n")
    ns = surface_count(p)
    out = torch.empty((n * ns, 3), dtype=torch.float64, device=corner.device)
    _check(lib().stkfmm_box_surfaces_device(int(p), float(alpha), n, ctypes.c_void_p(corner.data_ptr()),
                                            ctypes.c_void_p(level.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                            _stream_handle(stream)))
    return out


class NodeSurfaces:
    """One order-p surface per node of a Tree (all non-empty boxes), on the device, and the W- and
    X-list work lists that use them (the M2T and S2L evaluations of PVFMM's near field):

    * ``w_plan``: for every leaf B, the upward equivalent surfaces (radius RAD_INNER) of the nodes in
      W(B), carrying densities, evaluated at B's points (kernel k_m2t, ``get_m2t_uker``);
    * ``x_plan``: for every node C with a non-empty X(C), the points of the leaves in X(C) evaluated
      at C's downward check surface (radius RAD_INNER; kernel k_s2l, ``get_s2l_uker``).
    ``coord`` is [n_nodes * count, 3], node c at rows [c*count, (c+1)*count).
    """

    def __init__(self, tree, p, alpha=RAD_INNER, device=None, stream=None):
        torch = _torch()
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.tree, self.p, self.alpha = tree, int(p), float(alpha)
        self.count = surface_count(p)
        corner = torch.from_numpy(tree.node_corner).to(device)
        level = torch.from_numpy(tree.node_level.astype(np.int32)).to(device)
        self.coord = box_surfaces_device(p, alpha, corner, level, stream)

    def w_plan(self, uker, precision, tune=None):
        t = self.tree
        ns = self.count
        rb = t.w_node.astype(np.int64) * ns
        return Plan(uker, precision, t.n_nodes * ns, t.n, t.leaf_begin, t.leaf_end, t.leaf_level, t.w_ptr, rb,
                    rb + ns, tune)

    def x_plan(self, uker, precision, tune=None):
        t = self.tree
        ns = self.count
        nodes = np.nonzero(np.diff(t.x_ptr) > 0)[0]
        gb = nodes.astype(np.int64) * ns
        ptr = np.concatenate([[0], np.cumsum(np.diff(t.x_ptr)[nodes])])
        leaves = t.x_leaf
        return Plan(uker, precision, t.n, t.n_nodes * ns, gb, gb + ns, t.node_level[nodes], ptr,
                    t.leaf_begin[leaves], t.leaf_end[leaves], tune)


class LeafSurfaces:
    """One order-p surface per leaf of a Tree, on the device, and the per-leaf work lists that use it.

    * ``to_surface_plan``: every leaf's points (sources) evaluated at the leaf's own surface
      (targets) -- the P2P part of PVFMM's S2M step (sources -> upward check surface, radius
      RAD_OUTER, kernel k_s2m).
    * ``from_surface_plan``: every leaf's surface (sources, e.g. downward equivalent densities,
      radius RAD_OUTER) evaluated at the leaf's points -- the P2P part of L2T (kernel k_l2t).
    Both are ordinary batched plans (one target group per leaf, level-scaled in mixed mode); the
    surface points are ``coord`` ([n_leaves * count, 3], leaf l at rows [l*count, (l+1)*count)).
    """

    def __init__(self, tree, p, alpha=RAD_OUTER, device=None, stream=None):
        torch = _torch()
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.tree, self.p, self.alpha = tree, int(p), float(alpha)
        self.count = surface_count(p)
        corner = torch.from_numpy(tree.leaf_corner).to(device)
        level = torch.from_numpy(tree.leaf_level.astype(np.int32)).to(device)
        self.coord = box_surfaces_device(p, alpha, corner, level, stream)

    def _surf_ranges(self):
        nl = self.tree.n_leaves
        b = np.arange(nl, dtype=np.int64) * self.count
        return b, b + self.count

    def to_surface_plan(self, uker, precision, tune=None):
        t, nl = self.tree, self.tree.n_leaves
        sb, se = self._surf_ranges()
        return Plan(uker, precision, t.n, nl * self.count, sb, se, t.leaf_level, np.arange(nl + 1), t.leaf_begin,
                    t.leaf_end, tune)

    def from_surface_plan(self, uker, precision, tune=None):
        t, nl = self.tree, self.tree.n_leaves
        sb, se = self._surf_ranges()
        return Plan(uker, precision, nl * self.count, t.n, t.leaf_begin, t.leaf_end, t.leaf_level,
                    np.arange(nl + 1), sb, se, tune)


# ---- batched plans ------------------------------------------------------------------------------

class Plan:
    """A batched P2P work list on the current device (``stkfmm_plan_*``).

    Group g owns targets [grp_trg_begin[g], grp_trg_end[g]) and the source ranges
    [rng_begin[r], rng_end[r]) for r in [grp_rng_ptr[g], grp_rng_ptr[g+1]).
    """

    def __init__(self, uker, precision, n_src, n_trg, grp_trg_begin, grp_trg_end, grp_level, grp_rng_ptr,
                 rng_begin, rng_end, tune=None):
        i32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        gb, ge, gl, gp, rb, re = map(i32, (grp_trg_begin, grp_trg_end, grp_level, grp_rng_ptr, rng_begin,
                                           rng_end))
        ng = len(gb)
        if len(ge) != ng or len(gp) != ng + 1 or (gl is not None and len(gl) != ng) or len(rb) != len(re):
            raise ArgumentError("inconsistent work-list array lengths")
        if ng and gp[-1] > len(rb):
            raise ArgumentError("grp_rng_ptr points past the range arrays")
        ip = lambda a: None if a is None else a.ctypes.data_as(_capi.iptr)  # noqa: E731
        if isinstance(tune, str):
            if tune != "auto":
                raise ArgumentError(f"tune: expected a dict, None or 'auto', got {tune!r}")
            from .autotune import lookup, signature
            csum = np.concatenate([[0], np.cumsum(re.astype(np.int64) - rb.astype(np.int64))])
            per_group = csum[gp[1:]] - csum[gp[:-1]]
            pairs = int(np.sum(per_group * (ge.astype(np.int64) - gb.astype(np.int64))))
            tune = lookup(signature(uker, precision, n_trg, pairs, ng))
        t = None
        if tune is not None:
            t = StkTune(int(tune.get("targets_per_lane", 0)), int(tune.get("max_src_per_item", 0)),
                        int(tune.get("persistent", 0)), int(tune.get("waves_per_cu", 0)),
                        int(tune.get("split_tiles", 0)), int(tune.get("broadcast", 0)),
                        int(tune.get("queues", 0)), int(tune.get("sort_points", 0)))
        h = ctypes.c_void_p()
        _check(lib().stkfmm_plan_create(ctypes.byref(h), int(uker), int(precision), int(n_src), int(n_trg), ng,
                                        ip(gb), ip(ge), ip(gl), ip(gp), ip(rb), ip(re),
                                        ctypes.byref(t) if t is not None else None))
        self._h = h
        self.uker = UKER(int(uker))
        self.precision = PREC(int(precision))
        self.n_src, self.n_trg = int(n_src), int(n_trg)
        self.ks, self.kt = uker_dims(uker)
        self._keep = None

    @classmethod
    def from_tree(cls, tree, uker, precision, tune=None):
        """U-list plan over a Tree: every leaf is a target group against its U-list."""
        return cls(uker, precision, tree.n, tree.n, tree.leaf_begin, tree.leaf_end, tree.leaf_level,
                   tree.ulist_ptr, tree.rng_begin, tree.rng_end, tune)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().stkfmm_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def stats(self):
        pairs = ctypes.c_longlong()
        ni, nt, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().stkfmm_plan_stats(self._h, ctypes.byref(pairs), ctypes.byref(ni), ctypes.byref(nt),
                                       ctypes.byref(g)))
        return dict(pairs=pairs.value, n_items=ni.value, n_tiles=nt.value, grid=g.value)

    def set_timing(self, enable=True):
        _check(lib().stkfmm_plan_set_timing(self._h, 1 if enable else 0))

    def kernel_ms(self):
        """(average item-kernel ms, average reduce ms, count) over executes recorded since set_timing."""
        a, b, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        _check(lib().stkfmm_plan_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def _args(self, src_coord, src_value, trg_coord, trg_value, src_value2, perm):
        ks1 = 4 if self.uker == UKER.STOKES_PVEL_FUSED else self.ks
        _dev2d(src_coord, self.n_src, 3, "src_coord")
        _dev2d(src_value, self.n_src, ks1, "src_value")
        _dev2d(trg_coord, self.n_trg, 3, "trg_coord")
        if self.uker == UKER.STOKES_PVEL_FUSED:
            if src_value2 is None:
                raise ArgumentError("STOKES_PVEL_FUSED needs src_value2 (double-layer values)")
            _dev2d(src_value2, self.n_src, 9, "src_value2")
        if perm is not None:
            torch = _torch()
            if not perm.is_cuda or perm.dtype != torch.int32 or perm.numel() < self.n_trg:
                raise ArgumentError("perm: expected an int32 CUDA tensor of n_trg entries")
            n_out = trg_value.shape[0]
            _dev2d(trg_value, 0, self.kt, "trg_value")
            if self.n_trg and (int(perm[: self.n_trg].min()) < 0 or int(perm[: self.n_trg].max()) >= n_out):
                raise ArgumentError("perm: values outside the output rows")
        else:
            _dev2d(trg_value, self.n_trg, self.kt, "trg_value")
        sv2 = _sarr(src_value2) if src_value2 is not None else _sarr(src_value)
        args = (_sarr(src_coord), _sarr(src_value), sv2, _sarr(trg_coord), _oarr(trg_value))
        return args, (None if perm is None else perm.data_ptr())

    def execute(self, src_coord, src_value, trg_coord, trg_value, src_value2=None, perm=None, stream=None):
        """trg_value(perm[i]) += ScaleFactor * sum over the work list (device tensors, any 2-D strides)."""
        (a, b, c, d, e), p = self._args(src_coord, src_value, trg_coord, trg_value, src_value2, perm)
        _check(lib().stkfmm_plan_execute(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(d), ctypes.byref(e), p, _stream_handle(stream)))
        return trg_value

    def capture_graph(self, src_coord, src_value, trg_coord, trg_value, src_value2=None, perm=None, stream=None):
        """Capture execute(...) with these tensors into a HIP graph; replay with launch_graph()."""
        (a, b, c, d, e), p = self._args(src_coord, src_value, trg_coord, trg_value, src_value2, perm)
        _check(lib().stkfmm_plan_capture_graph(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                               ctypes.byref(d), ctypes.byref(e), p, _stream_handle(stream)))
        self._keep = (src_coord, src_value, trg_coord, trg_value, src_value2, perm)

    def launch_graph(self, stream=None):
        _check(lib().stkfmm_plan_launch_graph(self._h, _stream_handle(stream)))
