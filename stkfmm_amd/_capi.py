"""ctypes declarations of libstkfmm_p2p.so (the C ABI in include/stkfmm_p2p.h).

The library is loaded from this package directory; there is no fallback.  A missing library is an
ImportError naming the build step, so nothing silently runs another code path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstkfmm_p2p.so")

c_int, c_uint, c_double, c_void_p, c_longlong, c_float, c_char_p = (
    ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_float,
    ctypes.c_char_p)
dptr = ctypes.POINTER(ctypes.c_double)
iptr = ctypes.POINTER(ctypes.c_int)


class StkArray(ctypes.Structure):
    """stkfmm_array: element (i, k) at ptr[i*stride + k*fstride] (device pointer)."""
    _fields_ = [("ptr", c_void_p), ("stride", c_longlong), ("fstride", c_longlong)]


class StkOutArray(ctypes.Structure):
    _fields_ = [("ptr", c_void_p), ("stride", c_longlong), ("fstride", c_longlong)]


class StkTune(ctypes.Structure):
    _fields_ = [("targets_per_lane", c_int), ("max_src_per_item", c_int), ("persistent", c_int),
                ("waves_per_cu", c_int), ("split_tiles", c_int), ("broadcast", c_int), ("queues", c_int),
                ("sort_points", c_int)]


KER_T_NAMES = ["laplace_p", "laplace_dipolep", "laplace_pgrad", "laplace_dipolepgrad", "laplace_pgradgrad",
               "laplace_dipolepgradgrad", "stokes_vel", "stokes_pvel", "stokes_doublepvel", "stokes_regvel",
               "rpy_u", "rpy_ulapu", "stk_ulapu", "stokes_pvelgrad", "stokes_doublepvelgrad",
               "stokes_pvellaplacian", "stokes_doublelaplacian", "stokes_traction", "stokes_doubletraction",
               "stokes_regftvelomega", "stokes_velomega", "laplace_quadpgradgrad", "laplace_quadp", "stokes_regftvel"]
# micro-kernel id (STKFMM_UKER_*) of each Ker_t entry; 13 is the fused PVel kernel, which has no Ker_t entry
KER_T_UKER = {n: (i if i < 13 else i + 1) for i, n in enumerate(KER_T_NAMES)}

# every symbol include/stkfmm_p2p.h declares
EXPORTS = [
    "stkfmm_last_error", "stkfmm_get_kernel_dimension", "stkfmm_get_kernel_name", "stkfmm_get_uker",
    "stkfmm_uker_dims", "stkfmm_evaluate_kernel", "stkfmm_p2p_eval", "stkfmm_p2p_eval_device",
    "stkfmm_ker_eval", "stkfmm_plan_create", "stkfmm_plan_execute", "stkfmm_plan_capture_graph",
    "stkfmm_plan_launch_graph", "stkfmm_plan_stats", "stkfmm_plan_set_timing", "stkfmm_plan_kernel_ms",
    "stkfmm_plan_destroy", "stkfmm_tree_build", "stkfmm_tree_info", "stkfmm_tree_get", "stkfmm_tree_destroy",
    "stkfmm_p2p_vjp_device", "stkfmm_device_info", "stkfmm_tree_leaf_corners", "stkfmm_surface_count",
    "stkfmm_surface", "stkfmm_box_surfaces_device", "stkfmm_get_s2m_uker", "stkfmm_wall_evaluate",
    "stkfmm_wall_eval_device", "stkfmm_scale_coord", "stkfmm_scale_src", "stkfmm_scale_trg",
    "stkfmm_scale_coord_device", "stkfmm_scale_src_device", "stkfmm_scale_trg_device", "stkfmm_tree_node_info",
    "stkfmm_tree_get_nodes", "stkfmm_tree_get_wx", "stkfmm_get_s2l_uker", "stkfmm_get_m2t_uker",
    "stkfmm_p2p_eval_device_noalloc",
] + ["stkfmm_ker_" + n for n in KER_T_NAMES]


def _declare(L):
    L.stkfmm_last_error.restype = c_char_p
    L.stkfmm_last_error.argtypes = []
    L.stkfmm_get_kernel_dimension.argtypes = [c_uint, iptr]
    L.stkfmm_get_kernel_name.argtypes = [c_uint, c_char_p, c_int]
    L.stkfmm_get_uker.argtypes = [c_uint, c_uint, iptr]
    L.stkfmm_uker_dims.argtypes = [c_int, iptr, iptr]
    L.stkfmm_evaluate_kernel.argtypes = [c_uint, c_int, c_uint, c_int, dptr, dptr, c_int, dptr, dptr]
    L.stkfmm_p2p_eval.argtypes = [c_uint, c_uint, c_int, c_int, dptr, dptr, c_int, dptr, dptr]
    L.stkfmm_p2p_eval_device.argtypes = [c_uint, c_uint, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                         c_void_p, c_void_p]
    L.stkfmm_p2p_eval_device_noalloc.argtypes = [c_uint, c_uint, c_int, c_int, c_void_p, c_void_p, c_int,
                                                 c_void_p, c_void_p, c_void_p]
    L.stkfmm_ker_eval.argtypes = [c_int, dptr, c_int, dptr, c_int, dptr, c_int, dptr]
    for n in KER_T_NAMES:
        f = getattr(L, "stkfmm_ker_" + n)
        f.restype = None
        f.argtypes = [dptr, c_int, dptr, c_int, dptr, c_int, dptr, c_void_p]
    L.stkfmm_plan_create.argtypes = [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, iptr, iptr, iptr,
                                     iptr, iptr, iptr, ctypes.POINTER(StkTune)]
    arr = ctypes.POINTER(StkArray)
    oarr = ctypes.POINTER(StkOutArray)
    L.stkfmm_plan_execute.argtypes = [c_void_p, arr, arr, arr, arr, oarr, c_void_p, c_void_p]
    L.stkfmm_plan_capture_graph.argtypes = [c_void_p, arr, arr, arr, arr, oarr, c_void_p, c_void_p]
    L.stkfmm_plan_launch_graph.argtypes = [c_void_p, c_void_p]
    L.stkfmm_plan_stats.argtypes = [c_void_p, ctypes.POINTER(c_longlong), iptr, iptr, iptr]
    L.stkfmm_plan_set_timing.argtypes = [c_void_p, c_int]
    L.stkfmm_plan_kernel_ms.argtypes = [c_void_p, ctypes.POINTER(c_float), ctypes.POINTER(c_float), iptr]
    L.stkfmm_plan_destroy.argtypes = [c_void_p]
    L.stkfmm_tree_build.argtypes = [ctypes.POINTER(c_void_p), c_int, dptr, c_int]
    L.stkfmm_tree_info.argtypes = [c_void_p, iptr, iptr, iptr, ctypes.POINTER(c_longlong)]
    L.stkfmm_tree_get.argtypes = [c_void_p, iptr, iptr, iptr, iptr, iptr, iptr, iptr]
    L.stkfmm_tree_destroy.argtypes = [c_void_p]
    L.stkfmm_tree_leaf_corners.argtypes = [c_void_p, dptr]
    L.stkfmm_tree_node_info.argtypes = [c_void_p, iptr, iptr]
    L.stkfmm_tree_get_nodes.argtypes = [c_void_p, dptr, iptr, iptr, iptr, iptr, iptr]
    L.stkfmm_tree_get_wx.argtypes = [c_void_p, iptr, iptr, iptr, iptr]
    L.stkfmm_get_s2l_uker.argtypes = [c_uint, c_int, iptr]
    L.stkfmm_get_m2t_uker.argtypes = [c_uint, iptr]
    L.stkfmm_surface_count.argtypes = [c_int]
    L.stkfmm_surface.argtypes = [c_int, dptr, ctypes.c_double, c_int, dptr]
    L.stkfmm_box_surfaces_device.argtypes = [c_int, ctypes.c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    L.stkfmm_get_s2m_uker.argtypes = [c_uint, c_int, iptr]
    L.stkfmm_scale_coord.argtypes = [c_int, dptr, dptr, ctypes.c_double, c_uint]
    L.stkfmm_scale_src.argtypes = [c_uint, c_int, dptr, c_int, dptr, ctypes.c_double]
    L.stkfmm_scale_trg.argtypes = [c_uint, c_int, dptr, ctypes.c_double]
    L.stkfmm_scale_coord_device.argtypes = [c_int, c_void_p, dptr, ctypes.c_double, c_uint, c_void_p]
    L.stkfmm_scale_src_device.argtypes = [c_uint, c_int, c_void_p, c_int, c_void_p, ctypes.c_double, c_void_p]
    L.stkfmm_scale_trg_device.argtypes = [c_uint, c_int, c_void_p, ctypes.c_double, c_void_p]
    L.stkfmm_wall_evaluate.argtypes = [c_uint, c_int, c_int, dptr, dptr, c_int, dptr, dptr]
    L.stkfmm_wall_eval_device.argtypes = [c_uint, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                          c_void_p]
    L.stkfmm_p2p_vjp_device.argtypes = [c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p]
    L.stkfmm_device_info.argtypes = [iptr, iptr, c_char_p, c_int]
    return L


_lib = None


def lib():
    """The loaded library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"stkfmm_amd: native library {LIB_PATH} is missing; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C stkfmm_amd/csrc`")
        # torch ships its own HIP runtime (soname libamdhip64.so.7, but libtorch_hip asks for the
        # unversioned libamdhip64.so).  Loaded after this library, torch cannot match the runtime
        # already in the process and brings a second one; whichever initialises second then finds
        # no device.  Loading torch first lets this library bind to torch's runtime instead.
        try:
            import torch  # noqa: F401
        except (ImportError, OSError):  # torch absent or its own libraries unloadable: not needed here
            pass
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def last_error():
    return lib().stkfmm_last_error().decode()
