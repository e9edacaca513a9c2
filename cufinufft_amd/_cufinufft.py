"""Low-level ctypes bindings of libcufinufft.so (the in-tree HIP build).

Mirrors python/cufinufft/_cufinufft.py of cuFINUFFT v1.3: the same NufftOpts
structure, the same function objects (_default_opts, _make_plan[f],
_set_pts[f], _exec_plan[f], _destroy_plan[f]) with the same argtypes, plus the
library's extension entry points (_set_stream[f], _get_info[f],
_get_sort[f], _get_fseries[f], _get_kernel_poly[f]).

Loading fails loudly: there is no CPU fallback.
"""
import ctypes
import os
from ctypes import c_double, c_float, c_int, c_void_p

import numpy as np

c_int_p = ctypes.POINTER(c_int)
c_float_p = ctypes.POINTER(c_float)
c_double_p = ctypes.POINTER(c_double)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUFINUFFT_AMD_LIB", os.path.join(_HERE, "libcufinufft.so"))

if not os.path.exists(LIB_PATH):
    raise RuntimeError(
        f"libcufinufft.so not found at {LIB_PATH}; build it with "
        "`python cufinufft_amd/build.py` (hipcc, gfx950)")
lib = ctypes.cdll.LoadLibrary(LIB_PATH)


def _get_ctypes(dtype):
    """Return (real ctype, pointer ctype) for np.float32 / np.float64."""
    if dtype == np.float64:
        REAL_t = c_double
    elif dtype == np.float32:
        REAL_t = c_float
    else:
        raise TypeError("Expected np.float32 or np.float64.")
    return REAL_t, ctypes.POINTER(REAL_t)


def _get_NufftOpts():
    return [
        ('upsampfac', c_double),
        ('gpu_method', c_int),
        ('gpu_sort', c_int),
        ('gpu_binsizex', c_int),
        ('gpu_binsizey', c_int),
        ('gpu_binsizez', c_int),
        ('gpu_obinsizex', c_int),
        ('gpu_obinsizey', c_int),
        ('gpu_obinsizez', c_int),
        ('gpu_maxsubprobsize', c_int),
        ('gpu_nstreams', c_int),
        ('gpu_kerevalmeth', c_int),
        ('gpu_spreadinterponly', c_int),
        ('gpu_device_id', c_int)]


class NufftOpts(ctypes.Structure):
    pass


NufftOpts._fields_ = _get_NufftOpts()

CufinufftPlan = c_void_p
CufinufftPlanf = c_void_p
CufinufftPlan_p = ctypes.POINTER(CufinufftPlan)
CufinufftPlanf_p = ctypes.POINTER(CufinufftPlanf)
NufftOpts_p = ctypes.POINTER(NufftOpts)

_default_opts = lib.cufinufft_default_opts
_default_opts.argtypes = [c_int, c_int, NufftOpts_p]
_default_opts.restype = c_int

_make_plan = lib.cufinufft_makeplan
_make_plan.argtypes = [c_int, c_int, c_int_p, c_int, c_int, c_double, c_int,
                       CufinufftPlan_p, NufftOpts_p]
_make_plan.restype = c_int

_make_planf = lib.cufinufftf_makeplan
_make_planf.argtypes = [c_int, c_int, c_int_p, c_int, c_int, c_float, c_int,
                        CufinufftPlanf_p, NufftOpts_p]
_make_planf.restype = c_int

_make_plan_ext = lib.cufinufft_makeplan_ext
_make_plan_ext.argtypes = [c_int, c_int, c_int_p, c_int, c_int, c_double, c_int,
                           CufinufftPlan_p, NufftOpts_p, c_int]
_make_plan_ext.restype = c_int

_make_planf_ext = lib.cufinufftf_makeplan_ext
_make_planf_ext.argtypes = [c_int, c_int, c_int_p, c_int, c_int, c_float, c_int,
                            CufinufftPlanf_p, NufftOpts_p, c_int]
_make_planf_ext.restype = c_int

ACCURACY_FIRST = 1
TILE_SLOTS = 2
FULL_FFT = 4

_set_pts = lib.cufinufft_setpts
_set_pts.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_double_p,
                     c_double_p, c_double_p, c_void_p]
_set_pts.restype = c_int

_set_ptsf = lib.cufinufftf_setpts
_set_ptsf.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_float_p,
                      c_float_p, c_float_p, c_void_p]
_set_ptsf.restype = c_int

_exec_plan = lib.cufinufft_execute
_exec_plan.argtypes = [c_void_p, c_void_p, c_void_p]
_exec_plan.restype = c_int

_exec_planf = lib.cufinufftf_execute
_exec_planf.argtypes = [c_void_p, c_void_p, c_void_p]
_exec_planf.restype = c_int

_destroy_plan = lib.cufinufft_destroy
_destroy_plan.argtypes = [c_void_p]
_destroy_plan.restype = c_int

_destroy_planf = lib.cufinufftf_destroy
_destroy_planf.argtypes = [c_void_p]
_destroy_planf.restype = c_int

# ---- extensions -----------------------------------------------------------
_set_stream = lib.cufinufft_setstream
_set_stream.argtypes = [c_void_p, c_void_p]
_set_stream.restype = c_int
_set_streamf = lib.cufinufftf_setstream
_set_streamf.argtypes = [c_void_p, c_void_p]
_set_streamf.restype = c_int

_get_info = lib.cufinufft_get_info
_get_info.argtypes = [c_void_p, c_void_p]
_get_info.restype = c_int
_get_infof = lib.cufinufftf_get_info
_get_infof.argtypes = [c_void_p, c_void_p]
_get_infof.restype = c_int

_get_sort = lib.cufinufft_get_sort
_get_sort.argtypes = [c_void_p] + [c_void_p] * 5
_get_sort.restype = c_int
_get_sortf = lib.cufinufftf_get_sort
_get_sortf.argtypes = [c_void_p] + [c_void_p] * 5
_get_sortf.restype = c_int

_get_fseries = lib.cufinufft_get_fseries
_get_fseries.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_get_fseries.restype = c_int
_get_fseriesf = lib.cufinufftf_get_fseries
_get_fseriesf.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_get_fseriesf.restype = c_int

_get_kernel_poly = lib.cufinufft_get_kernel_poly
_get_kernel_poly.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_get_kernel_poly.restype = c_int
_get_kernel_polyf = lib.cufinufftf_get_kernel_poly
_get_kernel_polyf.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_get_kernel_polyf.restype = c_int

_set_timing = lib.cufinufft_set_timing
_set_timing.argtypes = [c_void_p, c_int]
_set_timing.restype = c_int
_set_timingf = lib.cufinufftf_set_timing
_set_timingf.argtypes = [c_void_p, c_int]
_set_timingf.restype = c_int
_get_timing = lib.cufinufft_get_timing
_get_timing.argtypes = [c_void_p, c_void_p, c_void_p]
_get_timing.restype = c_int
_get_timingf = lib.cufinufftf_get_timing
_get_timingf.argtypes = [c_void_p, c_void_p, c_void_p]
_get_timingf.restype = c_int

_exec_grad = lib.cufinufft_execute_grad
_exec_grad.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_exec_grad.restype = c_int
_exec_gradf = lib.cufinufftf_execute_grad
_exec_gradf.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
_exec_gradf.restype = c_int

for _suffix in ("", "f"):
    _f = getattr(lib, f"cufinufft{_suffix}_group_make")
    _f.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
    _f.restype = c_int
    _f = getattr(lib, f"cufinufft{_suffix}_group_execute")
    _f.argtypes = [c_void_p, c_int]
    _f.restype = c_int
    _f = getattr(lib, f"cufinufft{_suffix}_group_setstream")
    _f.argtypes = [c_void_p, c_void_p]
    _f.restype = c_int
    _f = getattr(lib, f"cufinufft{_suffix}_group_destroy")
    _f.argtypes = [c_void_p]
    _f.restype = c_int

# setpts without allocation (FP32 plans): sort scratch from the library's
# fixed device array; returns 13 beyond what is allocated
_set_pts_noallocf = lib.cufinufftf_setpts_noalloc
_set_pts_noallocf.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_void_p]
_set_pts_noallocf.restype = c_int

STAGES = ("sort", "grid_zero", "spread_interp", "fft", "deconv_amplify")

EXPORTED = [
    "cufinufft_default_opts", "cufinufftf_default_opts",
    "cufinufft_makeplan", "cufinufftf_makeplan",
    "cufinufft_makeplan_ext", "cufinufftf_makeplan_ext",
    "cufinufft_setpts", "cufinufftf_setpts",
    "cufinufft_execute", "cufinufftf_execute",
    "cufinufft_destroy", "cufinufftf_destroy",
    "cufinufft_setstream", "cufinufftf_setstream",
    "cufinufft_get_info", "cufinufftf_get_info",
    "cufinufft_get_sort", "cufinufftf_get_sort",
    "cufinufft_get_fseries", "cufinufftf_get_fseries",
    "cufinufft_get_kernel_poly", "cufinufftf_get_kernel_poly",
    "cufinufft_set_timing", "cufinufftf_set_timing",
    "cufinufft_get_timing", "cufinufftf_get_timing",
    "cufinufft_execute_grad", "cufinufftf_execute_grad",
    "cufinufft_group_make", "cufinufftf_group_make",
    "cufinufft_group_execute", "cufinufftf_group_execute",
    "cufinufft_group_setstream", "cufinufftf_group_setstream",
    "cufinufft_group_destroy", "cufinufftf_group_destroy",
    "cufinufftf_setpts_noalloc",
]
