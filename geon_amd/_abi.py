"""ctypes binding of include/geon_hip.h (libgeon_hip.so + geon_kernels.hsaco).

There is no fallback: importing this module on a machine without the built
library raises, and every call that reaches the GPU goes through the HIP code
object loaded by the native runtime.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libgeon_hip.so")
HSACO_PATH = os.path.join(LIB_DIR, "geon_kernels.hsaco")

OK = 0
E_INVALID, E_NOT_TIMELIKE, E_GRAM_SCHMIDT, E_SINGULAR, E_HIP = -1, -2, -3, -4, -5
E_NOMEM, E_ADMISSION, E_NO_MODULE, E_UNSUPPORTED, E_NOT_FOUND, E_PENDING = -6, -7, -8, -9, -10, -11

MINKOWSKI, SCHWARZSCHILD, ALCUBIERRE, MORRIS_THORNE = 0, 1, 2, 3
DIVERGED, CAPTURED, ESCAPED, MAXSTEPS = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
CURV_RIEMANN, CURV_RICCI, CURV_EINSTEIN, CURV_STRESS, CURV_SCALAR = 0, 1, 2, 3, 4


class GeonError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        super().__init__(f"{where}: {strerror(status)} ({status}){': ' + detail if detail else ''}")


class Metric(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("p", C.c_double * 6),
                ("horizon_pad", C.c_double), ("escape_radius", C.c_double),
                ("fd_step", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("observer", C.c_double * 4), ("fov_y_rad", C.c_double),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class RenderCfg(C.Structure):
    _fields_ = [("step", C.c_double), ("max_steps", C.c_uint32),
                ("samples_per_axis", C.c_uint32), ("captured_rgb", C.c_float * 3),
                ("diverged_rgb", C.c_float * 3), ("show_stars", C.c_uint32),
                ("precision", C.c_int32), ("fp64_band", C.c_double)]


class SchedCfg(C.Structure):
    _fields_ = [("block", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("rays_per_lane", C.c_uint32), ("refill", C.c_uint32)]


class ServeCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("quantum_steps", C.c_uint32),
                ("max_jobs", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("captured", C.c_uint64), ("escaped", C.c_uint64), ("diverged", C.c_uint64),
                ("max_null_residual", C.c_double), ("avg_steps", C.c_double),
                ("max_steps_hit", C.c_uint64), ("ray_steps", C.c_uint64),
                ("fp64_rays", C.c_uint64), ("kernel_ms", C.c_double),
                ("gamma_steps", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeResult(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("stats", RenderStats), ("ray_steps", C.c_uint64)]


# (name, restype, argtypes) of every entry point declared in include/geon_hip.h
_vp, _dp, _i, _u32, _sz = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_uint32, C.c_size_t
SIGNATURES = [
    ("geon_strerror", C.c_char_p, [_i]),
    ("geon_last_error", C.c_char_p, []),
    ("geon_version", _i, []),
    ("geon_metric_default", _i, [C.c_int32, C.POINTER(Metric)]),
    ("geon_metric_g", _i, [C.POINTER(Metric), _vp, _vp]),
    ("geon_metric_dg", _i, [C.POINTER(Metric), _vp, _vp]),
    ("geon_metric_g_inv", _i, [C.POINTER(Metric), _vp, _vp]),
    ("geon_metric_christoffel", _i, [C.POINTER(Metric), _vp, _vp]),
    ("geon_metric_terminate", _i, [C.POINTER(Metric), _vp, C.POINTER(C.c_int32)]),
    ("geon_metric_asymptotic_direction", _i, [C.POINTER(Metric), _vp, _vp, _vp]),
    ("geon_camera_new", _i, [_vp, C.c_double, _u32, _u32, C.POINTER(Camera)]),
    ("geon_sky_sample", _i, [_vp, C.c_int32, _vp]),
    ("geon_ctx_create", _i, [_i, C.c_char_p, C.POINTER(_vp)]),
    ("geon_ctx_destroy", None, [_vp]),
    ("geon_tetrad", _i, [C.POINTER(Metric), _vp, _vp]),
    ("geon_ray_state", _i, [C.POINTER(Camera), _vp, _u32, _u32, C.c_double, C.c_double, _vp]),
    ("geon_trace_rays", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_double, _u32,
                             _vp, _vp, _vp, _vp, _vp]),
    ("geon_trace_rays_dev", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_double, _u32,
                                 _vp, _vp, _vp, _vp, _vp, _vp]),
    ("geon_render", _i, [_vp, C.POINTER(Metric), C.POINTER(Camera), C.POINTER(RenderCfg),
                         _vp, _vp, _vp, C.POINTER(RenderStats)]),
    ("geon_render_dev", _i, [_vp, C.POINTER(Metric), C.POINTER(Camera), C.POINTER(RenderCfg),
                             _vp, _vp, _vp, C.POINTER(RenderStats), _vp]),
    ("geon_render_dev_noalloc", _i, [_vp, C.POINTER(Metric), C.POINTER(Camera),
                                     C.POINTER(RenderCfg), _vp, _vp, _vp, C.POINTER(RenderStats),
                                     _vp]),
    ("geon_render_dev_noalloc_max_rays", C.c_uint64, []),
    ("geon_curvature", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_int32, _vp]),
    ("geon_curvature_dev", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_int32, _vp, _vp]),
    ("geon_nec_contraction", _i, [_vp, C.POINTER(Metric), _vp, _sz, _vp, _vp]),
    ("geon_trace_rays_adaptive", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_double, C.c_double,
                                      C.c_double, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    ("geon_trace_rays_adaptive_dev", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_double, C.c_double,
                                          C.c_double, C.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("geon_nec_contraction_dev", _i, [_vp, C.POINTER(Metric), _vp, _sz, _vp, _vp, _vp]),
    ("geon_nec_scan", _i, [_vp, C.POINTER(Metric), _vp, _sz, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    ("geon_shadow_bisect", _i, [_vp, _vp, _vp, _sz, C.c_double, _u32, _u32, _vp, _vp]),
    ("geon_tonemap_u8", _i, [_vp, _sz, _vp]),
    ("geon_batch_create", _i, [_vp, C.POINTER(_vp)]),
    ("geon_batch_destroy", None, [_vp]),
    ("geon_batch_add_frame", _i, [_vp, C.POINTER(Metric), C.POINTER(Camera), C.POINTER(RenderCfg),
                                  _vp, _vp, _vp, C.POINTER(_u32)]),
    ("geon_batch_add_rays", _i, [_vp, C.POINTER(Metric), _vp, _sz, C.c_double, _u32, C.c_int32,
                                 C.c_int32, _vp, _vp, _vp, _vp, _vp, C.POINTER(_u32)]),
    ("geon_batch_run", _i, [_vp, _vp, _vp]),
    ("geon_batch_frame_stats", _i, [_vp, _u32, C.POINTER(RenderStats)]),
    ("geon_batch_totals", _i, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                               C.POINTER(C.c_double)]),
    ("geon_autotune", _i, [_vp, C.c_int32, C.c_int32, _vp]),
    ("geon_serve_create", _i, [_vp, C.POINTER(ServeCfg), C.POINTER(_vp)]),
    ("geon_serve_destroy", None, [_vp]),
    ("geon_serve_submit_render", _i, [_vp, C.POINTER(Metric), C.POINTER(Camera),
                                      C.POINTER(RenderCfg), C.c_int32, C.POINTER(C.c_uint64)]),
    ("geon_serve_submit_bisect", _i, [_vp, _vp, _vp, _sz, C.c_double, _u32, _u32, C.c_int32,
                                      C.POINTER(C.c_uint64)]),
    ("geon_serve_poll", _i, [_vp, C.c_uint64, C.POINTER(ServeResult)]),
    ("geon_serve_fetch_bisect", _i, [_vp, C.c_uint64, _vp, _vp]),
    ("geon_serve_fetch_render", _i, [_vp, C.c_uint64, _vp, _vp, _vp]),
    ("geon_serve_release", _i, [_vp, C.c_uint64]),
    ("geon_serve_drain", _i, [_vp]),
    ("geon_serve_counters", _i, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
]

_lib = None


def _one_hip_runtime():
    """A process must use a single HIP runtime. PyTorch wheels bundle their own
    libamdhip64.so.7; the dynamic linker binds every later request for that
    soname to whichever copy was loaded first. When torch is installed, load
    its copy (without importing torch) so that libgeon_hip.so and torch share
    it whatever the import order."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    for d in spec.submodule_search_locations:
        cand = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(cand):
            C.CDLL(cand, mode=C.RTLD_GLOBAL)
            return


def lib():
    """Load libgeon_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m geon_amd.build`")
        _one_hip_runtime()
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def strerror(status: int) -> str:
    return lib().geon_strerror(status).decode()


def check(status: int, where: str) -> None:
    if status != OK:
        detail = lib().geon_last_error().decode()
        raise GeonError(status, where, detail)
