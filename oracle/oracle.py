"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker. The product
package ``geon_amd`` never imports it.

The oracle restates the Rust reference (Arth-Singh/geon ``src/``) in plain C;
see the file:line citations in ``geon_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MINKOWSKI, SCHWARZSCHILD, ALCUBIERRE, MORRIS_THORNE = 0, 1, 2, 3
DIVERGED, CAPTURED, ESCAPED, MAXSTEPS = 0, 1, 2, 3
CURV_RIEMANN, CURV_RICCI, CURV_EINSTEIN, CURV_STRESS, CURV_SCALAR = 0, 1, 2, 3, 4


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


class RenderStats(C.Structure):
    _fields_ = [("captured", C.c_uint64), ("escaped", C.c_uint64), ("diverged", C.c_uint64),
                ("max_null_residual", C.c_double), ("avg_steps", C.c_double),
                ("max_steps_hit", C.c_uint64), ("ray_steps", C.c_uint64),
                ("fp64_rays", C.c_uint64), ("kernel_ms", C.c_double),
                ("gamma_steps", C.c_uint64)]


class TraceResult(C.Structure):
    _fields_ = [("cls", C.c_int32), ("reserved", C.c_int32), ("steps", C.c_uint64),
                ("y", C.c_double * 8), ("null_residual", C.c_double)]


_lib = None

# worker threads for the batched oracle calls (the GPU box grants 16 cores)
THREADS = max(1, min(16, os.cpu_count() or 1))


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.or_ricci_scalar.restype = C.c_double
        L.or_nec_contraction.restype = C.c_double
        L.or_dot4.restype = C.c_double
        L.or_trace_rays.restype = C.c_double
        L.or_curvature.restype = C.c_double
        L.or_alcubierre_nec_integral.restype = C.c_double
        L.or_trace.argtypes = [C.POINTER(Metric), dp, C.c_double, C.c_uint64,
                               C.POINTER(TraceResult)]
        L.or_rk4_step.argtypes = [C.POINTER(Metric), dp, C.c_double, dp]
        L.or_trace_rays.argtypes = [C.POINTER(Metric), dp, C.c_size_t, C.c_double, C.c_uint64,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
        L.or_ray_state_sub.argtypes = [C.POINTER(Camera), dp, C.c_uint64, C.c_uint64,
                                       C.c_double, C.c_double, dp]
        L.or_render.argtypes = [C.POINTER(Metric), C.POINTER(Camera), C.POINTER(RenderCfg),
                                C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(RenderStats)]
        L.or_curvature.argtypes = [C.POINTER(Metric), dp, C.c_size_t, C.c_int32, C.c_int, dp]
        L.or_is_captured.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64]
        L.or_validate_shadow.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                         dp, dp]
        L.or_equatorial_ray.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def metric(kind: int, **kw) -> Metric:
    m = Metric()
    lib().or_metric_default(C.c_int32(kind), C.byref(m))
    names = {SCHWARZSCHILD: ["mass"], ALCUBIERRE: ["bubble_radius", "bubble_thickness",
                                                   "ship_velocity", "ship_x"],
             MORRIS_THORNE: ["throat_radius"], MINKOWSKI: []}[kind]
    for k, v in kw.items():
        if k in names:
            m.p[names.index(k)] = float(v)
        elif k in ("horizon_pad", "escape_radius", "fd_step"):
            setattr(m, k, float(v))
        else:
            raise KeyError(k)
    return m


def camera(observer, fov_y_deg, width, height) -> Camera:
    """Camera::new (camera.rs:29-36): the field of view in degrees, held in radians."""
    c = Camera()
    obs = np.ascontiguousarray(observer, dtype=np.float64)
    lib().or_camera_new(_dp(obs), C.c_double(fov_y_deg), C.c_uint32(int(width)),
                        C.c_uint32(int(height)), C.byref(c))
    return c


def render_cfg(step=0.05, max_steps=20000, samples_per_axis=1, show_stars=True,
               precision=0) -> RenderCfg:
    r = RenderCfg()
    r.step, r.max_steps, r.samples_per_axis = step, max_steps, samples_per_axis
    r.captured_rgb[:] = [0.0, 0.0, 0.0]
    r.diverged_rgb[:] = [1.0, 0.0, 1.0]
    r.show_stars, r.precision = int(bool(show_stars)), precision
    return r


def _x4(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def dot4(gm, a, b):
    """types.rs:22-31."""
    g16 = np.ascontiguousarray(gm, dtype=np.float64).reshape(16)
    L = lib()
    L.or_dot4.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return L.or_dot4(g16.ctypes.data, _x4(a).ctypes.data, _x4(b).ctypes.data)


def g(m, x):
    out = np.zeros(16)
    lib().or_g(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4)


def dg(m, x):
    out = np.zeros(64)
    lib().or_dg(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4, 4)


def g_inv(m, x):
    out = np.zeros(16)
    st = lib().or_g_inv(C.byref(m), _dp(_x4(x)), _dp(out))
    return st, out.reshape(4, 4)


def inv4(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(16)
    out = np.zeros(16)
    st = lib().or_inv4(_dp(a), _dp(out))
    return st, out.reshape(4, 4)


def christoffel(m, x):
    out = np.zeros(64)
    lib().or_christoffel(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4, 4)


def terminate(m, x):
    return lib().or_terminate(C.byref(m), _dp(_x4(x)))


def asymptotic_direction(m, x, p):
    out = np.zeros(3)
    lib().or_asymptotic_direction(C.byref(m), _dp(_x4(x)), _dp(_x4(p)), _dp(out))
    return out


def rhs(m, y):
    out = np.zeros(8)
    lib().or_rhs(C.byref(m), _dp(_x4(y)), _dp(out))
    return out


def rk4_step(m, y, h):
    out = np.zeros(8)
    lib().or_rk4_step(C.byref(m), _dp(_x4(y)), h, _dp(out))
    return out


def trace(m, y0, step, max_steps):
    r = TraceResult()
    lib().or_trace(C.byref(m), _dp(_x4(y0)), step, max_steps, C.byref(r))
    return r.cls, r.steps, np.array(r.y[:]), r.null_residual


def trace_rays(m, y0_aos, step, max_steps, nthreads=THREADS):
    y0 = np.ascontiguousarray(y0_aos, dtype=np.float64)
    n = y0.shape[0]
    cls = np.zeros(n, np.int32)
    steps = np.zeros(n, np.uint64)
    yf = np.zeros((n, 8))
    resid = np.zeros(n)
    d = np.zeros((n, 3))
    el = lib().or_trace_rays(C.byref(m), _dp(y0), n, step, max_steps, nthreads,
                             cls.ctypes.data, steps.ctypes.data, yf.ctypes.data,
                             resid.ctypes.data, d.ctypes.data)
    return dict(cls=cls, steps=steps, y=yf, resid=resid, dir=d, seconds=el)


def ck_step(m, y, h):
    """One Cash–Karp step: (fifth-order solution, embedded error estimate)."""
    L = lib()
    L.or_ck_step.argtypes = [C.POINTER(Metric), C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
    L.or_ck_step.restype = None
    y = np.ascontiguousarray(y, dtype=np.float64)
    yn = np.zeros(8)
    err = np.zeros(8)
    L.or_ck_step(C.byref(m), y.ctypes.data, h, yn.ctypes.data, err.ctypes.data)
    return yn, err


def trace_rays_adaptive(m, y0_aos, h0, rtol, atol, max_steps):
    """Adaptive Cash–Karp RK4(5) trace (serial; no reference counterpart)."""
    y0 = np.ascontiguousarray(y0_aos, dtype=np.float64).reshape(-1, 8)
    n = y0.shape[0]
    cls = np.zeros(n, np.int32)
    steps = np.zeros(n, np.uint64)
    rej = np.zeros(n, np.uint64)
    yf = np.zeros((n, 8))
    d = np.zeros((n, 3))
    L = lib()
    L.or_trace_adaptive_rays.argtypes = [C.POINTER(Metric), C.c_void_p, C.c_size_t, C.c_double,
                                         C.c_double, C.c_double, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_trace_adaptive_rays(C.byref(m), y0.ctypes.data, n, h0, rtol, atol, max_steps,
                             cls.ctypes.data, steps.ctypes.data, rej.ctypes.data,
                             yf.ctypes.data, d.ctypes.data)
    return dict(cls=cls, steps=steps, rejected=rej, y=yf, dir=d)


def tetrad(m, observer):
    out = np.zeros(16)
    st = lib().or_tetrad(C.byref(m), _dp(_x4(observer)), _dp(out))
    return st, out.reshape(4, 4)


def ray_state_sub(cam, tet, i, j, du=0.5, dv=0.5):
    y = np.zeros(8)
    t = np.ascontiguousarray(tet, dtype=np.float64).reshape(16)
    lib().or_ray_state_sub(C.byref(cam), _dp(t), i, j, du, dv, _dp(y))
    return y


def camera_rays(m, cam):
    """Initial states of every pixel centre, AoS (W·H)×8 in row-major order."""
    st, tet = tetrad(m, cam.observer)
    if st != 0:
        raise RuntimeError(f"tetrad failed: {st}")
    n = cam.width * cam.height
    out = np.zeros((n, 8))
    for idx in range(n):
        out[idx] = ray_state_sub(cam, tet, idx % cam.width, idx // cam.width)
    return out


def sky_sample(d, show_stars=True):
    out = np.zeros(3, np.float32)
    lib().or_sky_sample(_dp(_x4(d)), int(show_stars), out.ctypes.data_as(C.c_void_p))
    return out


def tonemap_u8(px):
    px = np.ascontiguousarray(px, dtype=np.float32)
    out = np.zeros(px.size, np.uint8)
    lib().or_tonemap_u8(px.ctypes.data_as(C.c_void_p), C.c_size_t(px.size),
                        out.ctypes.data_as(C.c_void_p))
    return out.reshape(px.shape)


def render(m, cam, cfg, nthreads=THREADS):
    n = cam.width * cam.height
    rgb = np.zeros((n, 3), np.float32)
    cls = np.zeros(n, np.int32)
    d = np.zeros((n, 3))
    steps = np.zeros(n, np.uint64)
    stats = RenderStats()
    st = lib().or_render(C.byref(m), C.byref(cam), C.byref(cfg), nthreads, rgb.ctypes.data,
                         cls.ctypes.data, d.ctypes.data, steps.ctypes.data, C.byref(stats))
    if st != 0:
        raise RuntimeError(f"render failed: {st}")
    return dict(rgb=rgb, cls=cls, dir=d, steps=steps, stats=stats)


def riemann(m, x):
    out = np.zeros(256)
    lib().or_riemann(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4, 4, 4)


def ricci(m, x):
    out = np.zeros(16)
    lib().or_ricci(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4)


def ricci_scalar(m, x):
    return lib().or_ricci_scalar(C.byref(m), _dp(_x4(x)))


def einstein(m, x):
    out = np.zeros(16)
    lib().or_einstein(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4)


def stress_energy(m, x):
    out = np.zeros(16)
    lib().or_stress_energy(C.byref(m), _dp(_x4(x)), _dp(out))
    return out.reshape(4, 4)


def nec_contraction(m, x, k):
    return lib().or_nec_contraction(C.byref(m), _dp(_x4(x)), _dp(_x4(k)))


def curvature(m, x_aos, what, nthreads=THREADS):
    x = np.ascontiguousarray(x_aos, dtype=np.float64)
    n = x.shape[0]
    width = {CURV_RIEMANN: 256, CURV_SCALAR: 1}.get(what, 16)
    out = np.zeros((n, width))
    el = lib().or_curvature(C.byref(m), _dp(x), n, what, nthreads, _dp(out))
    return out, el


def set_tanh_ulp_shift(k: int) -> None:
    """Conditioning tests only: move every tanh of the Alcubierre shape function
    by one ulp towards ±∞ (k = ±1), or each evaluation up or down by one ulp as
    a hash of its argument and k decides (|k| ≥ 2); 0 restores the reference's
    arithmetic."""
    lib().or_set_tanh_ulp_shift(C.c_int(int(k)))


def equatorial_ray(mass, r_obs, alpha):
    y = np.zeros(8)
    lib().or_equatorial_ray(mass, r_obs, alpha, _dp(y))
    return y


def is_captured(mass, r_obs, alpha, step=0.01, max_steps=200_000):
    return bool(lib().or_is_captured(mass, r_obs, alpha, step, max_steps))


def validate_shadow(mass=1.0, r_obs=20.0, step=0.01, max_steps=200_000, iters=60):
    a = C.c_double()
    b = C.c_double()
    lib().or_validate_shadow(mass, r_obs, step, max_steps, iters, C.byref(a), C.byref(b))
    return a.value, b.value


def alcubierre_nec_integral(m):
    return lib().or_alcubierre_nec_integral(C.byref(m))
