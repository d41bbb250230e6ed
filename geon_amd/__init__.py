"""geon_amd — MI355X (gfx950) geodesic ray tracer and curvature engine.

Python face of the C ABI in include/geon_hip.h, named after the reference
crate's public surface (Arth-Singh/geon src/lib.rs:18-30): metric types with
the reference field names and defaults, Camera, RenderConfig, RenderStats,
render, trace, and the curvature functions. All compute runs in the HIP
code object; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (ALCUBIERRE, CAPTURED, CURV_EINSTEIN, CURV_RICCI, CURV_RIEMANN, CURV_SCALAR,
                   CURV_STRESS, DIVERGED, ESCAPED, MAXSTEPS, MINKOWSKI, MORRIS_THORNE,
                   PREC_FP64, PREC_MIXED, SCHWARZSCHILD, GeonError, check, lib)

__all__ = ["Minkowski", "Schwarzschild", "Alcubierre", "MorrisThorne", "Camera", "RenderConfig",
           "Context", "render", "ray_states", "trace_rays", "trace_rays_adaptive", "riemann", "ricci", "ricci_scalar", "einstein",
           "stress_energy", "nec_contraction", "nec_scan", "shadow_bisect", "tetrad", "GeonError", "Batch",
           "autotune", "Server", "tonemap_u8", "write_png"]


# ---- metrics (src/metric/*.rs) ----------------------------------------------

@dataclass
class _MetricBase:
    horizon_pad: float = 1.001
    escape_radius: float = 1000.0
    fd_step: float = 1e-5

    KIND = -1

    def _params(self):
        return []

    def to_c(self) -> _abi.Metric:
        m = _abi.Metric()
        m.kind = self.KIND
        for i, v in enumerate(self._params()):
            m.p[i] = float(v)
        m.horizon_pad = float(self.horizon_pad)
        m.escape_radius = float(self.escape_radius)
        m.fd_step = float(self.fd_step)
        return m


@dataclass
class Minkowski(_MetricBase):
    """minkowski.rs — flat Cartesian chart, escape x²+y²+z² > escape_radius²."""
    KIND = MINKOWSKI


@dataclass
class Schwarzschild(_MetricBase):
    """schwarzschild.rs — Schwarzschild::new(mass): horizon_pad 1.001, escape 1000 M."""
    mass: float = 1.0
    KIND = SCHWARZSCHILD

    def _params(self):
        return [self.mass]


@dataclass
class Alcubierre(_MetricBase):
    """alcubierre.rs — Alcubierre::default(): R=2, s=0.5, v=2, ship_x=0."""
    bubble_radius: float = 2.0
    bubble_thickness: float = 0.5
    ship_velocity: float = 2.0
    ship_x: float = 0.0
    KIND = ALCUBIERRE

    def _params(self):
        return [self.bubble_radius, self.bubble_thickness, self.ship_velocity, self.ship_x]


@dataclass
class MorrisThorne(_MetricBase):
    """morris_thorne.rs — MorrisThorne::default(): b0 = 1."""
    throat_radius: float = 1.0
    KIND = MORRIS_THORNE

    def _params(self):
        return [self.throat_radius]


def _cmetric(m):
    if isinstance(m, _MetricBase):
        return m.to_c()
    if isinstance(m, _abi.Metric):
        return m
    return _abi.Metric.from_buffer_copy(bytes(m))  # same-layout structure from another binding


# ---- camera / config (camera.rs, render.rs) ----------------------------------

@dataclass
class Camera:
    """Camera::new(observer, fov_y_deg, width, height) (camera.rs:29-36): the
    field of view is given in degrees and held in radians (``fov_y_rad``), as
    the reference's Camera holds it."""
    observer: tuple
    fov_y_deg: float
    width: int
    height: int

    @property
    def fov_y_rad(self) -> float:
        return self.to_c().fov_y_rad

    def to_c(self) -> _abi.Camera:
        c = _abi.Camera()
        obs = (C.c_double * 4)(*[float(v) for v in self.observer])
        check(lib().geon_camera_new(obs, float(self.fov_y_deg), int(self.width), int(self.height),
                                    C.byref(c)), "geon_camera_new")
        return c


@dataclass
class RenderConfig:
    """RenderConfig (render.rs:17-41) plus the sky toggle and precision."""
    step: float = 0.05
    max_steps: int = 20_000
    captured_rgb: tuple = (0.0, 0.0, 0.0)
    diverged_rgb: tuple = (1.0, 0.0, 1.0)
    samples_per_axis: int = 1
    show_stars: bool = True
    precision: int = PREC_FP64
    fp64_band: float = 0.0

    def to_c(self) -> _abi.RenderCfg:
        r = _abi.RenderCfg()
        r.step, r.max_steps, r.samples_per_axis = float(self.step), int(self.max_steps), \
            int(self.samples_per_axis)
        r.captured_rgb[:] = list(self.captured_rgb)
        r.diverged_rgb[:] = list(self.diverged_rgb)
        r.show_stars, r.precision = int(bool(self.show_stars)), int(self.precision)
        r.fp64_band = float(self.fp64_band)
        return r


# ---- context -------------------------------------------------------------------

class Context:
    """One device, one code object, one stream (geon_ctx_create)."""

    def __init__(self, device: int = 0, code_object: str | None = None):
        h = C.c_void_p()
        path = (code_object or _abi.HSACO_PATH).encode()
        check(lib().geon_ctx_create(device, path, C.byref(h)), "geon_ctx_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().geon_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def tetrad(metric, observer):
    """Camera::tetrad (camera.rs:47-97): rows e_(a)^μ."""
    out = np.zeros(16)
    obs = np.ascontiguousarray(observer, dtype=np.float64)
    check(lib().geon_tetrad(C.byref(_cmetric(metric)), _ptr(obs), _ptr(out)), "geon_tetrad")
    return out.reshape(4, 4)


def ray_states(metric, camera: Camera, du=0.5, dv=0.5):
    """Initial states of every pixel of the camera (Camera::tetrad, then
    Camera::ray at (i + du, j + dv), camera.rs:47-160), AoS (W·H)×8 in
    row-major pixel order."""
    tet = tetrad(metric, camera.observer)
    cam = camera.to_c()
    n = camera.width * camera.height
    out = np.zeros((n, 8))
    f, pc, pt, base = lib().geon_ray_state, C.byref(cam), _ptr(tet), out.ctypes.data
    du, dv, k = float(du), float(dv), 0
    for j in range(camera.height):
        for i in range(camera.width):
            rc = f(pc, pt, i, j, du, dv, base + 64 * k)
            if rc:
                check(rc, "geon_ray_state")
            k += 1
    return out


def trace_rays(metric, y0, step, max_steps, ctx: Context | None = None):
    """Trace initial states y0 (n×8, AoS as numpy) on the GPU (geodesic::trace per ray).

    Returns dict(cls, steps, y, resid, dir) with y n×8 and dir n×3.
    """
    ctx = ctx or default_context()
    y0 = np.ascontiguousarray(np.asarray(y0, dtype=np.float64).reshape(-1, 8).T)  # SoA
    n = y0.shape[1]
    cls = np.zeros(n, np.int32)
    steps = np.zeros(n, np.uint32)
    yf = np.zeros((8, n))
    resid = np.zeros(n)
    d = np.zeros((3, n))
    check(lib().geon_trace_rays(ctx.handle, C.byref(_cmetric(metric)), _ptr(y0), n, float(step),
                                int(max_steps), _ptr(cls), _ptr(steps), _ptr(yf), _ptr(resid),
                                _ptr(d)), "geon_trace_rays")
    return dict(cls=cls, steps=steps, y=yf.T.copy(), resid=resid, dir=d.T.copy())


def trace_rays_adaptive(metric, y0, h0, rtol, atol, max_steps, ctx: Context | None = None):
    """trace_rays with adaptive Cash–Karp RK4(5) steps (the adaptive stepping
    integrator.rs:5-7 names as the next step; not a reference method).

    Returns dict(cls, steps, rejected, y, dir) with y n×8 and dir n×3.
    """
    ctx = ctx or default_context()
    y0 = np.ascontiguousarray(np.asarray(y0, dtype=np.float64).reshape(-1, 8).T)  # SoA
    n = y0.shape[1]
    cls = np.zeros(n, np.int32)
    steps = np.zeros(n, np.uint32)
    rej = np.zeros(n, np.uint32)
    yf = np.zeros((8, n))
    d = np.zeros((3, n))
    check(lib().geon_trace_rays_adaptive(ctx.handle, C.byref(_cmetric(metric)), _ptr(y0), n,
                                         float(h0), float(rtol), float(atol), int(max_steps),
                                         _ptr(cls), _ptr(steps), _ptr(rej), _ptr(yf), _ptr(d)),
          "geon_trace_rays_adaptive")
    return dict(cls=cls, steps=steps, rejected=rej, y=yf.T.copy(), dir=d.T.copy())


def render(metric, camera: Camera, cfg: RenderConfig | None = None, ctx: Context | None = None):
    """render::render (render.rs:55-165): returns (rgb W·H×3 float32, RenderStats dict,
    per-pixel class W·H, per-pixel direction W·H×3)."""
    ctx = ctx or default_context()
    cfg = cfg or RenderConfig()
    n = camera.width * camera.height
    rgb = np.zeros((n, 3), np.float32)
    cls = np.zeros(n, np.int32)
    d = np.zeros((3, n))
    st = _abi.RenderStats()
    check(lib().geon_render(ctx.handle, C.byref(_cmetric(metric)), C.byref(camera.to_c()),
                            C.byref(cfg.to_c()), _ptr(rgb), _ptr(cls), _ptr(d), C.byref(st)),
          "geon_render")
    return rgb, st.as_dict(), cls, d.T.copy()


def render_dev(metric, camera: Camera, cfg: RenderConfig, rgb_ptr, cls_ptr, dir_ptr,
               stream_ptr=None, want_stats=True, ctx: Context | None = None):
    """Device-pointer render (geon_render_dev): outputs stay in HBM."""
    ctx = ctx or default_context()
    st = _abi.RenderStats()
    check(lib().geon_render_dev(ctx.handle, C.byref(_cmetric(metric)), C.byref(camera.to_c()),
                                C.byref(cfg.to_c()), C.c_void_p(rgb_ptr), C.c_void_p(cls_ptr),
                                C.c_void_p(dir_ptr), C.byref(st) if want_stats else None,
                                C.c_void_p(stream_ptr) if stream_ptr else None),
          "geon_render_dev")
    return st.as_dict()


class Batch:
    """Persistent-scheduler batch (geon_batch_*): camera frames and explicit ray
    sets of any metric kinds and precisions share persistent launches.
    Output arguments are device pointers (ints, e.g. torch ``tensor.data_ptr()``)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().geon_batch_create(self.ctx.handle, C.byref(h)), "geon_batch_create")
        self.handle = h

    def add_frame(self, metric, camera: Camera, cfg: RenderConfig, rgb=0, cls=0, dir=0) -> int:
        fid = C.c_uint32()
        check(lib().geon_batch_add_frame(self.handle, C.byref(_cmetric(metric)),
                                         C.byref(camera.to_c()), C.byref(cfg.to_c()),
                                         C.c_void_p(rgb or None), C.c_void_p(cls or None),
                                         C.c_void_p(dir or None), C.byref(fid)),
              "geon_batch_add_frame")
        return fid.value

    def add_rays(self, metric, y0_soa_ptr, n, step, max_steps, precision=PREC_FP64,
                 early_exit=False, cls=0, steps=0, yfin=0, resid=0, dir=0) -> int:
        jid = C.c_uint32()
        check(lib().geon_batch_add_rays(self.handle, C.byref(_cmetric(metric)),
                                        C.c_void_p(y0_soa_ptr), n, float(step), int(max_steps),
                                        int(precision), int(bool(early_exit)),
                                        C.c_void_p(cls or None), C.c_void_p(steps or None),
                                        C.c_void_p(yfin or None), C.c_void_p(resid or None),
                                        C.c_void_p(dir or None), C.byref(jid)),
              "geon_batch_add_rays")
        return jid.value

    def run(self, tune: dict | None = None, stream_ptr=None):
        cfg = None
        if tune:
            cfg = _abi.SchedCfg(**{k: int(v) for k, v in tune.items()})
        check(lib().geon_batch_run(self.handle, C.byref(cfg) if cfg else None,
                                   C.c_void_p(stream_ptr) if stream_ptr else None),
              "geon_batch_run")

    def frame_stats(self, fid: int) -> dict:
        st = _abi.RenderStats()
        check(lib().geon_batch_frame_stats(self.handle, fid, C.byref(st)), "geon_batch_frame_stats")
        return st.as_dict()

    def totals(self) -> dict:
        rs, fb, ms = C.c_uint64(), C.c_uint64(), C.c_double()
        check(lib().geon_batch_totals(self.handle, C.byref(rs), C.byref(fb), C.byref(ms)),
              "geon_batch_totals")
        return dict(executed_ray_steps=rs.value, fp64_rays=fb.value, device_ms=ms.value)

    def close(self):
        if self.handle:
            lib().geon_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (geon_serve_*): a native host thread admits
    requests by HBM footprint and runs all admitted rays together in quanta."""

    def __init__(self, ctx: Context | None = None, hbm_budget: int = 0, quantum_steps: int = 0,
                 max_jobs: int = 0):
        self.ctx = ctx or default_context()
        cfg = _abi.ServeCfg(hbm_budget=hbm_budget, quantum_steps=quantum_steps, max_jobs=max_jobs)
        h = C.c_void_p()
        check(lib().geon_serve_create(self.ctx.handle, C.byref(cfg), C.byref(h)), "geon_serve_create")
        self.handle = h

    def submit_render(self, metric, camera: Camera, cfg: RenderConfig, keep=False) -> int:
        t = C.c_uint64()
        check(lib().geon_serve_submit_render(self.handle, C.byref(_cmetric(metric)),
                                             C.byref(camera.to_c()), C.byref(cfg.to_c()),
                                             int(bool(keep)), C.byref(t)), "geon_serve_submit_render")
        return t.value

    def submit_bisect(self, mass, r_obs, step=0.01, max_steps=200_000, iters=40,
                      full_trace=False) -> int:
        m = np.ascontiguousarray(np.atleast_1d(mass), dtype=np.float64)
        r = np.ascontiguousarray(np.atleast_1d(r_obs), dtype=np.float64)
        t = C.c_uint64()
        check(lib().geon_serve_submit_bisect(self.handle, _ptr(m), _ptr(r), m.size, float(step),
                                             int(max_steps), int(iters), int(bool(full_trace)),
                                             C.byref(t)),
              "geon_serve_submit_bisect")
        return t.value

    def poll(self, ticket: int):
        """None while pending, else the result dict."""
        res = _abi.ServeResult()
        st = lib().geon_serve_poll(self.handle, ticket, C.byref(res))
        if st == _abi.E_PENDING:
            return None
        check(st, "geon_serve_poll")
        return dict(kind=res.kind, status=res.status, latency_ms=res.latency_ms,
                    queue_ms=res.queue_ms, stats=res.stats.as_dict(), ray_steps=res.ray_steps)

    def fetch_bisect(self, ticket: int, n: int):
        a, b = np.zeros(n), np.zeros(n)
        check(lib().geon_serve_fetch_bisect(self.handle, ticket, _ptr(a), _ptr(b)),
              "geon_serve_fetch_bisect")
        return a, b

    def fetch_render(self, ticket: int, width: int, height: int):
        n = width * height
        rgb, cls, d = np.zeros((n, 3), np.float32), np.zeros(n, np.int32), np.zeros((3, n))
        check(lib().geon_serve_fetch_render(self.handle, ticket, _ptr(rgb), _ptr(cls), _ptr(d)),
              "geon_serve_fetch_render")
        return rgb, cls, d.T.copy()

    def release(self, ticket: int):
        check(lib().geon_serve_release(self.handle, ticket), "geon_serve_release")

    def drain(self):
        check(lib().geon_serve_drain(self.handle), "geon_serve_drain")

    def counters(self) -> dict:
        q, d, r, b = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_double()
        check(lib().geon_serve_counters(self.handle, C.byref(q), C.byref(d), C.byref(r),
                                        C.byref(b)), "geon_serve_counters")
        return dict(quanta=q.value, done=d.value, ray_steps=r.value, busy_ms=b.value)

    def close(self):
        if self.handle:
            lib().geon_serve_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(kind: int, precision: int = PREC_MIXED, ctx: Context | None = None) -> dict:
    """Time launch shapes on a calibration frame; the winner becomes the default."""
    ctx = ctx or default_context()
    cfg = _abi.SchedCfg()
    check(lib().geon_autotune(ctx.handle, int(kind), int(precision), C.byref(cfg)), "geon_autotune")
    return {k: getattr(cfg, k) for k, _ in cfg._fields_}


def curvature(metric, x, what, ctx: Context | None = None):
    """Curvature at points x (n×4): what = CURV_RIEMANN|RICCI|EINSTEIN|STRESS|SCALAR."""
    ctx = ctx or default_context()
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 4).T)
    n = xs.shape[1]
    w = {CURV_RIEMANN: 256, CURV_SCALAR: 1}.get(what, 16)
    out = np.zeros((n, w))
    check(lib().geon_curvature(ctx.handle, C.byref(_cmetric(metric)), _ptr(xs), n, what,
                               _ptr(out)), "geon_curvature")
    return out


def riemann(metric, x, ctx=None):
    return curvature(metric, x, CURV_RIEMANN, ctx).reshape(-1, 4, 4, 4, 4)


def ricci(metric, x, ctx=None):
    return curvature(metric, x, CURV_RICCI, ctx).reshape(-1, 4, 4)


def ricci_scalar(metric, x, ctx=None):
    return curvature(metric, x, CURV_SCALAR, ctx).reshape(-1)


def einstein(metric, x, ctx=None):
    return curvature(metric, x, CURV_EINSTEIN, ctx).reshape(-1, 4, 4)


def stress_energy(metric, x, ctx=None):
    return curvature(metric, x, CURV_STRESS, ctx).reshape(-1, 4, 4)


def nec_contraction(metric, x, k, ctx=None):
    """T_μν k^μ k^ν (curvature.rs:127-136) at each point, on the device: T
    from the curvature kernel, the contraction summed in the reference's
    order. k is one vector or one per point."""
    ctx = ctx or default_context()
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 4).T)
    n = xs.shape[1]
    k = np.asarray(k, dtype=np.float64).reshape(-1, 4)
    out = np.zeros(n)
    if k.shape[0] == 1:
        kk = np.ascontiguousarray(k[0])
        check(lib().geon_nec_contraction(ctx.handle, C.byref(_cmetric(metric)), _ptr(xs), n,
                                         _ptr(kk), _ptr(out)), "geon_nec_contraction")
        return out
    if k.shape[0] != n:
        raise ValueError("k: one vector or one per point")
    for row in np.unique(k, axis=0):                # one launch per distinct k
        sel = np.nonzero((k == row).all(axis=1))[0]
        xsel = np.ascontiguousarray(xs[:, sel])
        kk = np.ascontiguousarray(row)
        part = np.zeros(sel.size)
        check(lib().geon_nec_contraction(ctx.handle, C.byref(_cmetric(metric)), _ptr(xsel),
                                         sel.size, _ptr(kk), _ptr(part)), "geon_nec_contraction")
        out[sel] = part
    return out


def nec_scan(metric, x, k, dx=0.0, ctx=None):
    """The energy-condition line scans of bin/energy_conditions.rs:70-153 over
    the points x (in order) for one null vector k: returns (nec per point,
    first strict minimum, its index, Σ nec·dx over the negative values)."""
    ctx = ctx or default_context()
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, 4).T)
    n = xs.shape[1]
    kk = np.ascontiguousarray(np.asarray(k, dtype=np.float64).reshape(4))
    nec = np.zeros(n)
    mv, mi, integ = C.c_double(), C.c_uint64(), C.c_double()
    check(lib().geon_nec_scan(ctx.handle, C.byref(_cmetric(metric)), _ptr(xs), n, _ptr(kk),
                              float(dx), _ptr(nec), C.byref(mv), C.byref(mi), C.byref(integ)),
          "geon_nec_scan")
    return nec, mv.value, int(mi.value), integ.value


def shadow_bisect(mass, r_obs, step=0.01, max_steps=200_000, iters=60, ctx=None):
    """validate_shadow.rs bisection for each (mass[k], r_obs[k]) → (alpha_c, b_c)."""
    ctx = ctx or default_context()
    m = np.ascontiguousarray(np.atleast_1d(mass), dtype=np.float64)
    r = np.ascontiguousarray(np.atleast_1d(r_obs), dtype=np.float64)
    n = m.size
    a = np.zeros(n)
    b = np.zeros(n)
    check(lib().geon_shadow_bisect(ctx.handle, _ptr(m), _ptr(r), n, float(step), int(max_steps),
                                   int(iters), _ptr(a), _ptr(b)), "geon_shadow_bisect")
    return a, b


def tonemap_u8(rgb):
    """write_png's per-channel map (render.rs:175-179)."""
    px = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros(px.size, np.uint8)
    check(lib().geon_tonemap_u8(_ptr(px), px.size, _ptr(out)), "geon_tonemap_u8")
    return out.reshape(px.shape)


def write_png(path, pixels, width: int, height: int):
    """render::write_png (render.rs:169-184): each linear channel through
    write_png's tonemap, written as an 8-bit RGB PNG, rows top to bottom. The
    decoded pixels are the reference's bytes; the compressed stream is
    zlib's, which the image crate's encoder need not match byte for byte.
    A buffer whose size is not 3·width·height is refused, as the reference's
    `from_raw(...).expect` refuses it."""
    import struct
    import zlib
    px = np.asarray(pixels, dtype=np.float32).reshape(-1)
    if width <= 0 or height <= 0 or px.size != 3 * width * height:
        raise ValueError("buffer size matches dimensions")
    rows = tonemap_u8(px).reshape(height, 3 * width)
    raw = b"".join(b"\x00" + r.tobytes() for r in rows)   # filter type 0 per row

    def chunk(tag, data):
        return (struct.pack(">I", len(data)) + tag + data +
                struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF))

    png = (b"\x89PNG\r\n\x1a\n" +
           chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 2, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)
