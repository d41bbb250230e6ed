"""geon_render_dev_noalloc: the FP64 render with no allocation on any call.

Through the C ABI, against geon_render_dev on the same inputs, bit for bit
(pixels, per-pixel classes and directions, every statistic but kernel_ms):
two calls in a row, the second smaller than the first, on a caller's stream;
a call at exactly the ray limit (the per-ray planes of the fixed device
array must not overlap); the limit and precision errors; and no change in
the device's free memory across calls once the array is in place.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAT_KEYS = ("captured", "escaped", "diverged", "max_null_residual", "avg_steps",
             "max_steps_hit", "ray_steps", "fp64_rays", "gamma_steps")


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def _outputs(n):
    import torch
    return (torch.full((3 * n,), math.nan, dtype=torch.float32, device="cuda"),
            torch.full((n,), -99, dtype=torch.int32, device="cuda"),
            torch.full((3 * n,), math.nan, dtype=torch.float64, device="cuda"))


def _call(name, ctx, metric, cam, cfg, outs, stream=None, want_stats=True):
    import geon_amd as geon
    fn = getattr(geon.lib(), name)
    st = geon._abi.RenderStats()
    m, c, r = geon._cmetric(metric), cam.to_c(), cfg.to_c()
    return fn(ctx.handle, C.byref(m), C.byref(c), C.byref(r), C.c_void_p(outs[0].data_ptr()),
              C.c_void_p(outs[1].data_ptr()), C.c_void_p(outs[2].data_ptr()),
              C.byref(st) if want_stats else None, C.c_void_p(stream)), st


def _host(outs):
    import torch
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


def _same(name, metric, cam, cfg, ctx, stream=None):
    """One call of each entry point on the same inputs, compared bit for bit."""
    import geon_amd as geon
    n = cam.width * cam.height
    print(f"case: {name} {type(metric).__name__} {cam.width}x{cam.height} "
          f"spa={cfg.samples_per_axis}", flush=True)
    a, b = _outputs(n), _outputs(n)
    rc, sa = _call("geon_render_dev", ctx, metric, cam, cfg, a)
    geon._abi.check(rc, "geon_render_dev")
    rc, sb = _call("geon_render_dev_noalloc", ctx, metric, cam, cfg, b, stream)
    geon._abi.check(rc, "geon_render_dev_noalloc")
    ha, hb = _host(a), _host(b)
    for x, y, what in zip(ha, hb, ("rgb", "cls", "dir")):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), what
    for k in STAT_KEYS:
        assert getattr(sa, k) == getattr(sb, k), k
    assert sb.kernel_ms == 0.0
    return hb, sb


def test_noalloc_matches_render_dev_twice(ctx):
    import torch

    import geon_amd as geon
    s = torch.cuda.Stream()
    big = (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                           escape_radius=60.0),
           geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 160, 120),
           geon.RenderConfig(step=0.05, max_steps=3000, samples_per_axis=2))
    small = (geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
             geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 64, 48),
             geon.RenderConfig(step=0.05, max_steps=4000))
    h1, s1 = _same("first", *big, ctx, s.cuda_stream)
    h2, s2 = _same("second, smaller", *small, ctx, s.cuda_stream)
    # the smaller call's counts are its own, not the first call's left over
    assert s2.fp64_rays == 64 * 48 and s2.captured + s2.escaped + s2.diverged == 64 * 48
    assert s1.fp64_rays == 4 * 160 * 120
    assert (h2[1] >= 0).all() and np.isfinite(h2[0]).all()


def test_noalloc_at_the_ray_limit(ctx):
    import geon_amd as geon
    limit = int(geon.lib().geon_render_dev_noalloc_max_rays())
    assert limit == 1 << 22
    # 1024 × 1024 pixels at 2 × 2 samples = the limit: every plane is used
    # to its end, so overlapping planes would change classes or directions
    _same("limit", geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
          geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 1024, 1024),
          geon.RenderConfig(step=0.05, max_steps=400, samples_per_axis=2), ctx)


def test_noalloc_rejects_what_it_cannot_hold(ctx):
    import geon_amd as geon
    schw = geon.Schwarzschild(mass=1.0)
    outs = _outputs(16)
    # one pixel row more than the limit; 2049 samples per axis; FP32-mixed
    for cam, cfg, code in [
        (geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 2048, 2049), geon.RenderConfig(),
         geon._abi.E_INVALID),
        (geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 1, 1),
         geon.RenderConfig(samples_per_axis=2049), geon._abi.E_INVALID),
        (geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 4, 4),
         geon.RenderConfig(precision=geon.PREC_MIXED), geon._abi.E_UNSUPPORTED),
    ]:
        rc, _ = _call("geon_render_dev_noalloc", ctx, schw, cam, cfg, outs)
        assert rc == code, (cam, cfg, rc)
        assert geon.lib().geon_last_error()
    # the outputs were not touched
    h = _host(outs)
    assert (h[1] == -99).all() and np.isnan(h[0]).all()


def test_noalloc_leaves_free_memory_unchanged(ctx):
    import torch

    import geon_amd as geon
    m = geon.MorrisThorne(throat_radius=1.0, escape_radius=60.0)
    cfg = geon.RenderConfig(step=0.05, max_steps=2000)
    cams = [geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, w, h)
            for w, h in ((96, 64), (200, 150), (32, 32))]
    outs = [_outputs(c.width * c.height) for c in cams]
    rc, _ = _call("geon_render_dev_noalloc", ctx, m, cams[0], cfg, outs[0])
    geon._abi.check(rc, "geon_render_dev_noalloc")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for cam, o in zip(cams[1:], outs[1:]):
        rc, st = _call("geon_render_dev_noalloc", ctx, m, cam, cfg, o)
        geon._abi.check(rc, "geon_render_dev_noalloc")
        assert st.fp64_rays == cam.width * cam.height
        rc, _ = _call("geon_render_dev_noalloc", ctx, m, cam, cfg, o, want_stats=False)
        geon._abi.check(rc, "geon_render_dev_noalloc")
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
