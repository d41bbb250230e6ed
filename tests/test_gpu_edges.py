"""Edge cases on the GPU: empty inputs, odd frame shapes with supersampling
against the oracle's render (render.rs:55-165), and non-finite initial
states through the persistent scheduler in both precisions."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def test_empty_inputs(ctx):
    import geon_amd as geon
    m = geon.Schwarzschild()
    got = geon.trace_rays(m, np.zeros((0, 8)), 0.05, 100, ctx=ctx)
    assert got["cls"].shape == (0,) and got["y"].shape == (0, 8)
    assert geon.curvature(m, np.zeros((0, 4)), geon.CURV_STRESS, ctx=ctx).shape == (0, 16)
    a, b = geon.shadow_bisect(np.zeros(0), np.zeros(0), ctx=ctx)
    assert a.size == 0 and b.size == 0
    bt = geon.Batch(ctx)
    bt.add_rays(m, 0, 0, 0.05, 100)
    bt.run()
    assert bt.totals()["executed_ray_steps"] == 0
    bt.close()


@pytest.mark.parametrize("w,h,spa", [(1, 1, 1), (1, 7, 2), (33, 17, 3)])
def test_odd_frames_match_oracle_render(ctx, O, w, h, spa):
    """Pixel classes, RGB and statistics of small odd frames with spa² sub-rays
    per pixel, Schwarzschild r=12M looking at the hole (captures, escapes and
    the majority rule all occur)."""
    import geon_amd as geon
    metric = geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=60.0)
    cam = geon.Camera((0.0, 12.0, 1.4, 0.2), 70.0, w, h)
    cfg = geon.RenderConfig(step=0.05, max_steps=3000, samples_per_axis=spa)
    rgb, st, cls, d = geon.render(metric, cam, cfg, ctx=ctx)
    om = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=60.0)
    ref = O.render(om, O.camera(cam.observer, 70.0, w, h),
                   O.render_cfg(step=0.05, max_steps=3000, samples_per_axis=spa))
    assert np.array_equal(cls, ref["cls"])
    # sky colours agree except where a direction within ulps of a sky cell
    # edge lands in the neighbouring cell
    assert (np.abs(rgb - ref["rgb"]).max(axis=1) == 0).mean() >= 0.99
    assert st["captured"] == ref["stats"].captured
    assert st["escaped"] == ref["stats"].escaped
    assert st["diverged"] == ref["stats"].diverged
    # per-pixel integer step averages; a ray's step count can move by one
    # where transcendental ulps shift a threshold crossing
    assert abs(st["avg_steps"] - ref["stats"].avg_steps) <= 1.0


@pytest.mark.parametrize("precision", [0, 1])
def test_nonfinite_states_through_scheduler(ctx, O, precision):
    """NaN and infinite initial states end diverged with their steps at 0 in
    the scheduler, whatever their neighbours do."""
    import torch

    import geon_amd as geon
    f = 1 - 2 / 20.0
    good = np.array([0.0, 20.0, math.pi / 2, 0.0, 1 / math.sqrt(f), -0.9 * math.sqrt(f), 0.0,
                     0.4359 * math.sqrt(f) / 20.0])
    y0 = np.tile(good, (256, 1))
    y0[::7, 1] = math.inf
    y0[3::7, 5] = math.nan
    ref = O.trace_rays(O.metric(O.SCHWARZSCHILD), y0, 0.05, 3000)
    ys = torch.tensor(np.ascontiguousarray(y0.T), dtype=torch.float64, device="cuda")
    cls = torch.zeros(256, dtype=torch.int32, device="cuda")
    steps = torch.zeros(256, dtype=torch.int32, device="cuda")
    bt = geon.Batch(ctx)
    bt.add_rays(geon.Schwarzschild(), ys.data_ptr(), 256, 0.05, 3000, precision=precision,
                cls=cls.data_ptr(), steps=steps.data_ptr())
    bt.run()
    bt.close()
    got_cls, got_steps = cls.cpu().numpy(), steps.cpu().numpy()
    bad = ~np.isfinite(y0).all(axis=1)
    assert np.all(got_cls[bad] == O.DIVERGED) and np.all(got_steps[bad] == 0)
    assert np.array_equal(got_cls[~bad], ref["cls"][~bad])
