"""Persistent scheduler and FP32-mixed precision on the GPU.

- FP64 frames run by the scheduler are bit-identical to the plain one-lane-
  per-ray kernel (same device arithmetic, different launch structure).
- Results do not depend on what else is in the batch, on the rays-per-lane
  setting or the refill interval: per-ray arithmetic never depends on the
  slot or lane a ray lands in.
- FP32-mixed: pixel-class mismatch vs the FP64 oracle ≤ 1e-3 and the escape
  direction error is reported and bounded.
"""
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


def _dev(n, dtype):
    import torch
    return torch.zeros(n, dtype=dtype, device="cuda")


def _frame_outputs(cam):
    import torch
    n = cam.width * cam.height
    return (_dev(3 * n, torch.float32), _dev(n, torch.int32), _dev(3 * n, torch.float64))


def _run_frames(ctx, frames, tune=None):
    import geon_amd as geon
    b = geon.Batch(ctx)
    outs, ids = [], []
    for metric, cam, cfg in frames:
        rgb, cls, d = _frame_outputs(cam)
        ids.append(b.add_frame(metric, cam, cfg, rgb.data_ptr(), cls.data_ptr(), d.data_ptr()))
        outs.append((rgb, cls, d))
    b.run(tune)
    stats = [b.frame_stats(i) for i in ids]
    res = [(rgb.cpu().numpy().reshape(-1, 3), cls.cpu().numpy(),
            d.cpu().numpy().reshape(3, -1).T.copy()) for rgb, cls, d in outs]
    tot = b.totals()
    b.close()
    return res, stats, tot


def _frames():
    import geon_amd as geon
    return [
        (geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
         geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 96, 80),
         geon.RenderConfig(step=0.05, max_steps=4000)),
        (geon.MorrisThorne(throat_radius=1.0, escape_radius=60.0),
         geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 64, 64),
         geon.RenderConfig(step=0.05, max_steps=3000)),
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                         escape_radius=60.0),
         geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 64, 48),
         geon.RenderConfig(step=0.05, max_steps=3000)),
        (geon.Schwarzschild(mass=1.7, horizon_pad=1.005, escape_radius=80.0),
         geon.Camera((0.0, 25.0, 1.2, 0.3), 50.0, 40, 56),
         geon.RenderConfig(step=0.05, max_steps=2500, samples_per_axis=2)),
    ]


def test_fp64_scheduler_matches_plain_kernel(ctx):
    import geon_amd as geon
    frames = _frames()
    res, stats, tot = _run_frames(ctx, frames)
    for (metric, cam, cfg), (rgb, cls, d), st in zip(frames, res, stats):
        rgb0, st0, cls0, d0 = geon.render(metric, cam, cfg, ctx=ctx)
        assert np.array_equal(cls, cls0)
        assert np.array_equal(d, d0)
        assert np.array_equal(rgb, rgb0)
        for k in ("captured", "escaped", "diverged", "ray_steps", "max_steps_hit"):
            assert st[k] == st0[k], k
        assert st["max_null_residual"] == st0["max_null_residual"]
    assert tot["executed_ray_steps"] == sum(s["ray_steps"] for s in stats)


@pytest.mark.parametrize("tune", [dict(rays_per_lane=2), dict(refill=7, block=64),
                                  dict(waves_per_simd=1, block=128, refill=200)])
def test_results_independent_of_launch_shape_and_mix(ctx, tune):
    frames = _frames()
    alone = [_run_frames(ctx, [f])[0][0] for f in frames[:2]]
    mixed, _, _ = _run_frames(ctx, frames, tune)
    for a, m in zip(alone, mixed):
        for x, y in zip(a, m):
            assert np.array_equal(x, y)


def test_mixed_precision_independent_of_slot_and_shape(ctx):
    """FP32-mixed rays give the same bits whichever kernel instantiation (one
    or two rays per lane) and slot they run in: the FP32 step writes every
    fused multiply-add out, so its rounding does not depend on how the
    compiler treats each copy. A mix with a frame in front shifts every
    ray's queue position and so its slot."""
    import geon_amd as geon
    frames = [
        (geon.MorrisThorne(throat_radius=1.9, escape_radius=100.0),
         geon.Camera((0.0, 35.0, math.pi / 2, 0.0), 60.0, 160, 120),
         geon.RenderConfig(step=0.05, max_steps=5000, precision=geon.PREC_MIXED)),
        (geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
         geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 128, 96),
         geon.RenderConfig(step=0.05, max_steps=4000, precision=geon.PREC_MIXED)),
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.8,
                         escape_radius=80.0),
         geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 96, 96),
         geon.RenderConfig(step=0.05, max_steps=5000, precision=geon.PREC_MIXED)),
    ]
    lead = (geon.Schwarzschild(mass=1.3, horizon_pad=1.005, escape_radius=90.0),
            geon.Camera((0.0, 25.0, 1.2, 0.0), 50.0, 37, 29),
            geon.RenderConfig(step=0.05, max_steps=3000, precision=geon.PREC_MIXED))
    base, _, _ = _run_frames(ctx, frames, dict(rays_per_lane=2))
    for tune, mix in ((dict(rays_per_lane=1), False), (dict(rays_per_lane=2), True),
                      (dict(rays_per_lane=1, refill=5, block=128), True)):
        res, _, _ = _run_frames(ctx, ([lead] if mix else []) + frames, tune)
        res = res[1:] if mix else res
        for a, b in zip(base, res):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (tune, mix)


def _angle(a, b):
    cross = np.linalg.norm(np.cross(a, b), axis=-1)
    return np.arctan2(cross, (a * b).sum(axis=-1))


def test_mixed_precision_schwarzschild_vs_oracle(ctx, O):
    """Schwarzschild r=30M 60° 256², capture r<2.01M, escape 100M: FP32 RK4 with
    null re-projection, FP64 for |b − 3√3M| < 1e-3 M and suspicious endings."""
    import geon_amd as geon
    metric = geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    cam = geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 256, 256)
    cfg = geon.RenderConfig(step=0.05, max_steps=4000, precision=geon.PREC_MIXED)
    res, stats, tot = _run_frames(ctx, [(metric, cam, cfg)])
    rgb, cls, d = res[0]
    om = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    y0 = O.camera_rays(om, O.camera(cam.observer, 60.0, 256, 256))
    ref = O.trace_rays(om, y0, 0.05, 4000)
    mismatch = float((cls != ref["cls"]).mean())
    real = (cls == ref["cls"]) & (ref["cls"] == O.ESCAPED) & (ref["y"][:, 1] < 1000.0)
    ang = _angle(d[real], ref["dir"][real])
    print(f"mixed C1 256²: class mismatch {mismatch:.3e}, escape angle max {ang.max():.3e} "
          f"mean {ang.mean():.3e} rad, fp64 rays {stats[0]['fp64_rays']}")
    assert mismatch <= 1e-3
    assert ang.max() < 1e-3 and ang.mean() < 1e-5
    assert stats[0]["fp64_rays"] > 0


def test_mixed_precision_generic_metrics_vs_oracle(ctx, O):
    import geon_amd as geon
    cases = [
        (geon.MorrisThorne(throat_radius=1.0, escape_radius=60.0),
         O.metric(O.MORRIS_THORNE, throat_radius=1.0, escape_radius=60.0),
         (0.0, 15.0, math.pi / 2, 0.0)),
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                         escape_radius=60.0),
         O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                  escape_radius=60.0), (0.0, 10.0, 0.0, 0.0)),
        # narrow camera far out on the axis: long FP32 flat runs before the wall
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                         escape_radius=100.0),
         O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                  escape_radius=100.0), (0.0, 40.0, 0.0, 0.0), 8.0, 4000),
    ]
    for case in cases:
        gm, om, obs = case[:3]
        fov, max_steps = case[3:] if len(case) > 3 else (60.0, 3000)
        cam = geon.Camera(obs, fov, 96, 96)
        cfg = geon.RenderConfig(step=0.05, max_steps=max_steps, precision=geon.PREC_MIXED)
        (rgb, cls, d), = _run_frames(ctx, [(gm, cam, cfg)])[0]
        y0 = O.camera_rays(om, O.camera(obs, fov, 96, 96))
        ref = O.trace_rays(om, y0, 0.05, max_steps)
        mismatch = float((cls != ref["cls"]).mean())
        real = (cls == ref["cls"]) & (ref["cls"] == O.ESCAPED)
        ang = _angle(d[real], ref["dir"][real])
        print(f"mixed kind={om.kind}: mismatch {mismatch:.3e}, angle max {ang.max():.3e} "
              f"mean {ang.mean():.3e}")
        assert mismatch <= 1e-3
        assert ang.mean() < 1e-4 and ang.max() < 1e-3


def test_mixed_precision_morris_thorne_headline_frame(ctx):
    """FP32-mixed against the FP64 scheduler on the Morris–Thorne 1024² frame
    (b0 = 1, l = 15, 20,000 steps): no class mismatch and every escaped
    direction within 1e-3 rad, leaving out the pole-flung rays whose
    direction is (0, 0, 0) in both precisions (their states overflow the
    direction formula)."""
    import geon_amd as geon
    m = geon.MorrisThorne(throat_radius=1.0)
    cam = geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 1024, 1024)
    out = {}
    for prec in (geon.PREC_FP64, geon.PREC_MIXED):
        (res,), _, tot = _run_frames(ctx, [(m, cam, geon.RenderConfig(step=0.05, max_steps=20000,
                                                                        precision=prec))])
        out[prec] = (res, tot)
    (_, c64, d64), _ = out[geon.PREC_FP64]
    (_, cmx, dmx), tot = out[geon.PREC_MIXED]
    assert np.array_equal(c64, cmx)
    n64 = np.linalg.norm(d64, axis=1)
    flung = (c64 == 2) & (n64 == 0.0)
    assert np.array_equal(dmx[flung], d64[flung])
    esc = (c64 == 2) & ~flung
    ang = _angle(dmx[esc], d64[esc])
    print(f"MT 1024² mixed: escaped {int(esc.sum())}, flung {int(flung.sum())}, re-traced in FP64 "
          f"{tot['fp64_rays']}, angle max {ang.max():.3e}", flush=True)
    assert esc.sum() > 3000 and ang.max() < 1e-3


def test_explicit_rays_early_exit_matches_full_trace(ctx, O):
    """Capture predicate with the photon-sphere early exit (bisection rays),
    Schwarzschild M=1, r_obs=20, step 0.01, max 200,000."""
    import torch

    import geon_amd as geon
    alphas = np.linspace(0.01, math.pi / 2 - 2e-3, 301)
    alphas = alphas[np.abs(alphas - 0.24904141754185194) > 1e-4]
    y0 = np.array([O.equatorial_ray(1.0, 20.0, a) for a in alphas])
    ref = O.trace_rays(O.metric(O.SCHWARZSCHILD), y0, 0.01, 200_000)
    n = len(alphas)
    ys = torch.tensor(np.ascontiguousarray(y0.T), dtype=torch.float64, device="cuda")
    cls = torch.zeros(n, dtype=torch.int32, device="cuda")
    steps = torch.zeros(n, dtype=torch.int32, device="cuda")
    b = geon.Batch(ctx)
    b.add_rays(geon.Schwarzschild(), ys.data_ptr(), n, 0.01, 200_000, early_exit=True,
               cls=cls.data_ptr(), steps=steps.data_ptr())
    b.run()
    tot = b.totals()
    b.close()
    got = cls.cpu().numpy()
    # the fixed-step full trace flings some captured rays out at the horizon
    # (a stage lands on r ≈ 2M); the early exit decides them before that
    artefact = (ref["cls"] == O.ESCAPED) & (ref["y"][:, 1] > 2000.0)
    assert np.array_equal((got == O.CAPTURED)[~artefact], (ref["cls"] == O.CAPTURED)[~artefact])
    assert np.all(got[artefact] == O.CAPTURED)
    assert tot["executed_ray_steps"] < 0.2 * int(ref["steps"].sum())


def test_autotune(ctx):
    import geon_amd as geon
    cfg = geon.autotune(geon.SCHWARZSCHILD, geon.PREC_MIXED, ctx=ctx)
    print("autotuned schwarzschild mixed:", cfg)
    assert cfg["block"] in (64, 128, 256) and cfg["rays_per_lane"] in (1, 2)
    assert 1 <= cfg["waves_per_simd"] <= 8 and cfg["refill"] >= 1
