"""GPU parity: the HIP path through the C ABI vs the CPU oracle on the same inputs.

The FP64 bar is parity_util's (stated there): classes identical on every ray
except rays shown to lie within 1e-9 of a capture boundary in impact
parameter, step counts identical, escape directions within 1e-9 rad, null
residuals equal to the oracle's (and ≤ 1e-12 wherever the oracle's is). The
device code follows the reference's operation order without FMA
contraction; its sin/cos/tanh may differ from glibc's in the last bit.
"""
import math

import numpy as np
import pytest

from parity_util import ANGLE_TOL, angle as _angle, compare_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def test_render_schwarzschild_c1(ctx, O):
    """Plain FP64 Schwarzschild render, 256², camera r=30M, 60°, capture r<2.01M,
    escape r>100M, step 0.05, max 4000, stars on: every one of the 65,536 rays."""
    import geon_amd as geon
    metric = geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    cam = geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 256, 256)
    cfg = geon.RenderConfig(step=0.05, max_steps=4000)
    rgb, stats, cls, d = geon.render(metric, cam, cfg, ctx=ctx)
    om = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    ocam = O.camera(cam.observer, 60.0, 256, 256)
    y0 = O.camera_rays(om, ocam)
    ref = O.trace_rays(om, y0, 0.05, 4000)
    # per-ray steps, states and residuals from trace_rays on the same initial states
    tr = geon.trace_rays(metric, y0, 0.05, 4000, ctx=ctx)
    assert np.array_equal(tr["cls"], cls)
    assert np.array_equal(tr["dir"], d)
    compare_rays(O, om, y0, tr, ref, 0.05, 4000, "C1 render", esc_r=100.0)
    # statistics follow render.rs:146-163
    assert stats["captured"] + stats["escaped"] + stats["diverged"] == 65536
    assert stats["ray_steps"] == int(tr["steps"].astype(np.int64).sum())
    assert stats["gamma_steps"] == stats["ray_steps"]       # Γ at every Schwarzschild step
    assert stats["avg_steps"] == stats["ray_steps"] / 65536
    fin = np.isfinite(tr["resid"])
    assert stats["max_null_residual"] == tr["resid"][fin].max()
    # colour of escaped pixels equals the sky at the reference direction, up to
    # ulp-level differences in acos/atan2 at cell edges
    esc = (cls == O.ESCAPED) & (ref["cls"] == O.ESCAPED)
    ref_rgb = np.array([O.sky_sample(v, True) for v in ref["dir"][esc][:4000]])
    assert (np.abs(rgb[esc][:4000] - ref_rgb).max(axis=1) == 0).mean() > 0.999
    assert np.all(rgb[cls == O.CAPTURED] == 0.0)


@pytest.mark.parametrize("kind,kw,obs", [
    (0, {}, (0.0, 20.0, 0.0, 0.0)),
    (1, dict(mass=1.0), (0.0, 30.0, math.pi / 2, 0.0)),
    (1, dict(mass=1.0), (0.0, 12.0, 1.3, 0.4)),
    (2, dict(bubble_thickness=0.125, ship_velocity=0.5), (0.0, -10.0, 0.0, 0.0)),
    (2, {}, (0.0, 8.0, 0.0, 0.0)),
    (3, dict(throat_radius=1.0), (0.0, 15.0, math.pi / 2, 0.0)),
])
def test_camera_rays_are_null(ctx, O, kind, kw, obs):
    """Camera rays leave the observer on the null cone: |g_μν p^μ p^ν| ≤ 1e-12 at
    the initial state (trace with max_steps = 0 reports the residual of the
    state it was given), equal to the oracle's residual of the same states."""
    import geon_amd as geon
    om = O.metric(kind, **kw)
    y0 = O.camera_rays(om, O.camera(obs, 60.0, 48, 40))
    got = geon.trace_rays(om, y0, 0.05, 0, ctx=ctx)
    ref = O.trace_rays(om, y0, 0.05, 0)
    assert np.array_equal(got["y"], y0)
    assert got["resid"].max() <= 1e-12 and ref["resid"].max() <= 1e-12
    assert np.abs(got["resid"] - ref["resid"]).max() <= 1e-15


@pytest.mark.parametrize("kind,kw,obs,esc_r,max_steps", [
    (3, dict(throat_radius=1.0), (0.0, 15.0, math.pi / 2, 0.0), 60.0, 3000),
    (2, dict(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5), (0.0, 10.0, 0.0, 0.0),
     60.0, 3000),
    (0, {}, (0.0, 20.0, 0.0, 0.0), 60.0, 3000),
    (1, dict(mass=1.0), (0.0, 20.0, 1.3, 0.4), 80.0, 6000),
    (1, dict(mass=1.0), (0.0, 12.0, 1.3, 0.4), 80.0, 8000),
])
def test_trace_rays_generic_metrics(ctx, O, kind, kw, obs, esc_r, max_steps):
    import geon_amd as geon
    om = O.metric(kind, escape_radius=esc_r, **kw)
    y0 = O.camera_rays(om, O.camera(obs, 60.0, 64, 48))
    ref = O.trace_rays(om, y0, 0.05, max_steps)
    got = geon.trace_rays(om, y0, 0.05, max_steps, ctx=ctx)
    compare_rays(O, om, y0, got, ref, 0.05, max_steps, f"trace kind={kind} {kw} obs={obs}",
                 esc_r=esc_r)
    # final states of escaped rays (rays still orbiting at max_steps, or
    # captured next to the horizon, amplify last-bit differences without bound)
    far = np.abs(ref["y"][:, 1]) if kind in (1, 3) else np.linalg.norm(ref["y"][:, 1:4], axis=1)
    fin = (got["cls"] == ref["cls"]) & (ref["cls"] == O.ESCAPED) & (far < 10 * esc_r)
    rel = np.abs(got["y"][fin] - ref["y"][fin]) / np.maximum(1.0, np.abs(ref["y"][fin]))
    assert rel.max() < 1e-8


@pytest.mark.parametrize("e", [-300, -110, 110, 300])
def test_schwarzschild_mass_scale_takes_tested_steps(ctx, O, e):
    """M = 2^e outside [2^-100, 2^100]: every stage's quotient range test fails,
    so every step is taken again with the tested stages (IEEE divisions).
    Scaling M, t, r and the step by 2^e maps the reference's discrete map onto
    itself exactly — every operation of a Schwarzschild step either scales by
    a power of two or is scale-free, and nothing leaves the normal range for
    |e| ≤ 300 — and the oracle shows it on this camera (the same classes, step
    counts, directions and scaled final states at every e). So the GPU must
    give its own M = 1 results bit for bit, through the other code path, and
    meet the oracle's bar."""
    import geon_amd as geon
    s = 2.0 ** e
    obs1 = (0.0, 30.0, math.pi / 2, 0.0)
    om1 = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    y01 = O.camera_rays(om1, O.camera(obs1, 60.0, 64, 48))
    g1 = geon.trace_rays(om1, y01, 0.05, 4000, ctx=ctx)
    om = O.metric(O.SCHWARZSCHILD, mass=s, horizon_pad=1.005, escape_radius=100.0)
    y0 = O.camera_rays(om, O.camera((0.0, 30.0 * s, math.pi / 2, 0.0), 60.0, 64, 48))
    scale = np.array([s, s, 1.0, 1.0, 1.0, 1.0, 1.0 / s, 1.0 / s])
    assert np.array_equal(y0, y01 * scale)
    got = geon.trace_rays(om, y0, 0.05 * s, 4000, ctx=ctx)
    assert np.array_equal(got["cls"], g1["cls"])
    assert np.array_equal(got["steps"], g1["steps"])
    assert np.array_equal(got["dir"], g1["dir"], equal_nan=True)
    assert np.array_equal(got["y"], g1["y"] * scale, equal_nan=True)
    assert np.array_equal(got["resid"], g1["resid"], equal_nan=True)
    ref = O.trace_rays(om, y0, 0.05 * s, 4000)
    compare_rays(O, om, y0, got, ref, 0.05 * s, 4000, f"Schwarzschild M=2^{e}", esc_r=100.0 * s)


@pytest.mark.parametrize("e", [-130, 130])
def test_morris_thorne_tested_inverse(ctx, O, e):
    """b0 = 2^±130 puts |det g| = (b0² + l²)² sin²θ outside [2^-500, 2^500]
    (about 2^-504 and 2^535 on this camera), so every Morris–Thorne step is
    taken again with the tested inverse (IEEE divisions). The finite
    differences keep their absolute h = 1e-5, so this is not a rescaled
    frame; it is held to the oracle's bar on the same discrete map."""
    import geon_amd as geon
    s = 2.0 ** e
    om = O.metric(O.MORRIS_THORNE, throat_radius=s, escape_radius=60.0 * s)
    y0 = O.camera_rays(om, O.camera((0.0, 15.0 * s, math.pi / 2, 0.0), 60.0, 64, 48))
    det = np.prod(np.diag(O.g(om, y0[0][:4])))
    assert not (2.0 ** -500 < abs(det) < 2.0 ** 500)
    ref = O.trace_rays(om, y0, 0.05 * s, 3000)
    got = geon.trace_rays(om, y0, 0.05 * s, 3000, ctx=ctx)
    compare_rays(O, om, y0, got, ref, 0.05 * s, 3000, f"Morris-Thorne b0=2^{e}",
                 esc_r=60.0 * s)


def test_superluminal_alcubierre_conditioning(ctx, O):
    """The superluminal bubble (v = 2, s = 0.5) has horizons in its walls: there
    the reference's own result is ill-conditioned in the last bit of tanh.
    Classes and step counts must agree on every ray. Escape directions agree
    within 1e-9 rad except on rays that are ill-conditioned in that last bit,
    shown per ray with perturbations planted in the oracle: every tanh moved
    by one ulp up, down, or up or down at random per evaluation (the device
    tanh differs from glibc's by about that). Such a ray's own one-ulp
    spread is above 1e-10 rad and the GPU lies within 30× of it."""
    import geon_amd as geon
    om = O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0,
                  escape_radius=60.0)
    y0 = O.camera_rays(om, O.camera((0.0, 8.0, 0.0, 0.0), 60.0, 64, 48))
    ref = O.trace_rays(om, y0, 0.05, 3000)
    got = geon.trace_rays(om, y0, 0.05, 3000, ctx=ctx)
    assert np.array_equal(got["cls"], ref["cls"])
    assert np.array_equal(got["steps"].astype(np.int64), ref["steps"].astype(np.int64))
    esc = (ref["cls"] == O.ESCAPED) & (np.linalg.norm(ref["y"][:, 1:4], axis=1) < 600.0)
    a = _angle(got["dir"][esc], ref["dir"][esc])
    try:
        sens = np.zeros(esc.sum())
        for k in (-1, 1, 2, 3, 4, 5):
            O.set_tanh_ulp_shift(k)
            pert = O.trace_rays(om, y0[esc], 0.05, 3000)
            sens = np.maximum(sens, _angle(pert["dir"], ref["dir"][esc]))
    finally:
        O.set_tanh_ulp_shift(0)
    off = a > ANGLE_TOL
    ratio = a[off] / np.maximum(sens[off], 1e-300)
    print(f"superluminal: escaped {esc.sum()}, above 1e-9 rad {off.sum()} (max {a.max():.3e}); "
          f"their one-ulp spread min {sens[off].min() if off.any() else 0:.3e}, GPU/spread max "
          f"{ratio.max() if off.any() else 0:.2f}; rays whose spread exceeds 1e-9 rad "
          f"{(sens > ANGLE_TOL).sum()}", flush=True)
    assert (sens[off] > 1e-10).all() and (ratio <= 30.0).all()


def _c2_sample(O, om, cam, n, seed):
    st, tet = O.tetrad(om, cam.observer)
    assert st == 0
    idx = np.random.default_rng(seed).choice(cam.width * cam.height, size=n, replace=False)
    y0 = np.array([O.ray_state_sub(cam, tet, int(i) % cam.width, int(i) // cam.width) for i in idx])
    return idx, y0


@pytest.mark.parametrize("label,kind,kw,obs,max_steps", [
    ("alcubierre x=-10 (as given)", 2, dict(bubble_radius=2.0, bubble_thickness=0.125,
                                             ship_velocity=0.5), (0.0, -10.0, 0.0, 0.0), 20_000),
    ("alcubierre x=+10 (bubble in view)", 2, dict(bubble_radius=2.0, bubble_thickness=0.125,
                                                   ship_velocity=0.5), (0.0, 10.0, 0.0, 0.0), 40_000),
    ("morris-thorne l=15", 3, dict(throat_radius=1.0), (0.0, 15.0, math.pi / 2, 0.0), 20_000),
])
def test_c2_render_1024_sampled(ctx, O, label, kind, kw, obs, max_steps):
    """The generic-metric FP64 workload at its configured size: 1024² at 60°,
    step 0.05, reference escape radius (1000), fd_step 1e-5. The whole frame is
    rendered on the GPU; 2048 seeded pixels are traced by the oracle and held
    to the FP64 bar (class, steps, direction, residual), and the frame's
    statistics are checked against the per-ray outputs."""
    import geon_amd as geon
    om = O.metric(kind, **kw)
    cam = O.camera(obs, 60.0, 1024, 1024)
    gm = geon.Alcubierre(**kw) if kind == 2 else geon.MorrisThorne(**kw)
    gcam = geon.Camera(obs, 60.0, 1024, 1024)
    rgb, stats, cls, d = geon.render(gm, gcam, geon.RenderConfig(step=0.05, max_steps=max_steps),
                                     ctx=ctx)
    idx, y0 = _c2_sample(O, om, cam, 2048, seed=kind * 10 + int(obs[1] > 0))
    print(f"{label}: frame classes {np.bincount(cls, minlength=4).tolist()}, ray-steps "
          f"{stats['ray_steps']} (Γ-evaluating {stats['gamma_steps']}), kernel "
          f"{stats['kernel_ms']:.1f} ms", flush=True)
    ref = O.trace_rays(om, y0, 0.05, max_steps)
    tr = geon.trace_rays(gm, y0, 0.05, max_steps, ctx=ctx)
    # the render's pixels are the traced rays (one sample per pixel)
    assert np.array_equal(cls[idx], tr["cls"])
    assert np.array_equal(d[idx], tr["dir"])
    compare_rays(O, om, y0, tr, ref, 0.05, max_steps, label, esc_r=1000.0)
    assert stats["captured"] + stats["escaped"] + stats["diverged"] == 1024 * 1024
    assert 0 < stats["gamma_steps"] <= stats["ray_steps"] or kind == 2
    if kind == 3:
        assert stats["gamma_steps"] == stats["ray_steps"]


def test_c2_curvature_full_launches(ctx, O):
    """C2 curvature as configured: Alcubierre (R=2, s=1/8, v=0.5) T_μν on the
    whole 128³ grid over [−4, 4]³ (x_i = −4 + 8i/127, t = 0) in one launch,
    4096 seeded grid points against the oracle; Morris–Thorne (b0 = 1) on the
    4096-point line l ∈ [−4, 4], θ = π/2, every point against the oracle.
    Tolerance 1e-5 of the largest |T| (the FD of an FD amplifies rounding by
    ~1/h²), and the Morris–Thorne closed form T_tt = −b0²/(8π(b0² + l²)²)."""
    import geon_amd as geon
    alc = O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5)
    g = -4.0 + 8.0 * np.arange(128) / 127.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([np.zeros(X.size), X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    T = geon.stress_energy(alc, pts, ctx=ctx)
    assert T.shape == (128 ** 3, 4, 4) and np.isfinite(T).all()
    idx = np.random.default_rng(5).choice(len(pts), 4096, replace=False)
    ref, _ = O.curvature(alc, pts[idx], O.CURV_STRESS)
    scale = max(1.0, np.abs(ref).max())
    err = np.abs(T[idx].reshape(-1, 16) - ref).max()
    print(f"alcubierre 128^3: max |T| {np.abs(T).max():.4e}, sampled max |Δ| {err:.3e} "
          f"(scale {scale:.3e})", flush=True)
    assert err <= 1e-5 * scale
    # T is symmetric; far outside the walls the metric is flat and T vanishes
    assert np.abs(T - T.transpose(0, 2, 1)).max() <= 1e-5 * scale
    far = np.linalg.norm(pts[:, 1:], axis=1) > 2.0 + 22 * 0.125
    assert np.abs(T[far]).max() == 0.0

    mt = O.metric(O.MORRIS_THORNE, throat_radius=1.0)
    l = -4.0 + 8.0 * np.arange(4096) / 4095.0
    line = np.stack([np.zeros(4096), l, np.full(4096, math.pi / 2), np.zeros(4096)], axis=1)
    Tm = geon.stress_energy(mt, line, ctx=ctx)
    refm, _ = O.curvature(mt, line, O.CURV_STRESS)
    errm = np.abs(Tm.reshape(-1, 16) - refm).max()
    print(f"morris-thorne 4096-point line: max |Δ| {errm:.3e}", flush=True)
    assert errm <= 1e-5 * max(1.0, np.abs(refm).max())
    closed = -1.0 / (8 * math.pi * (1.0 + l * l) ** 2)
    assert np.abs(Tm[:, 0, 0] - closed).max() <= 1e-6


def test_alcubierre_ricci_quads_match_riemann_team(ctx):
    """The Ricci family of Alcubierre runs on quads (4 lanes per point) and
    Riemann on 8-lane teams. R_σν = Σ_α R^α_σαν (curvature.rs:72-85) formed
    on the host from the team kernel's Riemann, summed from 0.0 over α in
    order, equals the quad kernel's Ricci bit for bit: both form every
    R^α_σαν by the same operations. Points: a seeded sample of the 128³ grid
    plus the wall, the centre and the flat region, for two bubbles."""
    import geon_amd as geon
    rng = np.random.default_rng(11)
    g = -4.0 + 8.0 * np.arange(128) / 127.0
    idx = rng.integers(0, 128, (3000, 3))
    pts = np.c_[np.zeros(len(idx)), g[idx[:, 0]], g[idx[:, 1]], g[idx[:, 2]]]
    pts = np.r_[pts, [[0, 2.0, 0, 0], [0, 0, 0, 0], [0, -1.9, 0.3, 0.1], [0, 6.0, 1.0, -2.0]]]
    for m in (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5),
              geon.Alcubierre()):
        riem = geon.riemann(m, pts, ctx=ctx).reshape(len(pts), 4, 4, 4, 4)
        ric = geon.ricci(m, pts, ctx=ctx).reshape(len(pts), 4, 4)
        want = np.zeros((len(pts), 4, 4))
        for a in range(4):
            want = want + riem[:, a, :, a, :]
        assert np.array_equal(ric.view(np.uint64), want.view(np.uint64))
        assert np.abs(ric).max() > 0.0


def test_alcubierre_long_flat_runs(ctx, O):
    """A narrow camera 40 units out on the bubble's axis: every ray first
    crosses ~35 units of exactly flat space (flatness tested once per run of
    up to ~700 steps), then most of them are deflected by the wall. Classes
    and escape directions against the oracle."""
    import geon_amd as geon
    om = O.metric(O.ALCUBIERRE, bubble_thickness=0.125, ship_velocity=0.5, escape_radius=100.0)
    y0 = O.camera_rays(om, O.camera((0.0, 40.0, 0.0, 0.0), 8.0, 48, 48))
    ref = O.trace_rays(om, y0, 0.05, 4000)
    got = geon.trace_rays(om, y0, 0.05, 4000, ctx=ctx)
    compare_rays(O, om, y0, got, ref, 0.05, 4000, "alcubierre far camera", esc_r=100.0)
    # the wall deflected most rays (their final direction is not the initial one)
    p0 = y0[:, 5:8] / np.linalg.norm(y0[:, 5:8], axis=1)[:, None]
    defl = np.arccos(np.clip((p0 * ref["dir"]).sum(axis=1), -1.0, 1.0))
    assert (ref["cls"] == O.ESCAPED).all() and (defl > 1e-6).mean() > 0.5


def test_trace_semantics_on_gpu(ctx, O):
    """terminate-first, NaN guard with the pre-step state, max_steps."""
    import geon_amd as geon
    om = O.metric(O.SCHWARZSCHILD)
    f = 1 - 2 / 10.0
    base = np.array([0.0, 10.0, math.pi / 2, 0.0, 1 / math.sqrt(f), -math.sqrt(f), 0.0, 0.0])
    y0 = np.tile(base, (4, 1))
    y0[0, 1] = 1500.0          # escaped before any step
    y0[1, 1] = math.inf        # non-finite r
    y0[2, 6] = math.nan        # NaN momentum → diverged at n = 0, state kept
    got = geon.trace_rays(om, y0, 0.05, 7, ctx=ctx)
    ref = O.trace_rays(om, y0, 0.05, 7)
    assert list(got["cls"]) == list(ref["cls"]) == [O.ESCAPED, O.DIVERGED, O.DIVERGED, O.MAXSTEPS]
    assert list(got["steps"]) == [0, 0, 0, 7]
    assert got["y"][2, 1] == 10.0 and math.isnan(got["y"][2, 6])
    assert np.array_equal(got["y"][3], ref["y"][3]) or np.abs(got["y"][3] - ref["y"][3]).max() < 1e-14


def test_curvature_points(ctx, O):
    import geon_amd as geon
    rng = np.random.default_rng(1)
    cases = [
        (O.metric(O.SCHWARZSCHILD), np.c_[np.zeros(40), rng.uniform(3, 30, 40),
                                          rng.uniform(0.3, 2.8, 40), rng.uniform(-3, 3, 40)]),
        (O.metric(O.MORRIS_THORNE), np.c_[np.zeros(40), rng.uniform(-4, 4, 40),
                                          rng.uniform(0.3, 2.8, 40), np.zeros(40)]),
        (O.metric(O.ALCUBIERRE, bubble_thickness=0.125, ship_velocity=0.5),
         np.c_[np.zeros(40), rng.uniform(-4, 4, (40, 3))]),
        (O.metric(O.ALCUBIERRE), np.c_[np.zeros(40), rng.uniform(-4, 4, (40, 3))]),
        (O.metric(O.MINKOWSKI), np.c_[np.zeros(8), rng.uniform(-4, 4, (8, 3))]),
    ]
    for om, x in cases:
        for what, width in [(O.CURV_STRESS, 16), (O.CURV_RIEMANN, 256), (O.CURV_SCALAR, 1),
                            (O.CURV_RICCI, 16), (O.CURV_EINSTEIN, 16)]:
            ref, _ = O.curvature(om, x, what)
            got = geon.curvature(om, x, what, ctx=ctx)
            scale = max(1.0, np.abs(ref).max())
            # FD-of-FD noise: transcendental ulps are amplified by ~1/h² ≈ 1e10
            err = np.abs(got - ref).max()
            print(f"curvature kind={om.kind} what={what}: max |Δ| {err:.3e} (scale {scale:.3e})")
            assert err <= 1e-5 * scale, (om.kind, what, err)


def test_energy_condition_anchors_on_gpu(ctx, anchors):
    """docs/energy_conditions.txt rows, recomputed by the GPU curvature kernel."""
    import geon_amd as geon
    rows = anchors["energy_conditions"]["rows"]
    k = np.array([1.0, 1.0, 0.0, 0.0])
    mt = geon.MorrisThorne(throat_radius=1.0)
    alc = geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0)
    for r in rows:
        if r["kind"] == "morris_thorne_nec":
            v = geon.nec_contraction(mt, [0.0, r["coord"], math.pi / 2, 0.0], k, ctx=ctx)[0]
        elif r["kind"] == "alcubierre_nec":
            v = geon.nec_contraction(alc, [0.0, r["coord"], 0.0, 0.0], k, ctx=ctx)[0]
        elif r["kind"].startswith("schwarzschild"):
            rr = r["coord"]
            x = [0.0, rr, math.pi / 2, 0.0]
            f = 1 - 2 / rr
            if r["kind"].endswith("nec"):
                v = geon.nec_contraction(geon.Schwarzschild(), x,
                                         [1 / math.sqrt(f), -math.sqrt(f), 0, 0], ctx=ctx)[0]
            else:
                v = geon.ricci_scalar(geon.Schwarzschild(), x, ctx=ctx)[0]
            assert abs(v) < 1e-10           # vacuum
            continue
        else:
            continue
        assert abs(v - r["value"]) <= 1e-4 * abs(r["value"]) + 1e-9, (r, v)
    T = geon.stress_energy(mt, [0.0, 0.0, math.pi / 2, 0.0], ctx=ctx)[0]
    assert float(f"{T[0, 0]:.3g}") == float(f"{-1 / (8 * math.pi):.3g}")


def test_energy_condition_scans_on_gpu(ctx, O, anchors):
    """The line scans of bin/energy_conditions.rs on the device (geon_nec_scan):
    the Alcubierre minimum over x = −4 + 0.2n and the ∫ of the negative part
    over x = −5, −5 + 0.05, … ≤ 5 (accumulated as the harness does), and the
    Morris–Thorne minimum at the throat, against docs/energy_conditions.txt
    and against the oracle point by point. The Alcubierre minimum has a mirror
    point at −2.40 whose value differs from +2.40's only by FD noise, so the
    scan may keep either."""
    import geon_amd as geon
    rows = {r["kind"]: r for r in anchors["energy_conditions"]["rows"]}
    k = [1.0, 1.0, 0.0, 0.0]
    alc = geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0)
    oalc = O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0,
                    ship_x=0.0)
    xs = np.array([[0.0, -4.0 + 0.2 * n, 0.0, 0.0] for n in range(41)])
    nec, mv, mi, _ = geon.nec_scan(alc, xs, k, ctx=ctx)
    ref = np.array([O.nec_contraction(oalc, x, k) for x in xs])
    print(f"alcubierre scan: min {mv:.6e} at x = {xs[mi, 1]:.2f}; max |Δ| vs oracle "
          f"{np.abs(nec - ref).max():.3e}", flush=True)
    assert np.abs(nec - ref).max() <= 1e-5 * np.abs(ref).max()
    want = rows["alcubierre_nec_min"]
    assert abs(mv - want["value"]) <= 1e-4 * abs(want["value"])
    assert f"{abs(xs[mi, 1]):.2f}" == "2.40" and nec[mi] == nec.min()
    pts, x_pos = [], -5.0
    while x_pos <= 5.0:
        pts.append([0.0, x_pos, 0.0, 0.0])
        x_pos += 0.05
    _, _, _, integral = geon.nec_scan(alc, pts, k, dx=0.05, ctx=ctx)
    want = rows["alcubierre_nec_integral"]["value"]
    print(f"alcubierre ∫ negative part: {integral:.6e} (oracle {O.alcubierre_nec_integral(oalc):.6e})")
    assert abs(integral - want) <= 1e-4 * abs(want)
    assert abs(integral - O.alcubierre_nec_integral(oalc)) <= 1e-6 * abs(want)
    mt = geon.MorrisThorne(throat_radius=1.0)
    xs = np.array([[0.0, -4.0 + 0.2 * n, math.pi / 2, 0.0] for n in range(41)])
    nec, mv, mi, _ = geon.nec_scan(mt, xs, k, ctx=ctx)
    want = rows["morris_thorne_nec_min"]
    assert abs(mv - want["value"]) <= 1e-4 * abs(want["value"]) and mi == 20


@pytest.mark.parametrize("kind,kw,obs", [
    (0, dict(escape_radius=100.0), (0.0, 20.0, 0.0, 0.0)),
    (1, dict(horizon_pad=1.005, escape_radius=100.0), (0.0, 30.0, math.pi / 2, 0.0)),
    (2, dict(bubble_thickness=0.125, ship_velocity=0.5, escape_radius=60.0), (0.0, 10.0, 0.0, 0.0)),
    (3, dict(escape_radius=60.0), (0.0, 15.0, math.pi / 2, 0.0)),
])
def test_adaptive_trace_on_gpu(ctx, O, kind, kw, obs):
    """geon_trace_rays_adaptive against the oracle's restatement of the same
    Cash–Karp controller: Minkowski (no transcendental anywhere) bit for bit;
    elsewhere the library's sin/cos and tanh differ by ulps, so classes must
    agree and escape directions lie within 1e-8 rad. Morris–Thorne's kernel Γ
    differs from the oracle's by more ulps (stage sines by rotation); where an
    error estimate sits at the acceptance threshold the two take different
    step sequences, each accurate to the method's own error (a few 1e-7 rad
    against a 10× finer fixed-step trace), so there the bound is 1e-6 rad."""
    import geon_amd as geon
    om = O.metric(kind, **kw)
    y0 = np.array(O.camera_rays(om, O.camera(obs, 60.0, 24, 16)))
    ref = O.trace_rays_adaptive(om, y0, 0.05, 1e-9, 1e-12, 100000)
    got = geon.trace_rays_adaptive(om, y0, 0.05, 1e-9, 1e-12, 100000, ctx=ctx)
    assert np.array_equal(got["cls"], ref["cls"])
    esc = ref["cls"] == O.ESCAPED
    ang = _angle(got["dir"][esc], ref["dir"][esc]) if esc.any() else np.zeros(1)
    print(f"adaptive kind={kind}: rays {len(y0)}, steps {int(got['steps'].sum())} "
          f"(oracle {int(ref['steps'].sum())}), rejected {int(got['rejected'].sum())}, "
          f"max angle {ang.max():.3e}", flush=True)
    if kind == 0:
        assert np.array_equal(got["y"].view(np.uint64), ref["y"].view(np.uint64))
        assert np.array_equal(got["steps"], ref["steps"].astype(np.uint32))
    assert ang.max() < (1e-6 if kind == 3 else 1e-8)

    # the device-pointer entry point on torch's stream: the same kernel, the same bits
    import torch
    n = len(y0)
    dy0 = torch.from_numpy(np.ascontiguousarray(y0.T)).cuda()
    dcls = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    dsteps = torch.zeros(n, dtype=torch.int32, device="cuda")
    drej = torch.zeros(n, dtype=torch.int32, device="cuda")
    dyf = torch.zeros(8 * n, dtype=torch.float64, device="cuda")
    ddir = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    # a stream of the caller's own (torch's default stream is the null handle,
    # which the library takes as "use the context's stream")
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    geon._abi.check(geon.lib().geon_trace_rays_adaptive_dev(
        ctx.handle, geon.C.byref(geon._cmetric(om)), dy0.data_ptr(), n, 0.05, 1e-9, 1e-12, 100000,
        dcls.data_ptr(), dsteps.data_ptr(), drej.data_ptr(), dyf.data_ptr(), ddir.data_ptr(),
        st.cuda_stream), "geon_trace_rays_adaptive_dev")
    st.synchronize()
    assert np.array_equal(dcls.cpu().numpy(), got["cls"])
    assert np.array_equal(dsteps.cpu().numpy().view(np.uint32), got["steps"])
    assert np.array_equal(drej.cpu().numpy().view(np.uint32), got["rejected"])
    assert np.array_equal(dyf.cpu().numpy().reshape(8, n).T.view(np.uint64), got["y"].view(np.uint64))
    assert np.array_equal(ddir.cpu().numpy().reshape(3, n).T.view(np.uint64),
                          got["dir"].view(np.uint64))


def test_shadow_bisection_readme(ctx, O, anchors):
    """validate_shadow: 60 halvings, step 0.01, max 200,000 (README.md:43-45)."""
    import geon_amd as geon
    a, b = geon.shadow_bisect([1.0], [20.0], step=0.01, max_steps=200_000, iters=60, ctx=ctx)
    rd = anchors["readme"]
    assert f"{b[0]:.6f}" == rd["validate_shadow_measured"]["b_c_text"]
    assert f"{math.degrees(a[0]):.4f}" == rd["validate_shadow_measured"]["alpha_c_deg_text"]
    a_ref, b_ref = O.validate_shadow(1.0, 20.0, 0.01, 200_000, 60)
    assert abs(b[0] - b_ref) < 1e-9
