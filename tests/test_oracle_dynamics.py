"""Oracle integration layer (rhs, rk4_step, trace) pinned by what the mathematics fixes.

- Minkowski: rays are straight lines; the escape step count is the first n
  with |x0 + n h p|² > R², and |g(p,p)| stays at rounding level (README.md:50).
- Schwarzschild radial null rays: p^r is constant and
  t(r) = (r0 − r) + 2M ln((r0 − 2M)/(r − 2M)), which checks the RK4
  coefficients through the 4th-order convergence of t.
- Conserved E = f p^t and L = r² sin²θ p^φ drift only at the RK4 rate.
- trace semantics (geodesic.rs:62-117): terminate checked before stepping,
  NaN guard returns the pre-step state, max_steps → Diverged.
- Capture vs escape of camera rays follows the analytic b_c = 3√3 M.
"""
import math

import numpy as np
import pytest


def test_minkowski_straight_lines_and_escape_count(O):
    m = O.metric(O.MINKOWSKI)
    rng = np.random.default_rng(3)
    h = 0.05
    for _ in range(20):
        x0 = rng.uniform(-20, 20, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        y0 = np.concatenate([[0.0], x0, [1.0], n])
        cls, steps, y, resid = O.trace(m, y0, h, 100_000)
        assert cls == O.ESCAPED
        # first n with |x0 + n h p|² > 1e6, computed in extended precision
        a = float(h * h)
        b = 2 * h * float(x0 @ n)
        c = float(x0 @ x0) - 1e6
        n_exact = (-b + math.sqrt(b * b - 4 * a * c)) / (2 * a)
        assert steps == math.floor(n_exact) + 1
        assert np.allclose(y[1:4], x0 + steps * h * n, rtol=0, atol=1e-9)
        assert np.array_equal(y[4:], y0[4:])
        assert resid < 1e-15


def _radial(r0, sign, mass=1.0):
    f = 1.0 - 2.0 * mass / r0
    return np.array([0.0, r0, math.pi / 2, 0.0, 1.0 / math.sqrt(f), sign * math.sqrt(f), 0.0, 0.0])


def test_rk4_fourth_order_radial_infall(O):
    m = O.metric(O.SCHWARZSCHILD, mass=1.0)
    r0, lam = 10.0, 6.0
    errs, rerr = [], []
    for h in (0.2, 0.1, 0.05):
        y = _radial(r0, -1.0)
        E = math.sqrt(1 - 2 / r0)
        for _ in range(int(round(lam / h))):
            y = O.rk4_step(m, y, h)
        r = r0 - E * lam
        t_exact = (r0 - r) + 2.0 * math.log((r0 - 2.0) / (r - 2.0))
        errs.append(abs(y[0] - t_exact))
        rerr.append(abs(y[1] - r))           # p^r conserved → r linear in λ
    assert 12.0 < errs[0] / errs[1] < 20.0 and 12.0 < errs[1] / errs[2] < 20.0
    assert rerr[0] < 1e-7 and 12.0 < rerr[0] / rerr[1] < 20.0


def test_radial_capture_and_escape_step_counts(O):
    m = O.metric(O.SCHWARZSCHILD, mass=1.0)
    h = 0.05
    for r0 in (7.3, 12.9, 31.7):
        E = math.sqrt(1 - 2 / r0)
        cls, steps, y, _ = O.trace(m, _radial(r0, -1.0), h, 100_000)
        assert cls == O.CAPTURED
        assert steps == math.floor((r0 - 2.002) / (E * h)) + 1
        cls, steps, y, _ = O.trace(m, _radial(r0, +1.0), h, 100_000)
        assert cls == O.ESCAPED
        assert steps == math.floor((1000.0 - r0) / (E * h)) + 1
        assert y[1] > 1000.0


def test_energy_angular_momentum_drift(O):
    m = O.metric(O.SCHWARZSCHILD, mass=1.0)
    y0 = O.equatorial_ray(1.0, 20.0, 0.5)
    drift = []
    for h in (0.1, 0.05):
        y = y0.copy()
        E0 = (1 - 2 / y[1]) * y[4]
        L0 = y[1] ** 2 * y[7]
        for _ in range(int(round(20.0 / h))):
            y = O.rk4_step(m, y, h)
        drift.append(max(abs((1 - 2 / y[1]) * y[4] - E0), abs(y[1] ** 2 * y[7] - L0)))
    assert drift[0] < 1e-6 and 8.0 < drift[0] / drift[1] < 40.0


def test_trace_semantics(O):
    m = O.metric(O.SCHWARZSCHILD, mass=1.0)
    # terminate before any step
    cls, steps, y, _ = O.trace(m, _radial(1500.0, -1.0), 0.05, 10)
    assert (cls, steps) == (O.ESCAPED, 0)
    cls, steps, y, _ = O.trace(m, _radial(2.0015, 1.0), 0.05, 10)
    assert (cls, steps) == (O.CAPTURED, 0)
    # non-finite r → Diverged at n = 0
    y0 = _radial(10.0, -1.0)
    y0[1] = math.inf
    cls, steps, y, _ = O.trace(m, y0, 0.05, 10)
    assert (cls, steps) == (O.DIVERGED, 0)
    # NaN guard: returns the pre-step state and n
    y0 = _radial(10.0, -1.0)
    y0[6] = math.nan
    cls, steps, y, _ = O.trace(m, y0, 0.05, 10)
    assert (cls, steps) == (O.DIVERGED, 0)
    assert y[1] == 10.0 and math.isnan(y[6])
    # max_steps → reference Diverged with steps_taken == max_steps
    cls, steps, y, _ = O.trace(m, _radial(30.0, -1.0), 0.05, 7)
    assert (cls, steps) == (O.MAXSTEPS, 7)
    y7 = _radial(30.0, -1.0)
    for _ in range(7):
        y7 = O.rk4_step(m, y7, 0.05)
    assert np.array_equal(y, y7)


def test_camera_rays_capture_vs_analytic_shadow(O):
    """b = r sinα/√f: b > b_c + δ escapes; b < b_c − δ is captured, except where the
    reference's fixed step lands an RK4 stage on r ≈ 2M: f → 0 blows the state up to a
    huge finite r that terminate() then reads as Escaped (geodesic.rs:71-101). Those
    artefacts end far beyond the escape sphere, unlike genuine escapes."""
    m = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    cam = O.camera([0.0, 30.0, math.pi / 2, 0.0], 60.0, 48, 48)
    y0 = O.camera_rays(m, cam)
    res = O.trace_rays(m, y0, 0.05, 4000)
    f = 1 - 2 / 30.0
    st, tet = O.tetrad(m, cam.observer)
    e1 = tet[1]
    artefacts = 0
    for idx in range(len(y0)):
        p = y0[idx, 4:]
        n1 = (p[1] - tet[0][1]) / e1[1]          # local radial component of the direction
        sin_a = math.sqrt(max(0.0, 1 - n1 * n1))
        b = 30.0 * sin_a / math.sqrt(f)
        if b < 3 * math.sqrt(3) - 0.05:
            if res["cls"][idx] != O.CAPTURED:
                assert res["cls"][idx] == O.ESCAPED and res["y"][idx][1] > 1e4
                artefacts += 1
        elif b > 3 * math.sqrt(3) + 0.05:
            assert res["cls"][idx] == O.ESCAPED and res["y"][idx][1] < 100.1
            assert res["resid"][idx] < 1e-3
    assert artefacts < 0.05 * len(y0)


def test_trace_rays_matches_single_trace(O):
    m = O.metric(O.MORRIS_THORNE, throat_radius=1.0, escape_radius=40.0)
    cam = O.camera([0.0, 15.0, math.pi / 2, 0.0], 60.0, 8, 6)
    y0 = O.camera_rays(m, cam)
    res = O.trace_rays(m, y0, 0.05, 3000, nthreads=3)
    for i in range(0, len(y0), 5):
        cls, steps, y, resid = O.trace(m, y0[i], 0.05, 3000)
        assert cls == res["cls"][i] and steps == res["steps"][i]
        assert np.array_equal(y, res["y"][i]) and resid == res["resid"][i]
        if cls == O.ESCAPED:
            assert np.array_equal(res["dir"][i], O.asymptotic_direction(m, y[:4], y[4:]))


@pytest.mark.parametrize("e", [-300, -110, 110, 300])
def test_schwarzschild_step_is_scale_free(O, e):
    """Scaling M, t, r and the step by 2^e maps the Schwarzschild discrete map
    onto itself exactly: every Γ entry carries the power of M its dimension
    says (M/(r²f), Mf/r², r − 2M, 1/r, cot θ), so each rounded operation of a
    step scales by a power of two or not at all, and |e| ≤ 300 keeps every
    value normal. Classes, step counts, directions and the scaled final
    states must equal the M = 1 trace bit for bit. A term with a wrong power
    of r or M (M/r for M/r², r² for r − 2M) changes the trajectories at
    some e; the GPU test of the tested-division steps builds on this."""
    runs = []
    for k in (0, e):
        s = 2.0 ** k
        m = O.metric(O.SCHWARZSCHILD, mass=s, horizon_pad=1.005, escape_radius=100.0)
        y0 = O.camera_rays(m, O.camera((0.0, 30.0 * s, math.pi / 2, 0.0), 60.0, 24, 18))
        runs.append((s, y0, O.trace_rays(m, y0, 0.05 * s, 4000)))
    (_, y01, r1), (s, y0, r) = runs
    scale = np.array([s, s, 1.0, 1.0, 1.0, 1.0, 1.0 / s, 1.0 / s])
    assert np.array_equal(y0, y01 * scale)
    assert np.bincount(r1["cls"], minlength=4)[1] > 0 and np.bincount(r1["cls"], minlength=4)[2] > 0
    assert np.array_equal(r["cls"], r1["cls"])
    assert np.array_equal(r["steps"], r1["steps"])
    assert np.array_equal(r["dir"], r1["dir"], equal_nan=True)
    assert np.array_equal(r["y"], r1["y"] * scale, equal_nan=True)
