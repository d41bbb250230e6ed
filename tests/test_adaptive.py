"""Adaptive stepping (Cash–Karp RK4(5)): the step the reference names next
(integrator.rs:5-7), not a reference method, so its oracle restatement
(oracle/geon_oracle.c, or_trace_adaptive_rays) is pinned by physics instead:
straight lines in flat space, conserved energy and angular momentum in
Schwarzschild, agreement with fine fixed-step traces; and the kernels' code,
built for the host, equals that restatement bit for bit where the host build
of the fixed-step path equals the oracle (Minkowski, Schwarzschild,
Alcubierre)."""
import ctypes as C
import math

import numpy as np
import pytest

from tests.test_device_host import hd, _p  # noqa: F401  (fixture and helper)


def _adaptive_host(hd, m, y0, h0, rtol, atol, max_steps):
    hd.hd_trace_adaptive.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_uint32, C.c_void_p, C.c_void_p]
    Y = np.array(y0, dtype=np.float64)
    cls, steps, rej = [], [], []
    for i in range(len(Y)):
        y = np.ascontiguousarray(Y[i])
        s, r = C.c_uint32(), C.c_uint32()
        cls.append(hd.hd_trace_adaptive(C.addressof(m), y.ctypes.data, h0, rtol, atol, max_steps,
                                        C.addressof(s), C.addressof(r)))
        steps.append(s.value)
        rej.append(r.value)
        Y[i] = y
    return np.array(cls), np.array(steps), np.array(rej), Y


def test_minkowski_lines_are_exact_in_direction(O):
    m = O.metric(O.MINKOWSKI, escape_radius=100.0)
    y0 = np.array(O.camera_rays(m, O.camera([0, 20, 0, 0], 60.0, 8, 8)))
    r = O.trace_rays_adaptive(m, y0, 0.05, 1e-10, 1e-12, 100000)
    assert (r["cls"] == O.ESCAPED).all() and (r["rejected"] == 0).all()
    # Γ = 0: every step is the straight line, the momentum never changes
    assert np.array_equal(r["y"][:, 4:], y0[:, 4:])
    lam = (r["y"][:, 0] - y0[:, 0]) / y0[:, 4]
    assert np.abs(r["y"][:, 1:4] - (y0[:, 1:4] + lam[:, None] * y0[:, 5:8])).max() < 1e-10


def test_cash_karp_orders(O):
    """The tableau's orders, which pin its constants: a wrong coefficient
    drops the local error of the fifth-order solution from O(h⁶) or that of
    the embedded fourth-order one from O(h⁵). Local errors against the same
    step taken as 256 substeps (error ~256·(h/256)⁶, negligible), on a bent
    Schwarzschild orbit near r = 4M."""
    m = O.metric(O.SCHWARZSCHILD, horizon_pad=1.005, escape_radius=100.0)
    r0 = 4.0
    pt = 1.0 / (1 - 2 / r0)
    pphi = 0.3 / r0 ** 2
    pr = -math.sqrt((1 - 2 / r0) * ((1 - 2 / r0) * pt ** 2 - r0 ** 2 * pphi ** 2))
    y0 = np.array([0.0, r0, math.pi / 2, 0.0, pt, pr, 0.0, pphi])

    def fine(h, n=256):
        y = y0.copy()
        for _ in range(n):
            y, _ = O.ck_step(m, y, h / n)
        return y

    e5, est = [], []
    for h in (0.4, 0.2, 0.1):
        yn, err = O.ck_step(m, y0, h)
        e5.append(np.abs(yn - fine(h)).max())
        est.append(np.abs(err).max())
    # halving h divides the local error by 2⁶ and the estimate (the fourth-
    # order solution's local error) by 2⁵, up to the next terms
    for a, b in zip(e5, e5[1:]):
        assert 2 ** 5.3 < a / b < 2 ** 6.7, e5
    for a, b in zip(est, est[1:]):
        assert 2 ** 4.5 < a / b < 2 ** 5.5, est


def test_schwarzschild_conserves_energy_and_angular_momentum(O):
    m = O.metric(O.SCHWARZSCHILD, horizon_pad=1.005, escape_radius=100.0)
    # a horizontal fan: the rays stay in the equatorial plane with L = r² p^φ ≠ 0
    y0 = np.array(O.camera_rays(m, O.camera([0, 30, math.pi / 2, 0], 60.0, 24, 1)))
    assert np.abs(y0[:, 7]).max() > 0 and np.abs(y0[:, 6]).max() == 0
    r = O.trace_rays_adaptive(m, y0, 0.05, 1e-11, 1e-13, 200000)
    esc = r["cls"] == O.ESCAPED
    assert esc.sum() > 10
    def E(y):
        return (1 - 2 / y[:, 1]) * y[:, 4]
    def L(y):
        return y[:, 1] ** 2 * np.sin(y[:, 2]) ** 2 * y[:, 7]
    for q in (E, L):
        a, b = q(y0)[esc], q(r["y"])[esc]
        assert (np.abs(a - b) <= 1e-8 * np.maximum(1.0, np.abs(a))).all(), q.__name__


@pytest.mark.parametrize("kind,kw,obs", [
    ("schwarzschild", dict(horizon_pad=1.005, escape_radius=100.0), [0, 30, math.pi / 2, 0]),
    ("alcubierre", dict(bubble_thickness=0.125, ship_velocity=0.5, escape_radius=60.0), [0, 10, 0, 0]),
    ("morris_thorne", dict(escape_radius=60.0), [0, 15, math.pi / 2, 0]),
])
def test_agrees_with_fine_fixed_steps(O, kind, kw, obs):
    k = {"schwarzschild": O.SCHWARZSCHILD, "alcubierre": O.ALCUBIERRE,
         "morris_thorne": O.MORRIS_THORNE}[kind]
    m = O.metric(k, **kw)
    y0 = np.array(O.camera_rays(m, O.camera(obs, 60.0, 12, 12)))
    ad = O.trace_rays_adaptive(m, y0, 0.05, 1e-10, 1e-12, 200000)
    fx = O.trace_rays(m, y0, 0.005, 400000)
    assert np.array_equal(ad["cls"], fx["cls"])
    esc = ad["cls"] == O.ESCAPED
    d = np.clip((ad["dir"][esc] * fx["dir"][esc]).sum(1), -1, 1)
    assert np.arccos(d).max() < 1e-5


@pytest.mark.parametrize("kind,kw,obs", [
    ("minkowski", dict(escape_radius=100.0), [0, 20, 0, 0]),
    ("schwarzschild", dict(horizon_pad=1.005, escape_radius=100.0), [0, 30, math.pi / 2, 0]),
    ("alcubierre", dict(bubble_thickness=0.125, ship_velocity=0.5, escape_radius=60.0), [0, 10, 0, 0]),
])
def test_kernel_code_equals_oracle_bit_for_bit(hd, O, kind, kw, obs):
    k = {"minkowski": O.MINKOWSKI, "schwarzschild": O.SCHWARZSCHILD, "alcubierre": O.ALCUBIERRE}[kind]
    m = O.metric(k, **kw)
    y0 = np.array(O.camera_rays(m, O.camera(obs, 60.0, 10, 10)))
    ref = O.trace_rays_adaptive(m, y0, 0.05, 1e-9, 1e-12, 100000)
    cls, steps, rej, Y = _adaptive_host(hd, m, y0, 0.05, 1e-9, 1e-12, 100000)
    assert np.array_equal(cls, ref["cls"]) and np.array_equal(steps, ref["steps"])
    assert np.array_equal(rej, ref["rejected"])
    assert np.array_equal(Y.view(np.uint64), ref["y"].view(np.uint64))


def test_non_finite_trial_step_is_rejected(hd, O):
    """A Schwarzschild ray falling in from r = 2.5M with h0 = 2.5: the first
    Cash–Karp trial step reaches r ≤ 0 and comes out non-finite (its error
    norm drops the NaN under fmax). The controller rejects it with h/10 and
    retries instead of ending the ray as diverged: the ray is captured after
    a few rejections, in the oracle and in the kernels' code built for the
    host alike."""
    m = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.001, escape_radius=100.0)
    f = 1.0 - 2.0 / 2.5
    y0 = np.array([[0.0, 2.5, math.pi / 2, 0.0, 1.0 / f, -1.0, 0.0, 0.0]])
    yn, err = O.ck_step(m, y0[0], 2.5)
    assert not (np.isfinite(yn).all() and np.isfinite(err).all())
    ref = O.trace_rays_adaptive(m, y0, 2.5, 1e-9, 1e-12, 1000)
    assert ref["cls"][0] == O.CAPTURED and ref["rejected"][0] >= 1 and ref["steps"][0] > 0
    cls, steps, rej, Y = _adaptive_host(hd, m, y0, 2.5, 1e-9, 1e-12, 1000)
    assert cls[0] == ref["cls"][0] and steps[0] == ref["steps"][0] and rej[0] == ref["rejected"][0]
    assert np.array_equal(Y.view(np.uint64), ref["y"].view(np.uint64))
