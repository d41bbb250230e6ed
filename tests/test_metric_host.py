"""The Metric trait on the host through the C ABI (geon_metric_*), against the
oracle's restatement of src/metric/*.rs and the trait defaults of
src/metric/mod.rs:19-109, and against closed forms.

g, ∂g and Γ are compared bit for bit. g⁻¹ is compared by value: its
structural zeros are +0, where the reference's general cofactor expansion
(nalgebra, types.rs:34-46) can leave −0 off the diagonal.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

CASES = [
    (O.SCHWARZSCHILD, {}, lambda r: [0.0, r.uniform(2.5, 40), r.uniform(0.1, 3.0), r.uniform(-3, 3)]),
    (O.MORRIS_THORNE, {}, lambda r: [0.0, r.uniform(-20, 20), r.uniform(0.1, 3.0), r.uniform(-3, 3)]),
    (O.ALCUBIERRE, dict(bubble_thickness=0.125, ship_velocity=0.5), lambda r: [0.0, *r.uniform(-4, 4, 3)]),
    (O.ALCUBIERRE, {}, lambda r: [0.0, *r.uniform(-4, 4, 3)]),
    (O.MINKOWSKI, {}, lambda r: [0.0, *r.uniform(-4, 4, 3)]),
]


def _call(fn, om, x, width):
    import geon_amd as geon
    out = np.zeros(width)
    st = fn(C.byref(geon._cmetric(om)), np.ascontiguousarray(x).ctypes.data, out.ctypes.data)
    assert st == 0
    return out


@pytest.mark.parametrize("kind,kw,gen", CASES)
def test_metric_methods_match_oracle(kind, kw, gen):
    import geon_amd as geon
    L = geon.lib()
    om = O.metric(kind, **kw)
    rng = np.random.default_rng(kind * 7 + len(kw))
    for _ in range(200):
        x = np.array(gen(rng))
        for fn, ref, w in ((L.geon_metric_g, O.g, 16), (L.geon_metric_dg, O.dg, 64),
                           (L.geon_metric_christoffel, O.christoffel, 64)):
            got = _call(fn, om, x, w)
            want = np.asarray(ref(om, x)).reshape(-1)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (fn.__name__, x)
        got = _call(L.geon_metric_g_inv, om, x, 16)
        st, want = O.g_inv(om, x)
        assert st == 0 and np.array_equal(got, want.reshape(-1)), x


def test_metric_closed_forms_and_terminate():
    """g·g⁻¹ = I, Γ symmetric in its lower indices, Schwarzschild Γ^r_tt =
    M f / r², terminate and asymptotic_direction per metric."""
    import geon_amd as geon
    L = geon.lib()
    sm = geon.Schwarzschild(mass=1.0)
    x = np.array([0.0, 7.0, 1.1, 0.3])
    g = _call(L.geon_metric_g, sm, x, 16).reshape(4, 4)
    gi = _call(L.geon_metric_g_inv, sm, x, 16).reshape(4, 4)
    assert np.abs(g @ gi - np.eye(4)).max() < 1e-15
    gam = _call(L.geon_metric_christoffel, sm, x, 64).reshape(4, 4, 4)
    f = 1 - 2 / 7.0
    assert math.isclose(gam[1, 0, 0], f / 49.0, rel_tol=1e-15)
    assert np.array_equal(gam, gam.transpose(0, 2, 1))
    cls = C.c_int32()
    for m, pt, want in ((sm, [0, 2.0001, 1, 0], geon.CAPTURED), (sm, [0, 1000.5, 1, 0], geon.ESCAPED),
                        (sm, [0, 50.0, 1, 0], -1), (sm, [0, math.nan, 1, 0], geon.DIVERGED),
                        (geon.MorrisThorne(), [0, -1000.5, 1, 0], geon.ESCAPED),
                        (geon.Alcubierre(), [0, 600.0, 900.0, 0], geon.ESCAPED),
                        (geon.Minkowski(), [0, 1.0, 2.0, 3.0], -1)):
        pt = np.array(pt, dtype=np.float64)
        assert L.geon_metric_terminate(C.byref(geon._cmetric(m)), pt.ctypes.data, C.byref(cls)) == 0
        assert cls.value == want, (m, pt)
    # a radial outgoing ray at θ = π/2, φ = 0 points along +x; Alcubierre's
    # chart is Cartesian, so its direction is p's spatial part normalised
    d = np.zeros(3)
    xr = np.array([0.0, 50.0, math.pi / 2, 0.0])
    pr = np.array([1.0, 1.0, 0.0, 0.0])
    assert L.geon_metric_asymptotic_direction(C.byref(sm.to_c()), xr.ctypes.data, pr.ctypes.data,
                                              d.ctypes.data) == 0
    assert np.abs(d - [1.0, 0.0, 0.0]).max() < 1e-15
    pa = np.array([1.0, 0.0, 3.0, 4.0])
    assert L.geon_metric_asymptotic_direction(C.byref(geon.Alcubierre().to_c()), xr.ctypes.data,
                                              pa.ctypes.data, d.ctypes.data) == 0
    assert np.array_equal(d, [0.0, 0.6, 0.8])


def test_camera_new_radians():
    """Camera::new stores fov_y_deg.to_radians() (camera.rs:29-36)."""
    import geon_amd as geon
    for deg in (60.0, 8.0, 90.0, 1e-3, 179.0):
        cam = geon.Camera((0.0, 30.0, 1.0, 0.0), deg, 3, 2)
        assert cam.fov_y_rad == deg * (math.pi / 180.0)
        assert O.camera(cam.observer, deg, 3, 2).fov_y_rad == cam.fov_y_rad
