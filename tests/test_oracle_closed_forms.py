"""Oracle metric layer and curvature vs sympy closed forms (tests/golden/closed_forms.json).

The fixtures are derived symbolically from the line elements alone, so a
wrong index, sign or transposition in g, ∂g, g⁻¹, Γ or Riemann fails here.
Tolerances follow the reference's own numerics: ∂g and the generic Γ are
central differences with h = 1e-5 (error O(h²) + O(ε/h)); Riemann is a
difference of those (noise ~ε/h² relative to g).
"""
import numpy as np
import pytest

CASES = {
    "schwarzschild": (1, dict(mass=1.0)),
    "morris_thorne": (3, dict(throat_radius=1.0)),
    "morris_thorne_b2": (3, dict(throat_radius=2.0)),
    "alcubierre_ec": (2, dict(bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0,
                              ship_x=0.0)),
    "alcubierre_render": (2, dict(bubble_radius=2.0, bubble_thickness=0.125,
                                  ship_velocity=0.5, ship_x=0.0)),
}


def _points(closed_forms):
    for name, (kind, kw) in CASES.items():
        for p in closed_forms[name]["points"]:
            yield name, kind, kw, p


def test_g_ginv_dg_christoffel(O, closed_forms):
    n = 0
    for name, kind, kw, p in _points(closed_forms):
        m = O.metric(kind, **kw)
        x = p["x"]
        g = np.array(p["g"])
        assert np.allclose(O.g(m, x), g, rtol=1e-15, atol=1e-15), name
        st, gi = O.g_inv(m, x)
        assert st == 0
        assert np.allclose(gi, np.array(p["g_inv"]), rtol=1e-14, atol=1e-15), name
        dg = np.array(p["dg"])
        assert np.abs(O.dg(m, x) - dg).max() <= 1e-8 * max(1.0, np.abs(dg).max()), name
        gam = np.array(p["christoffel"])
        tol = (1e-14 if kind == 1 else 1e-8) * max(1.0, np.abs(gam).max())
        assert np.abs(O.christoffel(m, x) - gam).max() <= tol, (name, x)
        n += 1
    assert n == 16


def test_christoffel_symmetry_and_exact_zeros(O):
    """Γ^μ_αβ = Γ^μ_βα; derivatives along coordinates a metric ignores are exactly 0."""
    mt = O.metric(O.MORRIS_THORNE, throat_radius=1.0)
    d = O.dg(mt, [0.3, 1.7, 0.9, 2.2])
    assert np.all(d[0] == 0.0) and np.all(d[3] == 0.0)
    alc = O.metric(O.ALCUBIERRE, bubble_thickness=0.125, ship_velocity=0.5)
    assert np.all(O.dg(alc, [1.0, 1.9, 0.2, -0.3])[0] == 0.0)
    for m, x in [(mt, [0, 1.7, 0.9, 2.2]), (alc, [0, 1.9, 0.2, -0.3]),
                 (O.metric(O.SCHWARZSCHILD), [0, 4.0, 0.7, 0.1])]:
        G = O.christoffel(m, x)
        assert np.array_equal(G, np.transpose(G, (0, 2, 1)))


def test_inv4_identity_and_singular(O):
    rng = np.random.default_rng(7)
    for _ in range(50):
        a = rng.normal(size=(4, 4))
        a = a + a.T + 8 * np.eye(4)
        st, inv = O.inv4(a)
        assert st == 0
        assert np.abs(a @ inv - np.eye(4)).max() < 1e-14
        assert np.allclose(inv, np.linalg.inv(a), rtol=1e-13, atol=1e-15)
    st, _ = O.inv4(np.diag([1.0, 2.0, 0.0, 3.0]))
    assert st == -4  # GEON_E_SINGULAR (types.rs:39 panics here)


def test_curvature_vs_sympy(O, closed_forms):
    for name, kind, kw, p in _points(closed_forms):
        if "riemann" not in p:
            continue
        m = O.metric(kind, **kw)
        x = p["x"]
        R = np.array(p["riemann"])
        scale = max(1.0, abs(np.array(p["g"])).max())
        # nested FD noise: ~1e-16·g/h² plus the O(h²) truncation
        tol = 5e-5 * scale
        assert np.abs(O.riemann(m, x) - R).max() <= tol, (name, x)
        assert np.abs(O.einstein(m, x) - np.array(p["einstein"])).max() <= tol, (name, x)
        assert abs(O.ricci_scalar(m, x) - p["ricci_scalar"]) <= tol, (name, x)


def test_riemann_antisymmetry(O):
    for m, x in [(O.metric(O.MORRIS_THORNE), [0, 0.4, 1.1, 0]),
                 (O.metric(O.ALCUBIERRE), [0, 1.5, 0.7, -0.3]),
                 (O.metric(O.SCHWARZSCHILD), [0, 6.0, 1.0, 0.2])]:
        R = O.riemann(m, x)
        assert np.abs(R + np.transpose(R, (0, 1, 3, 2))).max() < 1e-9 * max(1, abs(R).max())
        ric = O.ricci(m, x)
        assert np.abs(ric - ric.T).max() < 1e-5


@pytest.mark.parametrize("l", [0.0, 0.7, -1.9, 3.0])
def test_morris_thorne_einstein_closed_form(O, l):
    """G_tt = G_ll = −b0²/(b0²+l²)², G_θθ = b0²/(b0²+l²), G_φφ = G_θθ sin²θ."""
    b0, th = 1.0, 1.2
    m = O.metric(O.MORRIS_THORNE, throat_radius=b0)
    G = O.einstein(m, [0.0, l, th, 0.0])
    r2 = b0 * b0 + l * l
    want = np.diag([-b0**2 / r2**2, -b0**2 / r2**2, b0**2 / r2, b0**2 / r2 * np.sin(th)**2])
    assert np.abs(G - want).max() < 5e-6
