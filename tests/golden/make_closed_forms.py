"""Closed-form fixtures for the oracle and the HIP kernels, derived with sympy.

Independent of both implementations: the metric components are written
from the line elements in the reference's doc comments
(schwarzschild.rs:3, alcubierre.rs:4-6, morris_thorne.rs:3) and everything
else (inverse metric, Christoffel symbols, their derivatives, Riemann,
Ricci, Einstein) is derived symbolically and evaluated at 40 digits.

    python tests/golden/make_closed_forms.py   ->  tests/golden/closed_forms.json
"""
import json
import os

import sympy as sp

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_forms.json")
DPS = 40

t, x1, x2, x3 = sp.symbols("t x1 x2 x3", real=True)
X = [t, x1, x2, x3]


def schwarzschild(M):
    r, th = x1, x2
    f = 1 - 2 * M / r
    return sp.diag(-f, 1 / f, r**2, r**2 * sp.sin(th) ** 2)


def morris_thorne(b0):
    l, th = x1, x2
    r2 = b0**2 + l**2
    return sp.diag(-1, 1, r2, r2 * sp.sin(th) ** 2)


def alcubierre(R, s, v, xs):
    # ds² = −dt² + (dx − v f dt)² + dy² + dz²
    rs = sp.sqrt((x1 - xs) ** 2 + x2**2 + x3**2)
    f = (sp.tanh((rs + R) / s) - sp.tanh((rs - R) / s)) / (2 * sp.tanh(R / s))
    g = sp.zeros(4, 4)
    g[0, 0] = -1 + v**2 * f**2
    g[0, 1] = g[1, 0] = -v * f
    g[1, 1] = g[2, 2] = g[3, 3] = 1
    return g


def christoffel_sym(g):
    ginv = g.inv()
    gam = [[[0] * 4 for _ in range(4)] for _ in range(4)]
    for mu in range(4):
        for a in range(4):
            for b in range(a, 4):
                e = sum(ginv[mu, s] * (sp.diff(g[s, b], X[a]) + sp.diff(g[s, a], X[b])
                                       - sp.diff(g[a, b], X[s])) for s in range(4)) / 2
                gam[mu][a][b] = gam[mu][b][a] = e
    return ginv, gam


def evalf(e, pt):
    return float(sp.N(e.subs(pt) if hasattr(e, "subs") else e, DPS))


def numeric_tables(g, pts, want_curvature):
    ginv, gam = christoffel_sym(g)
    rows = []
    for p in pts:
        sub = dict(zip(X, [sp.nsimplify(v, rational=True) for v in p]))
        row = dict(x=list(p))
        row["g"] = [[evalf(g[i, j], sub) for j in range(4)] for i in range(4)]
        row["g_inv"] = [[evalf(ginv[i, j], sub) for j in range(4)] for i in range(4)]
        row["dg"] = [[[evalf(sp.diff(g[m, n], X[a]), sub) for n in range(4)] for m in range(4)]
                     for a in range(4)]
        row["christoffel"] = [[[evalf(gam[m][a][b], sub) for b in range(4)] for a in range(4)]
                              for m in range(4)]
        if want_curvature:
            G = [[[[None] * 4 for _ in range(4)] for _ in range(4)] for _ in range(4)]
            gam_n = [[[sp.N(gam[m][a][b].subs(sub), DPS) for b in range(4)] for a in range(4)]
                     for m in range(4)]
            dgam = [[[[sp.N(sp.diff(gam[m][a][b], X[d]).subs(sub), DPS) for b in range(4)]
                      for a in range(4)] for m in range(4)] for d in range(4)]
            riem = [[[[0] * 4 for _ in range(4)] for _ in range(4)] for _ in range(4)]
            for rho in range(4):
                for sg in range(4):
                    for mu in range(4):
                        for nu in range(4):
                            val = dgam[mu][rho][nu][sg] - dgam[nu][rho][mu][sg]
                            for lam in range(4):
                                val += (gam_n[rho][mu][lam] * gam_n[lam][nu][sg]
                                        - gam_n[rho][nu][lam] * gam_n[lam][mu][sg])
                            riem[rho][sg][mu][nu] = val
            ric = [[sum(riem[a][m][a][n] for a in range(4)) for n in range(4)] for m in range(4)]
            gi = [[sp.N(ginv[i, j].subs(sub), DPS) for j in range(4)] for i in range(4)]
            gg = [[sp.N(g[i, j].subs(sub), DPS) for j in range(4)] for i in range(4)]
            R = sum(gi[m][n] * ric[m][n] for m in range(4) for n in range(4))
            row["ricci_scalar"] = float(R)
            row["einstein"] = [[float(ric[m][n] - gg[m][n] * R / 2) for n in range(4)]
                               for m in range(4)]
            row["riemann"] = [[[[float(riem[a][b][c][d]) for d in range(4)] for c in range(4)]
                               for b in range(4)] for a in range(4)]
            del G
        rows.append(row)
    return rows


def main():
    hp = sp.pi / 2
    fx = {}
    fx["schwarzschild"] = dict(
        params=dict(mass=1.0),
        points=numeric_tables(schwarzschild(1),
                              [(0.0, 5.0, float(hp), 0.3), (0.0, 3.7, 1.1, 2.0),
                               (0.0, 12.0, 0.4, 0.0), (0.0, 2.5, 2.9, -1.0)], True))
    fx["morris_thorne"] = dict(
        params=dict(throat_radius=1.0),
        points=numeric_tables(morris_thorne(1),
                              [(0.0, 0.0, float(hp), 0.0), (0.0, 1.3, 1.0, 0.5),
                               (0.0, -2.2, 2.5, 0.0), (0.0, 15.0, float(hp), 0.0)], True))
    fx["morris_thorne_b2"] = dict(
        params=dict(throat_radius=2.0),
        points=numeric_tables(morris_thorne(2), [(0.0, 0.5, 1.2, 0.0)], True))
    fx["alcubierre_ec"] = dict(
        params=dict(bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0, ship_x=0.0),
        points=numeric_tables(alcubierre(2, sp.Rational(1, 2), 2, 0),
                              [(0.0, 2.4, 0.0, 0.0), (0.0, 1.5, 0.7, -0.3),
                               (0.0, -2.1, 0.2, 0.5)], True))
    fx["alcubierre_render"] = dict(
        params=dict(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5, ship_x=0.0),
        points=numeric_tables(alcubierre(2, sp.Rational(1, 8), sp.Rational(1, 2), 0),
                              [(0.0, 2.05, 0.1, 0.0), (0.0, 1.9, 0.3, 0.2),
                               (0.0, 0.3, 0.1, 0.2), (0.0, 6.0, 0.0, 0.0)], True))
    # Shadow: b_c = 3√3 M, α_c = asin(b_c √(1−2M/r)/r) (validate_shadow.rs:45-50)
    bc = 3 * sp.sqrt(3)
    f = 1 - sp.Rational(2, 20)
    fx["shadow"] = dict(mass=1.0, r_obs=20.0, b_c=float(sp.N(bc, DPS)),
                        alpha_c=float(sp.N(sp.asin(bc * sp.sqrt(f) / 20), DPS)))
    with open(OUT, "w", encoding="utf-8") as fh:
        json.dump(fx, fh)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
