"""Shared checks of the GPU path against the oracle (imported by the -m gpu tests).

The FP64 bar (the tolerances each test states):
  * per-ray class identical on every ray, except rays whose impact parameter
    lies within 1e-9 of a capture boundary — shown per ray by tracing it with
    the oracle at b − 1e-9 and b + 1e-9 and getting two different classes;
  * step counts identical wherever the class is;
  * escape directions within 1e-9 rad;
  * null residual |g(p,p)|: the initial states are null to 1e-12 (tested
    on the camera rays); at the final state the residual the GPU reports
    equals the oracle's evaluation of that same state (within 1e-14 +
    1e-12·|r|: g's sin/tanh may differ in the last bit), and the GPU keeps
    the constraint as well as the reference does over the whole trace: its
    largest and its median final residual are at most twice the oracle's
    (+1e-12). The per-ray difference between the two traces' residuals is
    reported; it follows the final-state difference (for Morris–Thorne ~1e-11
    relative after thousands of steps, from its rotated stage sines), which
    the direction bar bounds. Captured rays end next to the horizon, where
    the residual grows without bound, and rays the reference's fixed step
    flings far beyond the escape sphere end in amplified rounding noise: both
    keep class and step parity, and their residuals are left out.
"""
from __future__ import annotations

import numpy as np

ANGLE_TOL = 1e-9
B_WINDOW = 1e-9


def angle(a, b):
    cross = np.linalg.norm(np.cross(a, b), axis=-1)
    return np.arctan2(cross, (a * b).sum(axis=-1))


def shifted_impact(y0, mass, db):
    """The Schwarzschild state with the same E = f p^t and position whose
    impact parameter b = L/E is b + db: p^θ and p^φ scaled by (b + db)/b,
    p^r re-solved from the null condition with its sign kept. NaN where no
    such ray exists (turning point)."""
    y = np.array(y0, dtype=np.float64, copy=True)
    r, th = y[:, 1], y[:, 2]
    f = 1.0 - 2.0 * mass / r
    s2 = np.sin(th) ** 2
    E = f * y[:, 4]
    L = r * r * np.sqrt(y[:, 6] ** 2 + s2 * y[:, 7] ** 2)
    b = L / E
    k = (b + db) / b
    y[:, 6] *= k
    y[:, 7] *= k
    pr2 = f * (f * y[:, 4] ** 2 - r * r * (y[:, 6] ** 2 + s2 * y[:, 7] ** 2))
    y[:, 5] = np.where(pr2 >= 0.0, np.sign(y[:, 5]) * np.sqrt(np.abs(pr2)), np.nan)
    return y


def near_capture_boundary(O, om, y0, step, max_steps, db=B_WINDOW):
    """True for rays whose oracle class changes between b − db and b + db."""
    if len(y0) == 0:
        return np.zeros(0, bool)
    lo = O.trace_rays(om, shifted_impact(y0, om.p[0], -db), step, max_steps)
    hi = O.trace_rays(om, shifted_impact(y0, om.p[0], +db), step, max_steps)
    return lo["cls"] != hi["cls"]


def compare_rays(O, om, y0, got, ref, step, max_steps, label, esc_r=None):
    """Assert the FP64 bar for traced rays. got/ref: dicts of cls, steps, dir
    and optionally resid and y (n×8). Returns a summary dict."""
    cls_g, cls_r = np.asarray(got["cls"]), np.asarray(ref["cls"])
    diff = np.nonzero(cls_g != cls_r)[0]
    explained = np.zeros(len(diff), bool)
    if len(diff) and om.kind == O.SCHWARZSCHILD:
        explained = near_capture_boundary(O, om, y0[diff], step, max_steps)
    assert explained.all(), (f"{label}: {int((~explained).sum())} class mismatches outside the "
                             f"1e-9 capture window: rays {diff[~explained][:10]}")
    same = cls_g == cls_r
    sg = np.asarray(got["steps"]).astype(np.int64)
    sr = np.asarray(ref["steps"]).astype(np.int64)
    assert np.array_equal(sg[same], sr[same]), f"{label}: step counts differ"
    esc = same & (cls_r == O.ESCAPED)
    flung = np.zeros(len(cls_r), bool)
    if esc_r is not None and "y" in ref:
        # the reference's fixed-step fling at the horizon: the state jumps far
        # past the escape sphere in one step
        yr = np.asarray(ref["y"])
        flung = esc & (np.linalg.norm(yr[:, 1:4], axis=1) > 10.0 * esc_r) \
            if om.kind in (O.MINKOWSKI, O.ALCUBIERRE) else esc & (np.abs(yr[:, 1]) > 10.0 * esc_r)
    real = esc & ~flung
    a = angle(np.asarray(got["dir"])[real], np.asarray(ref["dir"])[real]) if real.any() else np.zeros(1)
    af = angle(np.asarray(got["dir"])[flung], np.asarray(ref["dir"])[flung]) if flung.any() else np.zeros(1)
    out = dict(rays=len(cls_r), mismatches=len(diff), escaped=int(esc.sum()), flung=int(flung.sum()),
               max_angle=float(a.max()), max_angle_flung=float(af.max()))
    if "resid" in got and "resid" in ref and "y" in got:
        rg, rr = np.asarray(got["resid"]), np.asarray(ref["resid"])
        cmp = same & ~flung & (cls_r != O.CAPTURED) & np.isfinite(rr) & np.isfinite(rg)
        yg = np.ascontiguousarray(np.asarray(got["y"])[cmp])
        at_gpu = O.trace_rays(om, yg, step, 0)["resid"] if cmp.any() else np.zeros(0)
        de = np.abs(rg[cmp] - at_gpu)
        bad = de > 1e-14 + 1e-12 * np.abs(at_gpu)
        assert not bad.any(), (f"{label}: residual of the GPU state differs from the oracle's "
                               f"evaluation on {int(bad.sum())} rays (max |Δ| {de.max():.3e})")
        if cmp.any():
            assert rg[cmp].max() <= 2.0 * rr[cmp].max() + 1e-12, f"{label}: largest residual"
            assert np.median(rg[cmp]) <= 2.0 * np.median(rr[cmp]) + 1e-12, f"{label}: median residual"
        small = cmp & (rr <= 1e-12)
        d = np.abs(rg[cmp] - rr[cmp])
        rel = d / np.maximum(np.abs(rr[cmp]), 1e-300)
        out.update(resid_compared=int(cmp.sum()),
                   resid_eval_max_abs_diff=float(de.max()) if de.size else 0.0,
                   resid_trace_max_abs_diff=float(d.max()) if d.size else 0.0,
                   resid_trace_median_rel_diff=float(np.median(rel)) if rel.size else 0.0,
                   resid_oracle_at_most_1e12=int(small.sum()),
                   resid_gpu_at_most_1e12_there=int((rg[small] <= 1e-12).sum()))
    print(f"{label}: {out}", flush=True)
    assert out["max_angle"] <= ANGLE_TOL, f"{label}: escape direction {out['max_angle']:.3e} rad"
    return out
