// geon_fp32.hpp — FP32 side of the precision split.
//
// Bulk rays integrate in FP32 with the same RK4 tableau (integrator.rs:16-40)
// but with closed-form Christoffel symbols for every metric: a central
// difference with h = 1e-5 is pure rounding noise in FP32, so the generic
// FD ∂g of the reference is replaced by the analytic ∂g it approximates
// (its O(h²) truncation, ~1e-10 relative, is far below FP32 resolution).
// Every fused multiply-add is written out (fmaf) and nothing else is
// contracted: the compiler is free to fuse differently in each instantiation
// and unrolled copy of a step, which would make a ray's rounding depend on
// the kernel and slot it runs in; written out, it is fixed by the source.
// After each step p^t is re-projected onto the null cone g(p,p) = 0
// (future-directed root), which keeps FP32 drift out of the
// energy-carrying component — except in the Alcubierre wall, where the
// reference's own step leaves the cone by its truncation error and the
// re-projection would change the map that the FP64 trace follows (see
// AlcubierreF::reproject). Where Γ vanishes at every stage (Minkowski,
// Alcubierre beyond its wall) the step is the straight line y + h·p.
// Rays whose outcome is sensitive to rounding
// are handed to the FP64 path: a priori those whose impact parameter lies
// within a band of a photon orbit; a posteriori those that end diverged or
// flung far past the escape sphere, Schwarzschild captures (the reference
// decides them at a near-singular stage), those that dipped inside 3.2M
// (0.1 b0 at a wormhole throat), and those that passed near the coordinate
// pole: |sin θ| < 0.02 anywhere, or |sin θ| < 0.1 within 6M of a black hole
// (everywhere for the wormhole), where FP32 errors are amplified.
#pragma once
#include "geon_device.hpp"

namespace geon {

// float metric constants, filled per ray from the double constants
struct FParams {
    float M, cap_r, esc_r, esc2, esc_l;
    float R, inv_s, v, ship_x, inv_denom_s, denom_inv;
    float b0sq;
    float flat_rs;      // Alcubierre: beyond this r_s both FP32 tanh are exactly 1
    float near_orbit;   // a ray dipping below this radius (|l|) is re-traced in FP64
    float pole_r;       // within this radius (|l|) the wide pole band applies
};

// A ray whose |sin θ| drops below GEON_POLE_SIN passes the coordinate pole,
// where the spherical chart is ill-conditioned: the reference's fixed step
// turns sharply there and FP32 rounding changes the turn. Close to the hole
// the angular rates are larger and the band widens to GEON_POLE_SIN_NEAR
// (Schwarzschild 4096², r = 30M: escape directions off by radians at
// |sin θ| ≈ 0.02, by 8e-3 rad at 0.05 and 4M, below 5e-4 rad beyond 0.1
// or 5M).
#define GEON_POLE_SIN 0.02f
#define GEON_POLE_SIN_NEAR 0.1f

GEON_D FParams fparams(const geon_dmetric& P) {
    FParams f;
    f.M = (float)P.M;
    f.cap_r = (float)P.cap_r;
    f.esc_r = (float)P.esc_r;
    f.esc2 = (float)P.esc2;
    f.esc_l = (float)P.esc_l;
    f.R = (float)P.R;
    f.inv_s = (float)P.inv_s;
    f.v = (float)P.v;
    f.ship_x = (float)P.ship_x;
    f.denom_inv = (float)(1.0 / P.denom);              // 1 / (2 tanh(R/s))
    f.inv_denom_s = (float)(P.inv_s / P.denom);        // 1 / (2 s tanh(R/s))
    f.b0sq = (float)(P.b0 * P.b0);
    // 1 − 2/(e^{2u} + 1) rounds to 1.0f for u ≥ 9.1; 10 wall widths of margin
    f.flat_rs = (float)(P.R + 10.0 * fabs(P.s));
    f.near_orbit = P.kind == 1 ? (float)(3.2 * P.M) : (P.kind == 3 ? (float)(0.1 * P.b0) : 0.0f);
    // the wide band near the hole or the throat only (Morris–Thorne: 62.7 M
    // rays of 89 frames against FP64, the same worst escape direction, 5.0e-4
    // rad, with it at every radius or within 6 b0; without it, 3 rad)
    f.pole_r = P.kind == 1 ? (float)(6.0 * P.M) : (P.kind == 3 ? (float)(6.0 * P.b0) : 1e30f);
    return f;
}

// the path came near the pole (min |sin θ|) with the closest radius min_r
GEON_D bool pole_pass(const FParams& P, float min_s, float min_r) {
    return min_s < GEON_POLE_SIN || (min_s < GEON_POLE_SIN_NEAR && min_r < P.pole_r);
}

GEON_D float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
GEON_D float fsqrt(float x) { return __builtin_sqrtf(x); }
// a² + b² + c² as fmaf(a, a, fmaf(b, b, c·c))
GEON_D float norm3sq(float a, float b, float c) {
    return __builtin_fmaf(a, a, __builtin_fmaf(b, b, c * c));
}

// sin x and cos x in FP32 for |x| < 2^13 (else the library sincosf): the
// Cody–Waite reduction x = k·π/2 + r with π/2 in three floats (each product
// exact inside an FMA), Taylor polynomials of sin to r^9 and of cos to r^10
// on |r| ≤ π/4 (truncation below 1e-9), and the quadrant k mod 4.
#ifndef GEON_OPCOUNT   // the op-counting build tallies this as one sine and one cosine
GEON_D void sincosf_cw(float x, float* s, float* c) {
    if (!(fabsf(x) < 8192.0f)) {
        sincosf(x, s, c);
        return;
    }
    const float k = rintf(x * 0x1.45f306p-1f);                  // x·(2/π)
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = __builtin_fmaf(0x1.71de3ap-19f, r2, -0x1.a01a02p-13f);   //  1/9!, −1/7!
    ps = __builtin_fmaf(ps, r2, 0x1.111112p-7f);                         //  1/5!
    ps = __builtin_fmaf(ps, r2, -0x1.555556p-3f);                        // −1/3!
    const float sr = __builtin_fmaf(r * r2, ps, r);
    float pc = __builtin_fmaf(-0x1.27e4fcp-22f, r2, 0x1.a01a02p-16f);  // −1/10!, 1/8!
    pc = __builtin_fmaf(pc, r2, -0x1.6c16c2p-10f);                       // −1/6!
    pc = __builtin_fmaf(pc, r2, 0x1.555556p-5f);                         //  1/4!
    const float cr = __builtin_fmaf(r2 * r2, pc, __builtin_fmaf(-0.5f, r2, 1.0f));
    const int q = (int)k & 3;
    const float sv = (q & 1) ? cr : sr;
    const float cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}
#endif

// tanh via one exp: tanh(u) = 1 − 2/(e^{2u} + 1), exact saturation at ±1
GEON_D float ftanh(float u) {
    const float e = __expf(2.0f * u);
    return __builtin_fmaf(-2.0f, frcp(e + 1.0f), 1.0f);
}

// ---- Minkowski --------------------------------------------------------------
struct MinkowskiF {
    static constexpr int flat_test = 1;
    GEON_D static void accel(const FParams&, const float*, float* dp, float, float) {
        dp[0] = dp[1] = dp[2] = dp[3] = 0.0f;
    }
    GEON_D static void reproject(const FParams&, float* y) {
        y[4] = fsqrt(norm3sq(y[5], y[6], y[7]));
    }
    GEON_D static int terminate(const FParams& P, const float* x) {
        const float r2 = norm3sq(x[1], x[2], x[3]);
        return (r2 > P.esc2) ? ESCAPED : RUNNING;
    }
    GEON_D static bool suspicious(const FParams&, int cls, const float*, float, float) {
        return cls == DIVERGED;
    }
    GEON_D static float radius(const float*) { return 1e30f; }
    static constexpr bool uses_theta = false;
};

// ---- Schwarzschild (analytic Γ, schwarzschild.rs:61-107) --------------------
struct SchwarzschildF {
    static constexpr bool uses_theta = true;
    static constexpr int flat_test = 0;
    // 1/r, 1/(r − 2M) and cot θ from one reciprocal of r·(r − 2M)·sin θ
    // (reciprocals issue at a quarter of the VALU rate). No pole guard: a
    // stage this close to the pole has |sin θ| below the pole band, so the
    // ray goes to the FP64 path whatever this step gives (at sin θ = 0 it
    // gives NaN, which ends the ray as diverged).
    GEON_D static void accel(const FParams& P, const float* y, float* dp, float s, float c) {
        const float r = y[1];
        const float pt = y[4], pr = y[5], pth = y[6], pph = y[7];
        const float rf = __builtin_fmaf(-2.0f, P.M, r);   // r − 2M = r·f
        const float rr = r * rf;
        const float q = frcp(rr * s);
        const float inv_r = (rf * s) * q;
        const float inv_rr = s * q;                  // 1/(r(r − 2M))
        const float f = rf * inv_r;
        const float a = (P.M * inv_r) * inv_r;       // M/r²
        const float gtr = P.M * inv_rr;              // Γ^t_tr = M/(r²f)
        const float grtt = a * f;                    // Γ^r_tt
        const float cot = (c * rr) * q;
        const float pph2 = pph * pph;
        dp[0] = ((-2.0f * gtr) * pt) * pr;
        // −(Γ^r_tt pt² − M/(r²f)·pr²) + (r − 2M)(pθ² + sin²θ pφ²)
        dp[1] = __builtin_fmaf(rf, __builtin_fmaf(pth, pth, (s * s) * pph2),
                               __builtin_fmaf(-grtt * pt, pt, (gtr * pr) * pr));
        dp[2] = __builtin_fmaf((-2.0f * inv_r) * pr, pth, (s * c) * pph2);
        dp[3] = (-2.0f * pph) * __builtin_fmaf(inv_r, pr, cot * pth);
    }
    GEON_D static void reproject(const FParams& P, float* y, float s) {
        const float r = y[1];
        const float rf = __builtin_fmaf(-2.0f, P.M, r);
        const float inv_f = (r * r) * frcp(r * rf);  // 1/f = r/(r − 2M)
        const float ang = __builtin_fmaf(y[6], y[6], (s * s) * (y[7] * y[7]));
        const float sp = __builtin_fmaf(y[5] * y[5], inv_f, (r * r) * ang);
        if (rf > 0.0f && sp >= 0.0f) y[4] = fsqrt(sp * inv_f);
    }
    GEON_D static int terminate(const FParams& P, const float* x) {
        const float r = x[1];
        if (r >= P.cap_r && r <= P.esc_r) return RUNNING;   // NaN fails both
        if (!__builtin_isfinite(r)) return DIVERGED;
        return (r < P.cap_r) ? CAPTURED : ESCAPED;
    }
    // Capture is decided by FP64 too: whether the reference's fixed step
    // crosses r = 2.01M smoothly or has a stage land on r ≈ 2M (and fling the
    // ray out) depends on FP64 rounding at that near-singular stage.
    GEON_D static bool suspicious(const FParams& P, int cls, const float* y, float min_s, float min_r) {
        if (cls == DIVERGED || cls == CAPTURED) return true;
        if (cls == ESCAPED && y[1] > 2.0f * P.esc_r) return true;
        return pole_pass(P, min_s, min_r) || min_r < P.near_orbit;
    }
    GEON_D static float radius(const float* y) { return y[1]; }
};

// ---- Morris–Thorne (analytic ∂g of morris_thorne.rs:22-36) ------------------
struct MorrisThorneF {
    static constexpr bool uses_theta = true;
    static constexpr int flat_test = 0;
    // no pole guard, as SchwarzschildF::accel
    GEON_D static void accel(const FParams& P, const float* y, float* dp, float s, float c) {
        const float l = y[1];
        const float pl = y[5], pth = y[6], pph = y[7];
        const float r2 = __builtin_fmaf(l, l, P.b0sq);
        const float q = frcp(r2 * s);                // 1/r² and cot θ from one reciprocal
        const float inv_r2 = s * q;
        const float lr = l * inv_r2;                 // Γ^θ_lθ = Γ^φ_lφ
        const float cot = (c * r2) * q;
        const float pph2 = pph * pph;
        dp[0] = 0.0f;
        // −(Γ^l_θθ pθ² + Γ^l_φφ pφ²)
        dp[1] = l * __builtin_fmaf(pth, pth, (s * s) * pph2);
        dp[2] = __builtin_fmaf((-2.0f * lr) * pl, pth, (s * c) * pph2);
        dp[3] = (-2.0f * pph) * __builtin_fmaf(lr, pl, cot * pth);
    }
    GEON_D static void reproject(const FParams& P, float* y, float s) {
        const float r2 = __builtin_fmaf(y[1], y[1], P.b0sq);
        const float ang = __builtin_fmaf(y[6], y[6], (s * s) * (y[7] * y[7]));
        y[4] = fsqrt(__builtin_fmaf(y[5], y[5], r2 * ang));
    }
    GEON_D static int terminate(const FParams& P, const float* x) {
        const float l = x[1];
        if (fabsf(l) <= P.esc_l) return RUNNING;   // NaN fails it
        return __builtin_isfinite(l) ? ESCAPED : DIVERGED;
    }
    GEON_D static bool suspicious(const FParams& P, int cls, const float* y, float min_s, float min_r) {
        if (cls == DIVERGED) return true;
        if (cls == ESCAPED && fabsf(y[1]) > 2.0f * P.esc_l) return true;
        return pole_pass(P, min_s, min_r) || min_r < P.near_orbit;
    }
    GEON_D static float radius(const float* y) { return fabsf(y[1]); }
};

// ---- Alcubierre (analytic ∂g of alcubierre.rs:33-63) ------------------------
// With n_i = ∂_i r_s, A_i = ∂_i g_tt = 2v²ff'n_i, B_i = ∂_i g_tx = −vf'n_i and
// g^00 = −1, g^01 = −vf, g^11 = 1 − v²f²:
//   Q0 = 2p⁰(A·p) + 2p¹(B·p), Q1 = −A_1 (p⁰)² + 2p⁰(B·p − B_1 p¹),
//   Qk = −A_k (p⁰)² − 2B_k p⁰p¹ (k = y, z);
//   dp⁰ = ½(Q0 + vf Q1), dp¹ = ½(vf Q0 − (1 − v²f²) Q1), dp^k = −½ Qk.
struct AlcubierreF {
    static constexpr bool uses_theta = false;
    static constexpr int flat_test = 2;
    // The number of consecutive steps from y, this one included, whose
    // stages all stay beyond flat_rs, where Γ is zero (0: not this one).
    // The momentum is constant there and each step moves the position by
    // h·p, so K steps stay within K·|h|·|p| of y. The allowance 1e-3·(1 + r)
    // on the distance covers the FP32 rounding of K ≤ 1024 positions
    // (≤ 1024·2^-24·r).
    GEON_D static uint32_t flat_steps(const FParams& P, const float* y, float h) {
        const float dx = y[1] - P.ship_x, dy = y[2], dz = y[3];
        const float r = fsqrt(norm3sq(dx, dy, dz));
        const float d = r - P.flat_rs;
        const float p2 = norm3sq(y[5], y[6], y[7]);
        if (!(d > 0.0f && d * d > (h * h) * p2)) return 0;   // one step
        const float k = __builtin_fmaf(-1e-3f, 1.0f + r, d) * frcp((fabsf(h) * fsqrt(p2)) * 1.001f);
        return k >= 1024.0f ? 1024u : (k >= 1.0f ? (uint32_t)k : 1u);
    }
    GEON_D static void shape(const FParams& P, const float* y, float& f, float& fp, float* n) {
        const float dx = y[1] - P.ship_x, dy = y[2], dz = y[3];
        const float rs = fsqrt(norm3sq(dx, dy, dz));
        const float inv_rs = frcp(fmaxf(rs, 1e-30f));
        n[0] = dx * inv_rs; n[1] = dy * inv_rs; n[2] = dz * inv_rs;
        const float tp = ftanh((rs + P.R) * P.inv_s);
        const float tm = ftanh((rs - P.R) * P.inv_s);
        f = (tp - tm) * P.denom_inv;
        fp = (__builtin_fmaf(-tp, tp, 1.0f) - __builtin_fmaf(-tm, tm, 1.0f)) * P.inv_denom_s;
    }
    GEON_D static void accel(const FParams& P, const float* y, float* dp, float, float) {
        float f, fp, n[3];
        shape(P, y, f, fp, n);
        const float v = P.v;
        const float p0 = y[4], p1 = y[5], p2 = y[6], p3 = y[7];
        const float np = __builtin_fmaf(n[0], p1, __builtin_fmaf(n[1], p2, n[2] * p3));
        const float Ac = (((2.0f * v) * v) * f) * fp;   // A_i = Ac n_i
        const float Bc = -v * fp;                       // B_i = Bc n_i
        const float Q0 = __builtin_fmaf((2.0f * p0) * Ac, np, ((2.0f * p1) * Bc) * np);
        const float Q1 = __builtin_fmaf((-Ac * n[0]) * p0, p0,
                                        ((2.0f * p0) * Bc) * __builtin_fmaf(-n[0], p1, np));
        const float vf = v * f;
        const float b2 = (2.0f * Bc) * p0 * p1;
        dp[0] = 0.5f * __builtin_fmaf(vf, Q1, Q0);
        dp[1] = 0.5f * __builtin_fmaf(vf, Q0, -(__builtin_fmaf(-vf, vf, 1.0f) * Q1));
        dp[2] = 0.5f * (n[1] * __builtin_fmaf(Ac * p0, p0, b2));
        dp[3] = 0.5f * (n[2] * __builtin_fmaf(Ac * p0, p0, b2));
    }
    // No null re-projection here. Inside the wall the reference's RK4 step
    // drifts off the null cone by its own truncation error, and the FP64
    // trace it is compared with keeps that drift; moving p^t back onto the
    // cone every step changes the discrete map itself, and the change grows
    // across the wall. On the CPU, an FP32 emulation of this path against
    // the oracle (v = 0.9, 4000 rays through the wall region): with the
    // re-projection 1496 escape directions off by more than 1e-3 rad (max
    // 2.4 rad), without it 10 (max 4.0e-3), all after more than 256 wall
    // steps. Outside the wall the momentum is constant and there is nothing
    // to re-project.
    GEON_D static void reproject(const FParams&, float*) {}
    GEON_D static int terminate(const FParams& P, const float* x) {
        const float r2 = norm3sq(x[1], x[2], x[3]);
        if (!__builtin_isfinite(r2)) return DIVERGED;
        return (r2 > P.esc2) ? ESCAPED : RUNNING;
    }
    // The second tracker counts the steps that evaluated Γ (wall steps).
    // Rays that stay in the wall region long — those that cross the bubble's
    // interior, or are caught near its walls — have escape directions that
    // amplify FP32 rounding, and trapped ones change class between escaped
    // and max-steps. A ray that took more than kWallSteps such steps, or is
    // still inside the escape sphere at max_steps, is re-traced in FP64.
    // Measured on the 94 Alcubierre frames of the 256-frame ragged batch
    // (v ∈ [0.1, 0.9], R = 2, s = 1/8, 53.8 M pixels), FP32-mixed against
    // FP64, no class mismatch at any threshold; escape directions above
    // 1e-3 rad: 564 pixels (max 1.7e-2 rad) at 192, 872 (max 1.7e-2) at 256
    // with 0.7 % of the rays re-traced, 2188 (max 5.5e-2) at 384. With the
    // re-projection the same frames needed 128 (10.7 % re-traced; 1.2 M
    // pixels above 1e-3 rad, max 0.52 rad, at 256).
    static constexpr bool counts_wall = true;
#ifndef GEON_FP32_WALL_STEPS
#define GEON_FP32_WALL_STEPS 256
#endif
    enum { kWallSteps = GEON_FP32_WALL_STEPS };
    GEON_D static bool suspicious(const FParams& P, int cls, const float* y, float, float wall) {
        if (cls == DIVERGED || cls == MAXSTEPS) return true;
        if (cls == ESCAPED)
            return norm3sq(y[1], y[2], y[3]) > 4.0f * P.esc2 || wall > (float)kWallSteps;
        return false;
    }
    GEON_D static float radius(const float*) { return 1e30f; }
};

// (sin, cos) of θ + δ from (sin θ, cos θ): rotation by δ with sin δ, cos δ
// from their Taylor series (|δ| ≤ 0.25: truncation below 2e-8), falling back
// to sincosf for a large stage increment.
GEON_D void rotate_sc(float s, float c, float theta_new, float d, float& so, float& co) {
    if (fabsf(d) > 0.25f) {
        sincosf(theta_new, &so, &co);
        return;
    }
    const float d2 = d * d;
    const float sd = d * __builtin_fmaf(-d2, __builtin_fmaf(-d2, 1.0f / 120.0f, 1.0f / 6.0f), 1.0f);
    const float cd = __builtin_fmaf(
        -d2, __builtin_fmaf(-d2, __builtin_fmaf(-d2, 1.0f / 720.0f, 1.0f / 24.0f), 0.5f), 1.0f);
    so = __builtin_fmaf(s, cd, c * sd);
    co = __builtin_fmaf(c, cd, -(s * sd));
}

// FP32 RK4 step with the reference tableau. (s, c) carry sin/cos of y[2] in
// and of out[2] out: one accurate sincosf per step, the stage angles θ + δ
// come from rotations by the small stage increments δ. p^t is re-projected
// at the end of the step. Returns the smallest |sin θ| of the new state and
// the step's stages (1 for Cartesian charts).
// A run of straight-line (Γ = 0) FP32 steps: the steps still known to be
// flat, the steps taken since the run began and the position it began at.
// Step j of a run puts the ray at a + (j·h)·p, one rounding from the exact
// line: adding h·p step by step would round the position every step, and
// far from the origin (an FP32 ulp is 4e-6 at |x| = 40) that drift alone
// moved escape directions by up to 2e-3 rad after a few hundred steps.
struct FlatRunF {
    uint32_t left = 0;
    uint32_t j = 0;
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool nf = false;   // the flatness test of this step already ran and failed
};

template <class MF>
GEON_D float rk4_step_f(const FParams& P, const float* y, float h, float* out, float& s, float& c,
                        FlatRunF& fr) {
    if constexpr (MF::flat_test != 0) {
        // Γ = 0 at every stage: the RK4 step is the straight line y + h·p
        // (momenta, p^t included, unchanged)
        bool cont = fr.j > 0, flat = true;
        if constexpr (MF::flat_test == 2) {
            if (fr.left > 0) {
                --fr.left;
            } else if (fr.nf) {
                fr.nf = false;
                flat = false;
            } else {
                const uint32_t k = MF::flat_steps(P, y, h);
                fr.left = k > 0 ? k - 1 : 0;
                flat = k > 0;
                cont = false;
            }
        }
        if (flat) {
            if (!cont) {
#pragma unroll
                for (int i = 0; i < 4; ++i) fr.a[i] = y[i];
                fr.j = 0;
            }
            ++fr.j;
            const float jh = (float)fr.j * h;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                out[i] = __builtin_fmaf(jh, y[4 + i], fr.a[i]);
                out[4 + i] = y[4 + i];
            }
            return 1.0f;
        }
        fr.j = 0;
    }
    float acc[8], yt[8], dp[4];
    float ss = s, cs = c;
    const float hh = 0.5f * h;
    MF::accel(P, y, dp, s, c);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[i] = y[4 + i];
        acc[4 + i] = dp[i];
        yt[i] = __builtin_fmaf(hh, y[4 + i], y[i]);
        yt[4 + i] = __builtin_fmaf(hh, dp[i], y[4 + i]);
    }
    float smin = 1.0f;   // the smallest stage |sin θ| (the pole band looks at stages too)
    if constexpr (MF::uses_theta) {
        rotate_sc(s, c, yt[2], hh * y[6], ss, cs);
        smin = fabsf(ss);
    }
    MF::accel(P, yt, dp, ss, cs);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float kx = yt[4 + i];
        acc[i] = __builtin_fmaf(2.0f, kx, acc[i]);
        acc[4 + i] = __builtin_fmaf(2.0f, dp[i], acc[4 + i]);
        yt[i] = __builtin_fmaf(hh, kx, y[i]);
        yt[4 + i] = __builtin_fmaf(hh, dp[i], y[4 + i]);
    }
    if constexpr (MF::uses_theta) {
        rotate_sc(s, c, yt[2], yt[2] - y[2], ss, cs);
        smin = fminf(smin, fabsf(ss));
    }
    MF::accel(P, yt, dp, ss, cs);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float kx = yt[4 + i];
        acc[i] = __builtin_fmaf(2.0f, kx, acc[i]);
        acc[4 + i] = __builtin_fmaf(2.0f, dp[i], acc[4 + i]);
        yt[i] = __builtin_fmaf(h, kx, y[i]);
        yt[4 + i] = __builtin_fmaf(h, dp[i], y[4 + i]);
    }
    if constexpr (MF::uses_theta) {
        rotate_sc(s, c, yt[2], yt[2] - y[2], ss, cs);
        smin = fminf(smin, fabsf(ss));
    }
    MF::accel(P, yt, dp, ss, cs);
    const float sixth = h * (1.0f / 6.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[i] += yt[4 + i];
        acc[4 + i] += dp[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = __builtin_fmaf(sixth, acc[i], y[i]);
    if constexpr (MF::uses_theta) {
        sincosf_cw(out[2], &s, &c);
        MF::reproject(P, out, s);
        return fminf(fabsf(s), smin);
    } else {
        MF::reproject(P, out);
        return 2.0f;   // a step that evaluated Γ (flat steps return 1)
    }
}

template <class MF>
GEON_D float rk4_step_f(const FParams& P, const float* y, float h, float* out, float& s, float& c) {
    FlatRunF fr;
    return rk4_step_f<MF>(P, y, h, out, s, c, fr);
}

// A span of flat steps for a state whose escape test just passed, at most
// `cap` steps: the anchored step above puts step j of a run at a + (j·h)·p
// whatever the steps before it did, so m steps are one evaluation at
// j + m — where every position up to it provably stays inside the escape
// sphere, so that no escape test or finiteness guard in between could have
// fired. Bound: |fl(j·h)| ≤ j|h|(1 + 2^-24) and the FMA rounds once, so with
// j|h||p_s|(1 + 1e-5) ≤ R(1 − 1e-5) − |a|(1 + 1e-5), R = √esc², every
// position is within R(1 − 9e-6) and its r² below esc². Momenta stay as
// they are. Returns the steps taken (0: take an ordinary step, which then
// knows the flatness test's outcome).
template <class MF>
GEON_D uint32_t flat_span_f(const FParams& P, float* y, float h, FlatRunF& fr, uint32_t cap) {
    uint32_t run = 1u << 22;
    if constexpr (MF::flat_test == 2) {
        run = fr.left;
        if (run == 0) {
            run = MF::flat_steps(P, y, h);
            if (run == 0) {
                fr.nf = true;
                return 0;
            }
            fr.j = 0;   // a new run: anchored at this state
        }
        fr.left = run;  // an ordinary step takes a flat step while > 0
    }
    if (fr.j == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) fr.a[i] = y[i];
    }
    const float ra = fsqrt(norm3sq(fr.a[1], fr.a[2], fr.a[3]));
    const float ps = fsqrt(norm3sq(y[5], y[6], y[7]));
    const float avail = __builtin_fmaf(-ra, 1.0f + 1e-5f, fsqrt(P.esc2) * (1.0f - 1e-5f));
    // j stays below 2^22, where (float)j is exact
    const float jlim = fminf(avail / ((fabsf(h) * ps) * (1.0f + 1e-5f)), 4194304.0f);
    if (!(avail > 0.0f && avail < 1e30f) || !(jlim >= (float)fr.j + 1.0f) ||
        !(__builtin_fmaf(8388608.0f, fabsf(h * y[4]), fabsf(fr.a[0])) < 1e30f))
        return 0;
    uint32_t m = (uint32_t)jlim - fr.j;
    if (m > run) m = run;
    if (m > cap) m = cap;
    fr.j += m;
    if constexpr (MF::flat_test == 2) fr.left = run - m;
    const float jh = (float)fr.j * h;
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = __builtin_fmaf(jh, y[4 + i], fr.a[i]);
    return m;
}

GEON_D bool finite8f(const float* y) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) ok = ok && __builtin_isfinite(y[i]);
    return ok;
}

// map a double metric policy to its FP32 counterpart
template <class MP> struct FP32Of;
template <> struct FP32Of<Minkowski> { using type = MinkowskiF; };
template <> struct FP32Of<Schwarzschild> { using type = SchwarzschildF; };
template <> struct FP32Of<MorrisThorne> { using type = MorrisThorneF; };
template <> struct FP32Of<Alcubierre> { using type = AlcubierreF; };

}  // namespace geon
