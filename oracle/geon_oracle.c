/*
 * geon_oracle.c — CPU restatement of Arth-Singh/geon's geodesic path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline leg of the bench. The product (libgeon_hip.so) never links it.
 *
 * Rules followed so that results match the Rust reference bit for bit:
 *   - every expression keeps the reference's association order
 *     (Rust evaluates a*b*c as (a*b)*c, as C does);
 *   - built with -ffp-contract=off on baseline x86-64 (SSE2): no FMA,
 *     matching rustc's release build for the baseline target
 *     (Cargo.toml:66-70);
 *   - transcendental calls go to glibc libm, which is what Rust's f64
 *     methods call on Linux.
 * The one dependency whose rounding cannot be pinned here is nalgebra
 * 0.33.3's 4×4 try_inverse (Cargo.lock:353-354, call site types.rs:34-46);
 * or_inv4 restates the cofactor (adjugate / determinant) expansion it uses.
 *
 * Parity pins (tests/): Schwarzschild curvature noise values and every
 * other line of docs/energy_conditions.txt, the README's validate_shadow
 * numbers, sympy closed forms for g, Γ and G_μν, Minkowski straight lines,
 * RK4 4th-order convergence, radial-infall step counts and tetrad
 * orthonormality.
 */
#define _GNU_SOURCE
#include "geon_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_PI  3.141592653589793      /* std::f64::consts::PI  */
#define OR_TAU 6.283185307179586      /* std::f64::consts::TAU */
#define OR_FRAC_PI_2 1.5707963267948966

/* ---------------------------------------------------------------- L0 */

/* types.rs:22-31 — Σ_i Σ_j g[i][j]·a[i]·b[j], i-major, left to right. */
double or_dot4(const double g[16], const double a[4], const double b[4]) {
    double s = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            s += g[i * 4 + j] * a[i] * b[j];
    return s;
}

/*
 * types.rs:34-46 — inv4 via nalgebra Matrix4::try_inverse. nalgebra 0.33.3
 * inverts 4×4 matrices with the closed-form cofactor expansion (the
 * well-known MESA gluInvertMatrix scheme): sixteen signed sums of triple
 * products over the column-major element array, det from the first
 * column's cofactors, each cofactor divided by det. Restated here from
 * that published scheme; bitwise agreement with nalgebra is unpinned.
 * Returns GEON_E_SINGULAR when det == 0 (the reference panics there).
 */
int or_inv4(const double a[16], double out[16]) {
    /* column-major element array: m[c*4+r] = A[r][c] */
    double m[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) m[c * 4 + r] = a[r * 4 + c];
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15]
           + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15]
           - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15]
           + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14]
            - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15]
           - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15]
           + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15]
           - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14]
            + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15]
           + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15]
           - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15]
            + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14]
            - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11]
           - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11]
           + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11]
            - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10]
            + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return GEON_E_SINGULAR;
    /* back to row-major: out[r][c] = inv[c*4+r] / det */
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = inv[c * 4 + r] / det;
    return GEON_OK;
}

/* ---------------------------------------------------------------- L1 */

/* Reference constructors: schwarzschild.rs:23-29, alcubierre.rs:21-30,
 * morris_thorne.rs:14-19, minkowski.rs:8; fd_step metric/mod.rs:73-75. */
void or_metric_default(int32_t kind, geon_metric_t* out) {
    memset(out, 0, sizeof(*out));
    out->kind = kind;
    out->horizon_pad = 1.001;
    out->escape_radius = 1000.0;
    out->fd_step = 1e-5;
    switch (kind) {
    case GEON_METRIC_SCHWARZSCHILD: out->p[0] = 1.0; break;
    case GEON_METRIC_ALCUBIERRE:
        out->p[0] = 2.0; out->p[1] = 0.5; out->p[2] = 2.0; out->p[3] = 0.0; break;
    case GEON_METRIC_MORRIS_THORNE: out->p[0] = 1.0; break;
    default: break;
    }
}

uint64_t or_sky_hash1(uint64_t x);

/* A planted perturbation for conditioning tests only (0 = the reference):
 * with k = ±1 every tanh of the shape function below is moved by one ulp
 * towards ±∞; with |k| ≥ 2 each evaluation is moved by one ulp up or down,
 * the direction drawn from a hash of its argument and k. A test measures how
 * far the reference's own results move when its tanh changes in the last
 * bit. */
static int or_tanh_ulp_shift = 0;
void or_set_tanh_ulp_shift(int k) { or_tanh_ulp_shift = k; }
static double shape_tanh(double v) {
    const double t = tanh(v);
    const int k = or_tanh_ulp_shift;
    if (k == 0) return t;
    int up = k > 0;
    if (k >= 2 || k <= -2) {
        uint64_t bits;
        memcpy(&bits, &v, sizeof bits);
        up = (int)(or_sky_hash1(bits ^ ((uint64_t)(int64_t)k * 0x9E3779B97F4A7C15ull)) & 1u);
    }
    return nextafter(t, up ? INFINITY : -INFINITY);
}

/* alcubierre.rs:33-44 — smooth top-hat f(r_s). */
static double alc_shape(const geon_metric_t* m, const double x[4]) {
    double dx = x[1] - m->p[3];
    double dy = x[2];
    double dz = x[3];
    double r_s = sqrt(dx * dx + dy * dy + dz * dz);
    double s = m->p[1];
    double r = m->p[0];
    double t_plus = shape_tanh((r_s + r) / s);
    double t_minus = shape_tanh((r_s - r) / s);
    double denom = 2.0 * tanh(r / s);
    return (t_plus - t_minus) / denom;
}

void or_g(const geon_metric_t* m, const double x[4], double g[16]) {
    memset(g, 0, 16 * sizeof(double));
    switch (m->kind) {
    case GEON_METRIC_MINKOWSKI: /* minkowski.rs:20-28 */
        g[0] = -1.0; g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
        break;
    case GEON_METRIC_SCHWARZSCHILD: { /* schwarzschild.rs:33-45 */
        double r = x[1], theta = x[2], mass = m->p[0];
        double f = 1.0 - 2.0 * mass / r;
        double sin_t = sin(theta);
        g[0] = -f;
        g[5] = 1.0 / f;
        g[10] = r * r;
        g[15] = r * r * sin_t * sin_t;
        break;
    }
    case GEON_METRIC_ALCUBIERRE: { /* alcubierre.rs:48-63 */
        double v = m->p[2];
        double f = alc_shape(m, x);
        g[0] = -1.0 + v * v * f * f;
        g[1] = -v * f;
        g[4] = -v * f;
        g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
        break;
    }
    case GEON_METRIC_MORRIS_THORNE: { /* morris_thorne.rs:22-36 */
        double l = x[1], theta = x[2], b0 = m->p[0];
        double r2 = b0 * b0 + l * l;
        double sin_t = sin(theta);
        g[0] = -1.0; g[5] = 1.0; g[10] = r2;
        g[15] = r2 * sin_t * sin_t;
        break;
    }
    default: break;
    }
}

/* metric/mod.rs:25-43 — central FD of g, out[α][μ][ν] = (g+ − g−)/(2h).
 * Minkowski overrides to zero (minkowski.rs:20-22). */
void or_dg(const geon_metric_t* m, const double x[4], double dg[64]) {
    memset(dg, 0, 64 * sizeof(double));
    if (m->kind == GEON_METRIC_MINKOWSKI) return;
    double h = m->fd_step;
    for (int alpha = 0; alpha < 4; ++alpha) {
        double xp[4], xm[4], gp[16], gm[16];
        memcpy(xp, x, sizeof xp);
        memcpy(xm, x, sizeof xm);
        xp[alpha] += h;
        xm[alpha] -= h;
        or_g(m, xp, gp);
        or_g(m, xm, gm);
        double two_h = 2.0 * h;
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu)
                dg[alpha * 16 + mu * 4 + nu] = (gp[mu * 4 + nu] - gm[mu * 4 + nu]) / two_h;
    }
}

/* g_inv: analytic for Schwarzschild (schwarzschild.rs:47-59), η for
 * Minkowski (minkowski.rs:24-26), inv4(g) otherwise (metric/mod.rs:46-48). */
int or_g_inv(const geon_metric_t* m, const double x[4], double out[16]) {
    if (m->kind == GEON_METRIC_SCHWARZSCHILD) {
        memset(out, 0, 16 * sizeof(double));
        double r = x[1], theta = x[2], mass = m->p[0];
        double f = 1.0 - 2.0 * mass / r;
        double sin_t = sin(theta);
        out[0] = -1.0 / f;
        out[5] = f;
        out[10] = 1.0 / (r * r);
        out[15] = 1.0 / (r * r * sin_t * sin_t);
        return GEON_OK;
    }
    if (m->kind == GEON_METRIC_MINKOWSKI) {
        or_g(m, x, out);
        return GEON_OK;
    }
    double g[16];
    or_g(m, x, g);
    return or_inv4(g, out);
}

/* Schwarzschild analytic Γ (schwarzschild.rs:61-107). */
static void schw_christoffel(const geon_metric_t* mm, const double x[4], double c[64]) {
    memset(c, 0, 64 * sizeof(double));
    double r = x[1], theta = x[2], m = mm->p[0];
    double f = 1.0 - 2.0 * m / r;
    double sin_t = sin(theta);
    double cos_t = cos(theta);
#define C(a, b, d) c[(a) * 16 + (b) * 4 + (d)]
    double gtr = m / (r * r * f);
    C(0, 0, 1) = gtr;
    C(0, 1, 0) = gtr;
    C(1, 0, 0) = m * f / (r * r);
    C(1, 1, 1) = -m / (r * r * f);
    C(1, 2, 2) = -(r - 2.0 * m);
    C(1, 3, 3) = -(r - 2.0 * m) * sin_t * sin_t;
    double gtheta_rt = 1.0 / r;
    C(2, 1, 2) = gtheta_rt;
    C(2, 2, 1) = gtheta_rt;
    C(2, 3, 3) = -sin_t * cos_t;
    C(3, 1, 3) = 1.0 / r;
    C(3, 3, 1) = 1.0 / r;
    double cot = (fabs(sin_t) < 1e-8) ? 0.0 : cos_t / sin_t;
    C(3, 2, 3) = cot;
    C(3, 3, 2) = cot;
#undef C
}

/* metric/mod.rs:51-69 — Γ^μ_αβ = ½ Σ_σ g^μσ(∂_α g_σβ + ∂_β g_σα − ∂_σ g_αβ). */
int or_christoffel(const geon_metric_t* m, const double x[4], double gam[64]) {
    if (m->kind == GEON_METRIC_SCHWARZSCHILD) {
        schw_christoffel(m, x, gam);
        return GEON_OK;
    }
    if (m->kind == GEON_METRIC_MINKOWSKI) { /* minkowski.rs:28-30 */
        memset(gam, 0, 64 * sizeof(double));
        return GEON_OK;
    }
    double ginv[16], dg[64];
    int st = or_g_inv(m, x, ginv);
    if (st != GEON_OK) {
        for (int i = 0; i < 64; ++i) gam[i] = NAN;
        return st;
    }
    or_dg(m, x, dg);
    for (int mu = 0; mu < 4; ++mu)
        for (int alpha = 0; alpha < 4; ++alpha)
            for (int beta = 0; beta < 4; ++beta) {
                double s = 0.0;
                for (int sigma = 0; sigma < 4; ++sigma)
                    s += ginv[mu * 4 + sigma] *
                         (dg[alpha * 16 + sigma * 4 + beta] + dg[beta * 16 + sigma * 4 + alpha] -
                          dg[sigma * 16 + alpha * 4 + beta]);
                gam[mu * 16 + alpha * 4 + beta] = 0.5 * s;
            }
    return GEON_OK;
}

/* terminate: schwarzschild.rs:109-121, minkowski.rs:32-38,
 * alcubierre.rs:65-75, morris_thorne.rs:38-48. Returns -1 for "continue". */
int or_terminate(const geon_metric_t* m, const double x[4]) {
    switch (m->kind) {
    case GEON_METRIC_SCHWARZSCHILD: {
        double r = x[1];
        if (!isfinite(r)) return GEON_DIVERGED;
        if (r < m->horizon_pad * 2.0 * m->p[0]) return GEON_CAPTURED;
        if (r > m->escape_radius * m->p[0]) return GEON_ESCAPED;
        return -1;
    }
    case GEON_METRIC_MINKOWSKI: {
        double r2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
        double esc2 = m->escape_radius * m->escape_radius;
        return (r2 > esc2) ? GEON_ESCAPED : -1;
    }
    case GEON_METRIC_ALCUBIERRE: {
        double r2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
        if (!isfinite(r2)) return GEON_DIVERGED;
        double esc2 = m->escape_radius * m->escape_radius;
        return (r2 > esc2) ? GEON_ESCAPED : -1;
    }
    case GEON_METRIC_MORRIS_THORNE: {
        double l = x[1];
        if (!isfinite(l)) return GEON_DIVERGED;
        return (fabs(l) > m->escape_radius) ? GEON_ESCAPED : -1;
    }
    default: return GEON_DIVERGED;
    }
}

/* Rust f64::max: NaN-ignoring maximum, same as C fmax. */
static void spherical_direction(double r, const double x[4], const double p[4], double out[3],
                                double sign_x) {
    double theta = x[2], phi = x[3];
    double sin_t = sin(theta), cos_t = cos(theta);
    double sin_p = sin(phi), cos_p = cos(phi);
    double pr = p[1], pt = p[2], pp = p[3];
    double vx = sin_t * cos_p * pr + r * cos_t * cos_p * pt - r * sin_t * sin_p * pp;
    double vy = sin_t * sin_p * pr + r * cos_t * sin_p * pt + r * sin_t * cos_p * pp;
    double vz = cos_t * pr - r * sin_t * pt;
    double n = fmax(sqrt(vx * vx + vy * vy + vz * vz), 1e-30);
    out[0] = sign_x * vx / n;
    out[1] = vy / n;
    out[2] = vz / n;
}

/* asymptotic_direction: default spherical chart (metric/mod.rs:91-108,
 * used by Schwarzschild and, as in the reference, by Minkowski), Cartesian
 * (alcubierre.rs:81-88), two-sheeted (morris_thorne.rs:54-77). */
void or_asymptotic_direction(const geon_metric_t* m, const double x[4], const double p[4],
                             double out[3]) {
    if (m->kind == GEON_METRIC_ALCUBIERRE) {
        double vx = p[1], vy = p[2], vz = p[3];
        double n = fmax(sqrt(vx * vx + vy * vy + vz * vz), 1e-30);
        out[0] = vx / n; out[1] = vy / n; out[2] = vz / n;
        return;
    }
    if (m->kind == GEON_METRIC_MORRIS_THORNE) {
        double l = x[1];
        double r_eff = fmax(fabs(l), 1e-30);
        /* sign * vx / n: Rust evaluates (sign*vx)/n */
        double theta = x[2], phi = x[3];
        double sin_t = sin(theta), cos_t = cos(theta);
        double sin_p = sin(phi), cos_p = cos(phi);
        double pl = p[1], pt = p[2], pp = p[3];
        double vx = sin_t * cos_p * pl + r_eff * cos_t * cos_p * pt - r_eff * sin_t * sin_p * pp;
        double vy = sin_t * sin_p * pl + r_eff * cos_t * sin_p * pt + r_eff * sin_t * cos_p * pp;
        double vz = cos_t * pl - r_eff * sin_t * pt;
        double n = fmax(sqrt(vx * vx + vy * vy + vz * vz), 1e-30);
        double sign = (l < 0.0) ? -1.0 : 1.0;
        out[0] = sign * vx / n;
        out[1] = vy / n;
        out[2] = vz / n;
        return;
    }
    spherical_direction(x[1], x, p, out, 1.0);
}

/* ---------------------------------------------------------------- L2 */

/* geodesic.rs:38-59 — dx = p, dp^μ = −Σ_αβ Γ^μ_αβ p^α p^β (dense). */
void or_rhs(const geon_metric_t* m, const double y[8], double out[8]) {
    double x[4] = {y[0], y[1], y[2], y[3]};
    double p[4] = {y[4], y[5], y[6], y[7]};
    double gam[64];
    or_christoffel(m, x, gam);
    for (int mu = 0; mu < 4; ++mu) out[mu] = p[mu];
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) s += gam[mu * 16 + a * 4 + b] * p[a] * p[b];
        out[4 + mu] = -s;
    }
}

/* integrator.rs:16-40 — classic RK4. */
void or_rk4_step(const geon_metric_t* m, const double y[8], double h, double out[8]) {
    double k1[8], k2[8], k3[8], k4[8], yt[8];
    or_rhs(m, y, k1);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + 0.5 * h * k1[i];
    or_rhs(m, yt, k2);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + 0.5 * h * k2[i];
    or_rhs(m, yt, k3);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * k3[i];
    or_rhs(m, yt, k4);
    double sixth = h / 6.0;
    for (int i = 0; i < 8; ++i)
        out[i] = y[i] + sixth * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
}

/* Adaptive stepping: no reference counterpart (integrator.rs:5-7 names it
 * as the next step). Cash–Karp RK4(5) (Cash & Karp, ACM TOMS 16 (1990)
 * 201-222) on or_rhs, the error test max_i |err_i|/(atol + rtol·max(|y_i|,
 * |y_new_i|)) ≤ 1, next step h·0.9·e^(-1/4) in [h/10, 5h] (5h below
 * e = 1.0498e-3), a trial step with a non-finite component rejected with
 * h/10, |h| in [1e-8, 8]·|h0|; every sum left to right in the
 * tableau's order. No reference result exists for it; the tests pin it by
 * the tableau's orders (local errors O(h^6) and O(h^5) of the fifth- and
 * embedded fourth-order solutions), closed forms (straight lines, conserved
 * energy and angular momentum) and fine fixed-step traces of or_trace. */
void or_ck_step(const geon_metric_t* m, const double y[8], double h, double yn[8], double err[8]) {
    const double a21 = 1.0 / 5.0;
    const double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    const double a41 = 3.0 / 10.0, a42 = -9.0 / 10.0, a43 = 6.0 / 5.0;
    const double a51 = -11.0 / 54.0, a52 = 5.0 / 2.0, a53 = -70.0 / 27.0, a54 = 35.0 / 27.0;
    const double a61 = 1631.0 / 55296.0, a62 = 175.0 / 512.0, a63 = 575.0 / 13824.0,
                 a64 = 44275.0 / 110592.0, a65 = 253.0 / 4096.0;
    const double b1 = 37.0 / 378.0, b3 = 250.0 / 621.0, b4 = 125.0 / 594.0, b6 = 512.0 / 1771.0;
    const double e1 = b1 - 2825.0 / 27648.0, e3 = b3 - 18575.0 / 48384.0,
                 e4 = b4 - 13525.0 / 55296.0, e5 = -277.0 / 14336.0, e6 = b6 - 1.0 / 4.0;
    double k1[8], k2[8], k3[8], k4[8], k5[8], k6[8], yt[8];
    or_rhs(m, y, k1);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a21 * k1[i]);
    or_rhs(m, yt, k2);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
    or_rhs(m, yt, k3);
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
    or_rhs(m, yt, k4);
    for (int i = 0; i < 8; ++i)
        yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    or_rhs(m, yt, k5);
    for (int i = 0; i < 8; ++i)
        yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    or_rhs(m, yt, k6);
    for (int i = 0; i < 8; ++i) {
        yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b6 * k6[i]);
        err[i] = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i]);
    }
}

/* n rays (AoS n×8), serial; cls, steps, rejected, y (n×8), dir (n×3) */
void or_trace_adaptive_rays(const geon_metric_t* m, const double* y0, size_t n, double h0,
                            double rtol, double atol, uint64_t max_steps, int32_t* cls,
                            uint64_t* steps, uint64_t* rejected, double* yf, double* dir) {
    for (size_t r = 0; r < n; ++r) {
        double y[8];
        memcpy(y, y0 + 8 * r, sizeof y);
        double h = h0;
        const double h_min = fabs(h0) * 1e-8, h_max = fabs(h0) * 8.0;
        uint64_t k = 0, rej = 0;
        int c = 3;   /* max_steps */
        while (k < max_steps) {
            const int t = or_terminate(m, y);
            if (t >= 0) { c = t; break; }
            double yn[8], err[8];
            or_ck_step(m, y, h, yn, err);
            double e = 0.0;
            for (int i = 0; i < 8; ++i) {
                const double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
                e = fmax(e, fabs(err[i]) / sc);
            }
            /* a non-finite trial step is rejected with the factor 0.1 (fmax
             * above drops a NaN from e); DIVERGED only below h_min */
            int fin = 1;
            for (int i = 0; i < 8; ++i) fin = fin && isfinite(yn[i]) && isfinite(err[i]);
            if (!fin || e > 1.0) {
                ++rej;
                h = h * (!fin ? 0.1 : fmax(0.9 / sqrt(sqrt(e)), 0.1));
                if (!(fabs(h) >= h_min)) { c = 0; break; }
                continue;
            }
            memcpy(y, yn, sizeof y);
            ++k;
            h = h * (e > 1.0498e-3 ? 0.9 / sqrt(sqrt(e)) : 5.0);
            if (fabs(h) > h_max) h = copysign(h_max, h);
        }
        cls[r] = c;
        steps[r] = k;
        rejected[r] = rej;
        memcpy(yf + 8 * r, y, sizeof y);
        double d[3] = {0.0, 0.0, 0.0};
        if (c == 2) or_asymptotic_direction(m, y, y + 4, d);
        memcpy(dir + 3 * r, d, sizeof d);
    }
}

static double null_resid(const geon_metric_t* m, const double y[8]) {
    double g[16];
    or_g(m, y, g);
    return fabs(or_dot4(g, y + 4, y + 4));
}

/* geodesic.rs:62-117 — terminate check first, step, NaN guard over all 8
 * components (pre-step state and n returned), max_steps → Diverged. */
void or_trace(const geon_metric_t* m, const double y0[8], double step, uint64_t max_steps,
              or_trace_result_t* res) {
    double y[8];
    memcpy(y, y0, sizeof y);
    for (uint64_t n = 0; n < max_steps; ++n) {
        int reason = or_terminate(m, y);
        if (reason >= 0) {
            res->cls = reason;
            memcpy(res->y, y, sizeof y);
            res->steps = n;
            res->null_residual = null_resid(m, y);
            return;
        }
        double yn[8];
        or_rk4_step(m, y, step, yn);
        int bad = 0;
        for (int i = 0; i < 8; ++i) bad |= !isfinite(yn[i]);
        if (bad) {
            res->cls = GEON_DIVERGED;
            memcpy(res->y, y, sizeof y);
            res->steps = n;
            res->null_residual = null_resid(m, y);
            return;
        }
        memcpy(y, yn, sizeof y);
    }
    res->cls = GEON_MAXSTEPS;
    memcpy(res->y, y, sizeof y);
    res->steps = max_steps;
    res->null_residual = null_resid(m, y);
}

/* ------------------------------------------------------ thread pool */

typedef struct {
    size_t n, chunk;
    size_t next; /* atomic */
    void (*fn)(void* arg, size_t begin, size_t end);
    void* arg;
} or_pool_t;

static void* or_pool_worker(void* p) {
    or_pool_t* pool = (or_pool_t*)p;
    for (;;) {
        size_t b = __atomic_fetch_add(&pool->next, pool->chunk, __ATOMIC_RELAXED);
        if (b >= pool->n) break;
        size_t e = b + pool->chunk;
        if (e > pool->n) e = pool->n;
        pool->fn(pool->arg, b, e);
    }
    return NULL;
}

/* rayon-style dynamic chunking over [0, n) (render.rs:73-75 analogue). */
static void or_parallel_for(size_t n, size_t chunk, int nthreads,
                            void (*fn)(void*, size_t, size_t), void* arg) {
    if (nthreads < 1) nthreads = 1;
    or_pool_t pool = {n, chunk ? chunk : 1, 0, fn, arg};
    if (nthreads == 1) {
        or_pool_worker(&pool);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    int started = 0;
    for (int t = 0; t < nthreads; ++t)
        if (pthread_create(&th[t], NULL, or_pool_worker, &pool) == 0) ++started;
    if (started == 0) or_pool_worker(&pool);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    free(th);
}

static double or_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
    const geon_metric_t* m;
    const double* y0;
    double step;
    uint64_t max_steps;
    int32_t* cls;
    uint64_t* steps;
    double *yf, *resid, *dir;
} trace_job_t;

static void trace_chunk(void* a, size_t b, size_t e) {
    trace_job_t* J = (trace_job_t*)a;
    for (size_t i = b; i < e; ++i) {
        or_trace_result_t r;
        or_trace(J->m, J->y0 + 8 * i, J->step, J->max_steps, &r);
        if (J->cls) J->cls[i] = r.cls;
        if (J->steps) J->steps[i] = r.steps;
        if (J->yf) memcpy(J->yf + 8 * i, r.y, sizeof r.y);
        if (J->resid) J->resid[i] = r.null_residual;
        if (J->dir) {
            double d[3] = {0.0, 0.0, 0.0};
            if (r.cls == GEON_ESCAPED) or_asymptotic_direction(J->m, r.y, r.y + 4, d);
            memcpy(J->dir + 3 * i, d, sizeof d);
        }
    }
}

double or_trace_rays(const geon_metric_t* m, const double* y0, size_t n, double step,
                     uint64_t max_steps, int nthreads, int32_t* cls, uint64_t* steps,
                     double* yf, double* resid, double* dir) {
    trace_job_t J = {m, y0, step, max_steps, cls, steps, yf, resid, dir};
    double t0 = or_now();
    or_parallel_for(n, 16, nthreads, trace_chunk, &J);
    return or_now() - t0;
}

/* ---------------------------------------------------------------- L3 */

/* camera.rs:47-97 — e0 = δ0/√(−g_tt); Gram–Schmidt of δ1, δ2, δ3 under g. */
int or_tetrad(const geon_metric_t* m, const double observer[4], double tet[16]) {
    double g[16];
    or_g(m, observer, g);
    memset(tet, 0, 16 * sizeof(double));
    double gtt = g[0];
    if (!(gtt < 0.0)) return GEON_E_NOT_TIMELIKE;
    double n0 = 1.0 / sqrt(-gtt);
    tet[0] = n0;
    for (int i = 0; i < 3; ++i) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        v[i + 1] = 1.0;
        for (int k = 0; k <= i; ++k) {
            const double* ek = tet + 4 * k;
            double proj = or_dot4(g, ek, v);
            double norm_k = or_dot4(g, ek, ek);
            double coeff = proj / norm_k;
            for (int mu = 0; mu < 4; ++mu) v[mu] -= coeff * ek[mu];
        }
        double n2 = or_dot4(g, v, v);
        if (!(n2 > 0.0)) return GEON_E_GRAM_SCHMIDT;
        double nn = sqrt(n2);
        for (int mu = 0; mu < 4; ++mu) v[mu] /= nn;
        memcpy(tet + 4 * (i + 1), v, sizeof v);
    }
    return GEON_OK;
}

/* Camera::new: fov_y_rad = fov_y_deg.to_radians() = deg * (PI/180)
 * (camera.rs:29-36); Camera holds the radians (camera.rs:21-26). */
void or_camera_new(const double observer[4], double fov_y_deg, uint32_t width, uint32_t height,
                   geon_camera_t* out) {
    memcpy(out->observer, observer, sizeof out->observer);
    out->fov_y_rad = fov_y_deg * (OR_PI / 180.0);
    out->width = width;
    out->height = height;
}
static double fov_rad(const geon_camera_t* cam) { return cam->fov_y_rad; }

/* camera.rs:113-160 — pixel (i,j)+(du,dv) → local dir (−1, −v, u)/‖·‖ →
 * p^μ = e0 + n1 e1 + n2 e2 + n3 e3. */
void or_ray_state_sub(const geon_camera_t* cam, const double tet[16], uint64_t i, uint64_t j,
                      double du, double dv, double y[8]) {
    double aspect = (double)cam->width / (double)cam->height;
    double h = tan(fov_rad(cam) * 0.5);
    double w = h * aspect;
    double u = (2.0 * ((double)i + du) / (double)cam->width - 1.0) * w;
    double v = -(2.0 * ((double)j + dv) / (double)cam->height - 1.0) * h;
    double dx = -1.0, dy = -v, dz = u;
    double norm = sqrt(dx * dx + dy * dy + dz * dz);
    double n1 = dx / norm, n2 = dy / norm, n3 = dz / norm;
    for (int mu = 0; mu < 4; ++mu) {
        y[mu] = cam->observer[mu];
        y[4 + mu] = tet[mu] + n1 * tet[4 + mu] + n2 * tet[8 + mu] + n3 * tet[12 + mu];
    }
}

/* ---------------------------------------------------------------- sky */

#define PHI_OFFSET 0.37
#define THETA_OFFSET 0.21

/* Rust `f64 as i64`: saturating, NaN → 0. */
static int64_t f2i64(double v) {
    if (isnan(v)) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}
static double rem_euclid_f(double a, double b) {
    double r = fmod(a, b);
    return (r < 0.0) ? r + fabs(b) : r;
}
static int64_t rem_euclid_i(int64_t a, int64_t b) {
    int64_t r = a % b;
    return (r < 0) ? r + (b < 0 ? -b : b) : r;
}
static double clampd(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* sky.rs:26-34 — SplitMix64 finaliser. */
uint64_t or_sky_hash1(uint64_t x);
uint64_t or_sky_hash1(uint64_t x) {
    x *= 0x9E3779B97F4A7C15ull;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}
/* sky.rs:36-39 */
static uint64_t hash2(int64_t a, int64_t b) {
    return or_sky_hash1(((uint64_t)a * 73856093ull) ^ ((uint64_t)b * 19349663ull));
}

/* sky.rs:43-68 */
static void stars(double theta, double phi, int show, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (!show) return;
    double n_phi = 512.0, n_theta = 256.0;
    double phi_shift = rem_euclid_f(phi + PHI_OFFSET, OR_TAU);
    double theta_shift = clampd(theta + THETA_OFFSET, 0.0, OR_PI);
    int64_t i = f2i64(floor(phi_shift / OR_TAU * n_phi));
    int64_t j = f2i64(floor(theta_shift / OR_PI * n_theta));
    uint64_t h = hash2(i, j);
    double prob = (double)(h & 0xFFFF) / 65535.0;
    if (prob < 0.004) {
        double brightness = 0.4 + 0.6 * ((double)((h >> 16) & 0xFF) / 255.0);
        double temp = (double)((h >> 24) & 0xFF) / 255.0;
        out[0] = (float)(brightness * (1.0 - 0.3 * temp));
        out[1] = (float)(brightness * (1.0 - 0.1 * fabs(temp - 0.5)));
        out[2] = (float)(brightness * (0.7 + 0.3 * temp));
    }
}

/* sky.rs:72-104 */
static void sectors(double theta, double phi, float out[3]) {
    static const double palette[6][3] = {
        {0.15, 0.08, 0.20}, {0.08, 0.15, 0.22}, {0.20, 0.15, 0.08},
        {0.10, 0.18, 0.10}, {0.22, 0.10, 0.10}, {0.10, 0.10, 0.22},
    };
    double phi_shift = rem_euclid_f(phi + PHI_OFFSET, OR_TAU);
    double theta_shift = clampd(theta + THETA_OFFSET, 0.0, OR_PI);
    int64_t phi_mod = f2i64(floor(phi_shift / OR_TAU * 6.0));
    int64_t band = f2i64(floor(theta_shift / OR_PI * 3.0));
    const double* base = palette[rem_euclid_i(phi_mod, 6)];
    double band_mod = (rem_euclid_i(band, 2) == 0) ? 1.0 : 0.6;
    int64_t ic = f2i64(floor(phi_shift / OR_TAU * 24.0));
    int64_t jc = f2i64(floor(theta_shift / OR_PI * 12.0));
    double check = (rem_euclid_i(ic + jc, 2) == 0) ? 1.0 : 0.55;
    for (int c = 0; c < 3; ++c) out[c] = (float)(base[c] * band_mod * check);
}

/* sky.rs:107-120 */
void or_sky_sample(const double dir[3], int show_stars, float rgb[3]) {
    double x = dir[0], y = dir[1], z = dir[2];
    double r = fmax(sqrt(x * x + y * y + z * z), 1e-30);
    double theta = acos(clampd(z / r, -1.0, 1.0));
    double phi = atan2(y, x);
    float s[3], st[3];
    sectors(theta, phi, s);
    stars(theta, phi, show_stars, st);
    for (int c = 0; c < 3; ++c) rgb[c] = fminf(s[c] + st[c], 1.0f);
}

/* render.rs:175-179 — gamma_inv is f32 (inferred from f32 powf). */
void or_tonemap_u8(const float* px, size_t n, uint8_t* out) {
    float gamma_inv = 1.0f / 2.2f;
    for (size_t i = 0; i < n; ++i) {
        float c = px[i];
        if (c < 0.0f) c = 0.0f;
        if (c > 1.0f) c = 1.0f;
        float v = powf(c, gamma_inv);
        float q = roundf(v * 255.0f);
        out[i] = (q <= 0.0f || isnan(q)) ? 0 : (q >= 255.0f ? 255 : (uint8_t)q);
    }
}

/* ---------------------------------------------------------------- L4 */

typedef struct {
    const geon_metric_t* m;
    const geon_camera_t* cam;
    const geon_render_cfg_t* cfg;
    double tet[16];
    float* rgb;
    int32_t* cls;
    double* dir;
    uint64_t* steps_pix;
    /* per-pixel scratch for the sequential reduction */
    int32_t* term;
    double* resid;
    uint64_t* maxhit;
    uint64_t* raysteps;
} render_job_t;

/* render.rs:76-143 — per pixel: spa² sub-rays, colour, majority class. */
static void render_chunk(void* a, size_t b, size_t e) {
    render_job_t* J = (render_job_t*)a;
    uint64_t w = J->cam->width;
    uint64_t spa = J->cfg->samples_per_axis ? J->cfg->samples_per_axis : 1;
    double inv_spa = 1.0 / (double)spa;
    for (size_t idx = b; idx < e; ++idx) {
        uint64_t i = idx % w, j = idx / w;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        uint64_t total = 0, hit = 0;
        double max_resid = 0.0;
        uint32_t nc = 0, ne = 0, nd = 0, nmax = 0;
        double first_dir[3] = {0.0, 0.0, 0.0};
        int have_dir = 0;
        for (uint64_t sj = 0; sj < spa; ++sj)
            for (uint64_t si = 0; si < spa; ++si) {
                double du = ((double)si + 0.5) * inv_spa;
                double dv = ((double)sj + 0.5) * inv_spa;
                double y0[8];
                or_ray_state_sub(J->cam, J->tet, i, j, du, dv, y0);
                or_trace_result_t r;
                or_trace(J->m, y0, J->cfg->step, J->cfg->max_steps, &r);
                float c[3];
                if (r.cls == GEON_CAPTURED) {
                    ++nc;
                    memcpy(c, J->cfg->captured_rgb, sizeof c);
                } else if (r.cls == GEON_ESCAPED) {
                    ++ne;
                    double d[3];
                    or_asymptotic_direction(J->m, r.y, r.y + 4, d);
                    if (!have_dir) { memcpy(first_dir, d, sizeof d); have_dir = 1; }
                    or_sky_sample(d, (int)J->cfg->show_stars, c);
                } else {
                    ++nd;
                    if (r.cls == GEON_MAXSTEPS) { ++nmax; ++hit; }
                    memcpy(c, J->cfg->diverged_rgb, sizeof c);
                }
                acc[0] += c[0]; acc[1] += c[1]; acc[2] += c[2];
                total += r.steps;
                if (isfinite(r.null_residual) && r.null_residual > max_resid)
                    max_resid = r.null_residual;
            }
        float nf = (float)(spa * spa);
        if (J->rgb) {
            J->rgb[3 * idx] = acc[0] / nf;
            J->rgb[3 * idx + 1] = acc[1] / nf;
            J->rgb[3 * idx + 2] = acc[2] / nf;
        }
        int32_t term;
        if (nd > 0 && nd >= nc && nd >= ne) term = (nmax == nd) ? GEON_MAXSTEPS : GEON_DIVERGED;
        else if (nc >= ne) term = GEON_CAPTURED;
        else term = GEON_ESCAPED;
        if (J->cls) J->cls[idx] = term;
        if (J->dir) memcpy(J->dir + 3 * idx, first_dir, sizeof first_dir);
        J->term[idx] = term;
        J->steps_pix[idx] = total / (spa * spa);
        J->resid[idx] = max_resid;
        J->maxhit[idx] = hit;
        J->raysteps[idx] = total;
    }
}

int or_render(const geon_metric_t* m, const geon_camera_t* cam, const geon_render_cfg_t* cfg,
              int nthreads, float* rgb, int32_t* cls, double* dir, uint64_t* steps_pix,
              geon_render_stats_t* stats) {
    render_job_t J;
    memset(&J, 0, sizeof J);
    J.m = m; J.cam = cam; J.cfg = cfg;
    int st = or_tetrad(m, cam->observer, J.tet);
    if (st != GEON_OK) return st;
    size_t n = (size_t)cam->width * cam->height;
    J.rgb = rgb; J.cls = cls; J.dir = dir;
    J.term = (int32_t*)malloc(n * sizeof(int32_t));
    J.resid = (double*)malloc(n * sizeof(double));
    J.maxhit = (uint64_t*)malloc(n * sizeof(uint64_t));
    J.raysteps = (uint64_t*)malloc(n * sizeof(uint64_t));
    J.steps_pix = steps_pix ? steps_pix : (uint64_t*)malloc(n * sizeof(uint64_t));
    double t0 = or_now();
    or_parallel_for(n, 4, nthreads, render_chunk, &J);
    double el = or_now() - t0;
    /* render.rs:146-163 — sequential stats reduction */
    geon_render_stats_t S;
    memset(&S, 0, sizeof S);
    uint64_t total = 0;
    for (size_t i = 0; i < n; ++i) {
        if (J.term[i] == GEON_CAPTURED) S.captured++;
        else if (J.term[i] == GEON_ESCAPED) S.escaped++;
        else S.diverged++;
        total += J.steps_pix[i];
        if (J.resid[i] > S.max_null_residual) S.max_null_residual = J.resid[i];
        S.max_steps_hit += J.maxhit[i];
        S.ray_steps += J.raysteps[i];
    }
    S.avg_steps = (double)total / (double)n;
    S.kernel_ms = el * 1e3;
    if (stats) *stats = S;
    free(J.term); free(J.resid); free(J.maxhit); free(J.raysteps);
    if (!steps_pix) free(J.steps_pix);
    return GEON_OK;
}

/* ------------------------------------------------------------- L2' */

/* curvature.rs:25-46 — ∂_α Γ by central FD, times 1/(2h). */
static void d_christoffel(const geon_metric_t* m, const double x[4], double out[256]) {
    double h = m->fd_step;
    for (int alpha = 0; alpha < 4; ++alpha) {
        double xp[4], xm[4], gp[64], gm[64];
        memcpy(xp, x, sizeof xp);
        memcpy(xm, x, sizeof xm);
        xp[alpha] += h;
        xm[alpha] -= h;
        or_christoffel(m, xp, gp);
        or_christoffel(m, xm, gm);
        double inv_2h = 1.0 / (2.0 * h);
        for (int i = 0; i < 64; ++i) out[alpha * 64 + i] = (gp[i] - gm[i]) * inv_2h;
    }
}

/* curvature.rs:48-70 — R^ρ_σμν = ∂_μΓ^ρ_νσ − ∂_νΓ^ρ_μσ + Γ^ρ_μλΓ^λ_νσ − Γ^ρ_νλΓ^λ_μσ. */
int or_riemann(const geon_metric_t* m, const double x[4], double out[256]) {
    double gam[64], dgam[256];
    or_christoffel(m, x, gam);
    d_christoffel(m, x, dgam);
#define G(a, b, c) gam[(a) * 16 + (b) * 4 + (c)]
#define DG(d, a, b, c) dgam[(d) * 64 + (a) * 16 + (b) * 4 + (c)]
    for (int rho = 0; rho < 4; ++rho)
        for (int sigma = 0; sigma < 4; ++sigma)
            for (int mu = 0; mu < 4; ++mu)
                for (int nu = 0; nu < 4; ++nu) {
                    double term1 = DG(mu, rho, nu, sigma);
                    double term2 = DG(nu, rho, mu, sigma);
                    double term3 = 0.0, term4 = 0.0;
                    for (int l = 0; l < 4; ++l) {
                        term3 += G(rho, mu, l) * G(l, nu, sigma);
                        term4 += G(rho, nu, l) * G(l, mu, sigma);
                    }
                    out[rho * 64 + sigma * 16 + mu * 4 + nu] = term1 - term2 + term3 - term4;
                }
#undef G
#undef DG
    return GEON_OK;
}

/* curvature.rs:72-85 — R_μν = R^α_μαν. */
static void ricci_from(const double r[256], double out[16]) {
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) {
            double s = 0.0;
            for (int a = 0; a < 4; ++a) s += r[a * 64 + mu * 16 + a * 4 + nu];
            out[mu * 4 + nu] = s;
        }
}

int or_ricci(const geon_metric_t* m, const double x[4], double out[16]) {
    double r[256];
    or_riemann(m, x, r);
    ricci_from(r, out);
    return GEON_OK;
}

/* curvature.rs:87-97 — R = Σ g^μν R_μν. */
static double scalar_from(const double ginv[16], const double ric[16]) {
    double s = 0.0;
    for (int i = 0; i < 16; ++i) s += ginv[i] * ric[i];
    return s;
}

double or_ricci_scalar(const geon_metric_t* m, const double x[4]) {
    double ginv[16], ric[16];
    or_g_inv(m, x, ginv);
    or_ricci(m, x, ric);
    return scalar_from(ginv, ric);
}

/* curvature.rs:99-110 — G_μν = R_μν − ½ g_μν R. The reference evaluates
 * Riemann twice (via ricci and via ricci_scalar); both evaluations are the
 * same deterministic computation, so one is reused here. */
int or_einstein(const geon_metric_t* m, const double x[4], double out[16]) {
    double g[16], ginv[16], ric[16];
    or_g(m, x, g);
    or_ricci(m, x, ric);
    or_g_inv(m, x, ginv);
    double r = scalar_from(ginv, ric);
    for (int i = 0; i < 16; ++i) out[i] = ric[i] - 0.5 * g[i] * r;
    return GEON_OK;
}

/* curvature.rs:112-122 — T = G · (1/(8π)). */
int or_stress_energy(const geon_metric_t* m, const double x[4], double out[16]) {
    double ein[16];
    or_einstein(m, x, ein);
    double factor = 1.0 / (8.0 * OR_PI);
    for (int i = 0; i < 16; ++i) out[i] = ein[i] * factor;
    return GEON_OK;
}

/* curvature.rs:127-136 */
double or_nec_contraction(const geon_metric_t* m, const double x[4], const double k[4]) {
    double t[16];
    or_stress_energy(m, x, t);
    double s = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) s += t[mu * 4 + nu] * k[mu] * k[nu];
    return s;
}

typedef struct {
    const geon_metric_t* m;
    const double* x;
    int32_t what;
    double* out;
} curv_job_t;

static void curv_chunk(void* a, size_t b, size_t e) {
    curv_job_t* J = (curv_job_t*)a;
    for (size_t i = b; i < e; ++i) {
        const double* x = J->x + 4 * i;
        switch (J->what) {
        case GEON_CURV_RIEMANN: or_riemann(J->m, x, J->out + 256 * i); break;
        case GEON_CURV_RICCI: or_ricci(J->m, x, J->out + 16 * i); break;
        case GEON_CURV_EINSTEIN: or_einstein(J->m, x, J->out + 16 * i); break;
        case GEON_CURV_STRESS: or_stress_energy(J->m, x, J->out + 16 * i); break;
        case GEON_CURV_SCALAR: J->out[i] = or_ricci_scalar(J->m, x); break;
        default: break;
        }
    }
}

double or_curvature(const geon_metric_t* m, const double* x, size_t n, int32_t what,
                    int nthreads, double* out) {
    curv_job_t J = {m, x, what, out};
    double t0 = or_now();
    or_parallel_for(n, 8, nthreads, curv_chunk, &J);
    return or_now() - t0;
}

/* ------------------------------------------------------------ harness */

/* validate_shadow.rs:20-35 */
void or_equatorial_ray(double mass, double r_obs, double alpha, double y[8]) {
    double f = 1.0 - 2.0 * mass / r_obs;
    double sin_a = sin(alpha), cos_a = cos(alpha);
    y[0] = 0.0; y[1] = r_obs; y[2] = OR_FRAC_PI_2; y[3] = 0.0;
    y[4] = 1.0 / sqrt(f);
    y[5] = -cos_a * sqrt(f);
    y[6] = 0.0;
    y[7] = sin_a / r_obs;
}

/* validate_shadow.rs:37-41 (Schwarzschild::new(mass): pad 1.001, esc 1000) */
int or_is_captured(double mass, double r_obs, double alpha, double step, uint64_t max_steps) {
    geon_metric_t m;
    or_metric_default(GEON_METRIC_SCHWARZSCHILD, &m);
    m.p[0] = mass;
    double y0[8];
    or_equatorial_ray(mass, r_obs, alpha, y0);
    or_trace_result_t r;
    or_trace(&m, y0, step, max_steps, &r);
    return r.cls == GEON_CAPTURED;
}

/* validate_shadow.rs:62-79 — bisection on α ∈ [0, π/2 − 1e-3]. */
void or_validate_shadow(double mass, double r_obs, double step, uint64_t max_steps, int iters,
                        double* alpha_c, double* b_c) {
    double f = 1.0 - 2.0 * mass / r_obs;
    double lo = 0.0, hi = OR_FRAC_PI_2 - 1e-3;
    for (int it = 0; it < iters; ++it) {
        double mid = 0.5 * (lo + hi);
        if (or_is_captured(mass, r_obs, mid, step, max_steps)) lo = mid;
        else hi = mid;
    }
    double a = 0.5 * (lo + hi);
    *alpha_c = a;
    *b_c = r_obs * sin(a) / sqrt(f);
}

/* energy_conditions.rs:104-118 — Σ_{x=-5, x+=0.05, x<=5} min(T_kk,0)·dx
 * with k = (1,1,0,0); the accumulated x keeps its float drift. */
double or_alcubierre_nec_integral(const geon_metric_t* m) {
    double k[4] = {1.0, 1.0, 0.0, 0.0};
    double integral = 0.0, dx = 0.05, x_pos = -5.0;
    while (x_pos <= 5.0) {
        double x[4] = {0.0, x_pos, 0.0, 0.0};
        double nec = or_nec_contraction(m, x, k);
        if (nec < 0.0) integral += nec * dx;
        x_pos += dx;
    }
    return integral;
}
