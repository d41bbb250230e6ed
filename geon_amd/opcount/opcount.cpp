// opcount.cpp — counts the operations of the device formulation.
//
// The gfx950 kernels' per-ray code (geon_device.hpp, FP64; geon_fp32.hpp,
// the FP32 side of the precision split) is compiled for the host with every
// `double` and `float` replaced by counting scalars, then one RK4 step per
// metric is evaluated. The tallies are the algorithmic operation counts of exactly
// the arithmetic the kernels execute (structural zeros already removed):
// add, sub, mul, div, sqrt, sin, cos, tanh count 1 each (sincos counts 2,
// an FMA counts as a mul and an add).
// Negation, |x|, comparisons, max/min are listed but not counted: on
// gfx950 they are operand modifiers or compare/select instructions.
//
// Build and run (prints JSON): see geon_amd/build.py (build_opcount).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <type_traits>
#include <utility>

#include <hip/hip_runtime.h>

struct OpTally {
    unsigned long long add, sub, mul, div, sqrt, sin, cos, tanh, acos, atan2, fmod, floor;
    unsigned long long neg, cmp, fabs, fmax;
    unsigned long long counted() const {
        return add + sub + mul + div + sqrt + sin + cos + tanh + acos + atan2 + fmod + floor;
    }
};
static OpTally T;

struct cnt_d {
    double v;
    cnt_d() = default;
    cnt_d(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
    cnt_d& operator+=(cnt_d o) { ++T.add; v += o.v; return *this; }
    cnt_d& operator-=(cnt_d o) { ++T.sub; v -= o.v; return *this; }
    cnt_d& operator*=(cnt_d o) { ++T.mul; v *= o.v; return *this; }
    cnt_d& operator/=(cnt_d o) { ++T.div; v /= o.v; return *this; }
};
inline cnt_d operator+(cnt_d a, cnt_d b) { ++T.add; return cnt_d(a.v + b.v); }
inline cnt_d operator-(cnt_d a, cnt_d b) { ++T.sub; return cnt_d(a.v - b.v); }
inline cnt_d operator*(cnt_d a, cnt_d b) { ++T.mul; return cnt_d(a.v * b.v); }
inline cnt_d operator/(cnt_d a, cnt_d b) { ++T.div; return cnt_d(a.v / b.v); }
inline cnt_d operator-(cnt_d a) { ++T.neg; return cnt_d(-a.v); }
inline bool operator<(cnt_d a, cnt_d b) { ++T.cmp; return a.v < b.v; }
inline bool operator>(cnt_d a, cnt_d b) { ++T.cmp; return a.v > b.v; }
inline bool operator<=(cnt_d a, cnt_d b) { ++T.cmp; return a.v <= b.v; }
inline bool operator>=(cnt_d a, cnt_d b) { ++T.cmp; return a.v >= b.v; }
inline bool operator==(cnt_d a, cnt_d b) { ++T.cmp; return a.v == b.v; }
inline bool operator!=(cnt_d a, cnt_d b) { ++T.cmp; return a.v != b.v; }
inline cnt_d sqrt(cnt_d a) { ++T.sqrt; return cnt_d(::sqrt(a.v)); }
inline cnt_d sin(cnt_d a) { ++T.sin; return cnt_d(::sin(a.v)); }
inline cnt_d cos(cnt_d a) { ++T.cos; return cnt_d(::cos(a.v)); }
inline void sincos(cnt_d a, cnt_d* s, cnt_d* c) {
    ++T.sin;
    ++T.cos;
    *s = cnt_d(::sin(a.v));
    *c = cnt_d(::cos(a.v));
}
inline cnt_d tanh(cnt_d a) { ++T.tanh; return cnt_d(::tanh(a.v)); }
inline cnt_d acos(cnt_d a) { ++T.acos; return cnt_d(::acos(a.v)); }
inline cnt_d atan2(cnt_d a, cnt_d b) { ++T.atan2; return cnt_d(::atan2(a.v, b.v)); }
inline cnt_d fmod(cnt_d a, cnt_d b) { ++T.fmod; return cnt_d(::fmod(a.v, b.v)); }
inline cnt_d floor(cnt_d a) { ++T.floor; return cnt_d(::floor(a.v)); }
inline cnt_d fabs(cnt_d a) { ++T.fabs; return cnt_d(::fabs(a.v)); }
inline cnt_d fmax(cnt_d a, cnt_d b) { ++T.fmax; return cnt_d(::fmax(a.v, b.v)); }
inline cnt_d cnt_fma(cnt_d a, cnt_d b, cnt_d c) {
    ++T.mul;
    ++T.add;
    return cnt_d(::fma(a.v, b.v, c.v));
}
inline bool cnt_isfinite(cnt_d a) { return ::isfinite(a.v); }
inline bool cnt_isnan(cnt_d a) { return ::isnan(a.v); }

// a quotient from a precomputed reciprocal is one algorithmic division
inline cnt_d div_rn(cnt_d a, cnt_d b, cnt_d) { ++T.div; return cnt_d(a.v / b.v); }
inline cnt_d div_rn_pz(cnt_d a, cnt_d b, cnt_d) { ++T.div; return cnt_d(a.v / b.v); }
inline cnt_d rcp_rn(cnt_d a) { ++T.div; return cnt_d(1.0 / a.v); }
inline cnt_d div_ns(cnt_d a, cnt_d b) { ++T.div; return cnt_d(a.v / b.v); }
// the kernels' own sincos is one sine and one cosine
inline void sincos_cw(cnt_d a, cnt_d* s, cnt_d* c) { sincos(a, s, c); }
// and its tanh one tanh
inline cnt_d tanh_cw(cnt_d a) { return tanh(a); }
#define GEON_OPCOUNT 1
#define __builtin_isfinite(x) cnt_isfinite(x)
#define __builtin_fma(a, b, c) cnt_fma(a, b, c)
#define __builtin_isnan(x) cnt_isnan(x)
#define double cnt_d
#include "../csrc/kernels/geon_device.hpp"
#undef double
#undef __builtin_isfinite
#undef __builtin_fma
#undef __builtin_isnan

// ---- FP32 side: the same for every `float` of geon_fp32.hpp --------------
// The reciprocal builtin counts as one division, __expf as one
// transcendental (tallied with tanh), sqrtf as sqrt.
struct cnt_f {
    float v;
    cnt_f() = default;
    cnt_f(float x) : v(x) {}
    explicit cnt_f(cnt_d x) : v((float)x.v) {}
    explicit operator float() const { return v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
    cnt_f& operator+=(cnt_f o) { ++T.add; v += o.v; return *this; }
    cnt_f& operator-=(cnt_f o) { ++T.sub; v -= o.v; return *this; }
    cnt_f& operator*=(cnt_f o) { ++T.mul; v *= o.v; return *this; }
};
inline cnt_f operator+(cnt_f a, cnt_f b) { ++T.add; return cnt_f(a.v + b.v); }
inline cnt_f operator-(cnt_f a, cnt_f b) { ++T.sub; return cnt_f(a.v - b.v); }
inline cnt_f operator*(cnt_f a, cnt_f b) { ++T.mul; return cnt_f(a.v * b.v); }
inline cnt_f operator/(cnt_f a, cnt_f b) { ++T.div; return cnt_f(a.v / b.v); }
inline cnt_f operator-(cnt_f a) { ++T.neg; return cnt_f(-a.v); }
inline bool operator<(cnt_f a, cnt_f b) { ++T.cmp; return a.v < b.v; }
inline bool operator>(cnt_f a, cnt_f b) { ++T.cmp; return a.v > b.v; }
inline bool operator<=(cnt_f a, cnt_f b) { ++T.cmp; return a.v <= b.v; }
inline bool operator>=(cnt_f a, cnt_f b) { ++T.cmp; return a.v >= b.v; }
inline bool operator==(cnt_f a, cnt_f b) { ++T.cmp; return a.v == b.v; }
inline cnt_f fabsf(cnt_f a) { ++T.fabs; return cnt_f(::fabsf(a.v)); }
inline cnt_f fmaxf(cnt_f a, cnt_f b) { ++T.fmax; return cnt_f(::fmaxf(a.v, b.v)); }
inline cnt_f fminf(cnt_f a, cnt_f b) { ++T.fmax; return cnt_f(::fminf(a.v, b.v)); }
inline void sincosf(cnt_f a, cnt_f* s, cnt_f* c) {
    ++T.sin;
    ++T.cos;
    *s = cnt_f(::sinf(a.v));
    *c = cnt_f(::cosf(a.v));
}
inline void sincosf_cw(cnt_f a, cnt_f* s, cnt_f* c) { sincosf(a, s, c); }
inline cnt_f cnt_rcpf(cnt_f a) { ++T.div; return cnt_f(1.0f / a.v); }
inline cnt_f cnt_sqrtf(cnt_f a) { ++T.sqrt; return cnt_f(::sqrtf(a.v)); }
inline cnt_f cnt_expf(cnt_f a) { ++T.tanh; return cnt_f(::expf(a.v)); }
inline bool cnt_isfinite(cnt_f a) { return ::isfinite(a.v); }
inline cnt_f cnt_fmaf(cnt_f a, cnt_f b, cnt_f c) {
    ++T.mul;
    ++T.add;
    return cnt_f(::fmaf(a.v, b.v, c.v));
}

#define __builtin_amdgcn_rcpf(x) cnt_rcpf(x)
#define __builtin_fmaf(a, b, c) cnt_fmaf(a, b, c)
#define __builtin_sqrtf(x) cnt_sqrtf(x)
#define __expf(x) cnt_expf(x)
#define __builtin_isfinite(x) cnt_isfinite(x)
#define float cnt_f
#include "../csrc/kernels/geon_fp32.hpp"
#undef float
#undef __builtin_fmaf
#undef __builtin_amdgcn_rcpf
#undef __builtin_sqrtf
#undef __expf
#undef __builtin_isfinite

using namespace geon;

static geon_dmetric metric(int kind) {
    geon_dmetric P;
    memset(&P, 0, sizeof P);
    P.kind = kind;
    P.M = 1.0;
    P.cap_r = 1.005 * 2.0;
    P.esc_r = 100.0;
    P.esc2 = 1e6;
    P.esc_l = 1000.0;
    P.fd_h = 1e-5;
    P.fd_two_h = 2e-5;
    P.fd_inv_2h = 1.0 / 2e-5;
    P.fd_sin_h = ::sin(1e-5);
    P.fd_cos_h = ::cos(1e-5);
    P.b0 = 1.0;
    P.R = 2.0;
    P.s = 0.125;
    P.s_pow2 = 1;
    P.inv_s = 8.0;
    P.v = 0.5;
    P.ship_x = 0.0;
    P.denom = 2.0 * ::tanh(16.0);
    P.flat_rs = 2.0 + 22.0 * 0.125 + 2e-5;
    P.mk = 1;   // as make_dmetric sets it for these parameters
    return P;
}

// flat_known: a step inside a run of steps already known to be flat (the
// steady state beyond the Alcubierre wall, where the flatness test is
// skipped)
template <class MP>
static OpTally step_ops(const geon_dmetric& P, const double* y0, bool flat_known = false) {
    cnt_d y[8], out[8];
    for (int i = 0; i < 8; ++i) y[i] = y0[i];
    FlatRun fr;
    fr.left = flat_known ? 1 : 0;
    memset(&T, 0, sizeof T);
    rk4_step<MP>(P, y, cnt_d(0.05), out, fr);
    return T;
}

template <class MF>
static OpTally step_ops_f(const geon_dmetric& Pd, const double* y0, bool flat_known = false) {
    const FParams P = fparams(Pd);
    cnt_f y[8], out[8];
    for (int i = 0; i < 8; ++i) y[i] = cnt_f((float)y0[i]);
    cnt_f s = cnt_f(::sinf((float)y0[2])), c = cnt_f(::cosf((float)y0[2]));
    FlatRunF fr;
    fr.left = flat_known ? 1 : 0;
    fr.j = flat_known ? 1 : 0;
    memset(&T, 0, sizeof T);
    rk4_step_f<MF>(P, y, cnt_f(0.05f), out, s, c, fr);
    return T;
}

// ---- curvature per point (curvature.rs:25-141) ----------------------------
// The formulation of the curvature kernels for one point with every
// structural zero left out: Γ at the point and at x ± h·e_d for each
// coordinate d that Γ depends on (the policies' own christoffel), the ∂Γ
// entries Ricci reads, each Ricci component as one signed sum of its nonzero
// terms, R = Σ g^μν R_μν, G = R_μν − ½ g_μν R and T = G/(8π) (what = 2 stops
// at G). A sum of k terms is k − 1 additions; terms that are structurally
// zero cost nothing.
struct SAcc {
    cnt_d v = cnt_d(0.0);
    bool any = false;
    void add(cnt_d t) { v = any ? v + t : t; any = true; }
    void sub(cnt_d t) { v = any ? v - t : -t; any = true; }
};

template <class MP>
static OpTally curvature_ops(const geon_dmetric& P, const double* x0, int what) {
    memset(&T, 0, sizeof T);
    cnt_d x[4];
    for (int i = 0; i < 4; ++i) x[i] = x0[i];
    cnt_d gam[64], D1[64], D2[64];
    MP::christoffel(P, x, gam);
    for (int d = 0; d < 4; ++d) {
        if (!MP::dep(d)) continue;
        cnt_d xp[4], xm[4], gp[64], gm[64];
        for (int i = 0; i < 4; ++i) xp[i] = xm[i] = x[i];
        xp[d] += cnt_d(P.fd_h);
        xm[d] -= cnt_d(P.fd_h);
        MP::christoffel(P, xp, gp);
        MP::christoffel(P, xm, gm);
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c) {
                const int i1 = d * 16 + b * 4 + c, i2 = b * 16 + b * 4 + c;
                if (MP::gam_nz(d, b, c)) D1[d * 16 + b * 4 + c] = (gp[i1] - gm[i1]) * cnt_d(P.fd_inv_2h);
                if (MP::gam_nz(b, b, c)) D2[d * 16 + b * 4 + c] = (gp[i2] - gm[i2]) * cnt_d(P.fd_inv_2h);
            }
    }
    cnt_d ric[16];
    bool ric_nz[16];
    for (int sg = 0; sg < 4; ++sg)
        for (int nu = 0; nu < 4; ++nu) {
            SAcc s;
            for (int a = 0; a < 4; ++a) {
                if (MP::dep(a) && MP::gam_nz(a, nu, sg)) s.add(D1[a * 16 + nu * 4 + sg]);
                if (MP::dep(nu) && MP::gam_nz(a, a, sg)) s.sub(D2[nu * 16 + a * 4 + sg]);
                for (int l = 0; l < 4; ++l) {
                    if (MP::gam_nz(a, a, l) && MP::gam_nz(l, nu, sg))
                        s.add(gam[a * 16 + a * 4 + l] * gam[l * 16 + nu * 4 + sg]);
                    if (MP::gam_nz(a, nu, l) && MP::gam_nz(l, a, sg))
                        s.sub(gam[a * 16 + nu * 4 + l] * gam[l * 16 + a * 4 + sg]);
                }
            }
            ric[sg * 4 + nu] = s.v;
            ric_nz[sg * 4 + nu] = s.any;
        }
    cnt_d gi[16], g[16];
    MP::ginv_at(P, x, gi);
    SAcc R;
    for (int i = 0; i < 16; ++i)
        if (MP::ginv_nz(i / 4, i % 4) && ric_nz[i]) R.add(gi[i] * ric[i]);
    MP::g(P, x, g);
    for (int i = 0; i < 16; ++i) {
        SAcc e;
        if (ric_nz[i]) e.add(ric[i]);
        if (MP::g_nz(i / 4, i % 4) && R.any) e.sub(cnt_d(0.5) * g[i] * R.v);
        if (e.any && what == 3) e.v = e.v * cnt_d(1.0 / (8.0 * M_PI));
    }
    return T;
}

static void emit(const char* name, const OpTally& t, bool last) {
    printf("    \"%s\": {\"ops\": %llu, \"add\": %llu, \"sub\": %llu, \"mul\": %llu, \"div\": %llu, "
           "\"sqrt\": %llu, \"sin\": %llu, \"cos\": %llu, \"tanh\": %llu, "
           "\"not_counted\": {\"neg\": %llu, \"cmp\": %llu, \"fabs\": %llu, \"fmax\": %llu}}%s\n",
           name, t.counted(), t.add, t.sub, t.mul, t.div, t.sqrt, t.sin, t.cos, t.tanh, t.neg, t.cmp,
           t.fabs, t.fmax, last ? "" : ",");
}

int main() {
    // representative states: off the pole (cot branch taken), inside the
    // Alcubierre wall (full Γ) and outside it (exact flat early-out)
    const double ys[8] = {0.0, 10.0, 1.2, 0.3, 1.3, -0.4, 0.02, 0.03};
    const double ym[8] = {0.0, 5.0, 1.2, 0.3, 1.0, -0.4, 0.02, 0.03};
    const double ya_wall[8] = {0.0, 2.0, 0.3, 0.2, 1.2, -0.5, 0.3, 0.1};
    const double ya_flat[8] = {0.0, 8.0, 0.3, 0.2, 1.2, -0.5, 0.3, 0.1};
    const double yk[8] = {0.0, 5.0, 1.0, 2.0, 1.0, -0.6, 0.0, 0.8};
    printf("{\n  \"unit\": \"FP64 operations per RK4 ray-step (rk4_step of geon_device.hpp) and per curvature point\",\n");
    printf("  \"ray_step\": {\n");
    emit("minkowski", step_ops<Minkowski>(metric(0), yk), false);
    emit("schwarzschild", step_ops<Schwarzschild>(metric(1), ys), false);
    emit("morris_thorne", step_ops<MorrisThorne>(metric(3), ym), false);
    emit("alcubierre_wall", step_ops<Alcubierre>(metric(2), ya_wall), false);
    emit("alcubierre_flat", step_ops<Alcubierre>(metric(2), ya_flat, true), true);
    printf("  },\n  \"ray_step_fp32\": {\n");
    emit("minkowski", step_ops_f<MinkowskiF>(metric(0), yk), false);
    emit("schwarzschild", step_ops_f<SchwarzschildF>(metric(1), ys), false);
    emit("morris_thorne", step_ops_f<MorrisThorneF>(metric(3), ym), false);
    emit("alcubierre_wall", step_ops_f<AlcubierreF>(metric(2), ya_wall), false);
    emit("alcubierre_flat", step_ops_f<AlcubierreF>(metric(2), ya_flat, true), true);
    // curvature sample points: no branch of any Γ depends on the point here
    const double xs[4] = {0.0, 10.0, 1.2, 0.3};
    const double xa[4] = {0.0, 1.9, 0.4, -0.3};
    const double xm[4] = {0.0, 0.7, 1.5707963267948966, 0.0};
    const double xk[4] = {0.0, 1.0, 2.0, 3.0};
    for (int what = 2; what <= 3; ++what) {
        printf("  },\n  \"%s\": {\n", what == 2 ? "curvature_einstein" : "curvature_stress");
        emit("minkowski", curvature_ops<Minkowski>(metric(0), xk, what), false);
        emit("schwarzschild", curvature_ops<Schwarzschild>(metric(1), xs, what), false);
        emit("morris_thorne", curvature_ops<MorrisThorne>(metric(3), xm, what), false);
        emit("alcubierre", curvature_ops<Alcubierre>(metric(2), xa, what), true);
    }
    printf("  }\n}\n");
    return 0;
}
