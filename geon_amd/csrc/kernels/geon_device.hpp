// geon_device.hpp — gfx950 device code for the per-ray geodesic path.
//
// Metric policies mirror the reference trait (src/metric/mod.rs:19-109):
// each supplies g(x) and either analytic Γ (Schwarzschild, Minkowski) or
// the trait's generic Γ from central-difference ∂g and g⁻¹ (Alcubierre,
// Morris–Thorne). Every structural zero of g, ∂g, g⁻¹ and Γ is a
// compile-time mask; the contractions walk the reference's loop order and
// skip masked terms at compile time. Adding an exact zero never changes an
// IEEE sum, so the skipped work leaves the results identical to the dense
// reference loops (geodesic.rs:49-57, metric/mod.rs:55-66) while costing
// only the nonzero terms. The FP64 build uses no FMA contraction.
// Deliberate exceptions, each at the level of a transcendental's last bit:
// Morris–Thorne's sin(θ ± h) come from angle addition (see
// MorrisThorne::christoffel), and on the GPU its sin/cos θ and Alcubierre's
// tanh come from the routines below (sincos_cw, tanh_cw) rather than the
// library's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

#include "../geon_dev_types.h"

#define GEON_D __device__ __forceinline__

namespace geon {

enum { DIVERGED = 0, CAPTURED = 1, ESCAPED = 2, MAXSTEPS = 3, RUNNING = -1 };

// ---- compile-time loops --------------------------------------------------
template <int N, class F, int... I>
GEON_D void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
GEON_D void sfor(F&& f) {
    sfor_impl<N>(f, std::make_integer_sequence<int, N>{});
}
#define CI(v) decltype(v)::value

// RN(a / b) given y = RN(1 / b), without a division (Markstein's theorem):
// q = a·y lies within an ulp of a/b, the remainder a − b·q is exact in one
// FMA, and q + (a − b·q)·y rounds to RN(a/b). A zero remainder means q is
// exact and already carries the right sign of zero. Valid while a, b and
// the quotient are normal numbers far from overflow, which each caller
// guarantees or guards. Saves the ~11-instruction IEEE division sequence
// where one reciprocal serves several quotients or is a constant.
#ifndef GEON_OPCOUNT   // the op-counting build supplies these, tallied as one division each
GEON_D double div_rn(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-q, b, a);
    return r == 0.0 ? q : __builtin_fma(r, y, q);
}
// The same for a numerator that is never −0: then q + r·y is exact for an
// exact q too (a +0 quotient stays +0), and the zero-remainder select goes.
GEON_D double div_rn_pz(double a, double b, double y) {
    const double q = a * y;
    return __builtin_fma(__builtin_fma(-q, b, a), y, q);
}
// RN(1/x) for 2^-500 < |x| < 2^500: the compiler's IEEE division sequence
// for 1.0/x (hardware reciprocal, two Newton steps, the remainder
// correction) without its operand scaling and its special-case fix-up, which
// leave such operands and quotients unchanged — the same bits in 7
// instructions instead of 11. Callers guard the range.
GEON_D double rcp_rn(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double r = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(r, y, y);
#else
    return 1.0 / x;
#endif
}
// RN(a/b) by the compiler's IEEE division sequence for a/b (hardware
// reciprocal, two Newton steps, the quotient and its FMA remainder
// correction) without the operand scaling and the special-case fix-up,
// which leave the operands and the quotient unchanged while a, b and a/b
// are normal numbers within 2^±900 of 1: the same bits in 8 instructions
// instead of 11. Callers guard the range.
GEON_D double div_ns(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
#else
    return a / b;
#endif
}
#endif

// A constant pinned to an SGPR pair where it is used. Loop-invariant FP64
// constants otherwise end up in VGPRs, and an FMA adding one becomes the
// two-address v_fmac behind a v_mov_b64 copy of the constant; as an SGPR
// operand it is a plain three-address v_fma_f64. The value is unchanged.
GEON_D double sconst(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+s"(c));
#endif
    return c;
}

// The library's sin and cos, for the arguments the routines below do not
// reduce (|x| ≥ 2^20: a Payne–Hanek reduction), as a call: inlined, its
// registers count against every ray kernel's step loop.
struct SinCos {
    double s, c;
};
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline)) inline SinCos sincos_lib(double x) {
#else
inline SinCos sincos_lib(double x) {
#endif
    SinCos r;
    sincos(x, &r.s, &r.c);
    return r;
}

// sin x and cos x together, for |x| < 2^20 (else the library sincos): the
// Cody–Waite reduction x = k·π/2 + r with π/2 split into two doubles and
// both products taken exactly inside FMAs, then Taylor polynomials of sin to
// r^17 and of cos to r^18 on |r| ≤ π/4 (truncation below 1e-19 relative)
// and the quadrant k mod 4. Within 1.5 ulp of the correctly rounded sine and
// cosine (the reduced argument's own rounding and the final sums; checked
// against 40-digit arithmetic by the test suite) at about half the
// instructions of the general library routine.
#ifndef GEON_OPCOUNT   // the op-counting build tallies these as one sine and cosine, one tanh
GEON_D void sincos_cw(double x, double* s, double* c) {
    if (!(fabs(x) < 0x1p20)) {
        const SinCos r = sincos_lib(x);
        *s = r.s;
        *c = r.c;
        return;
    }
    const double k = rint(x * 0x1.45f306dc9c883p-1);          // x·(2/π)
    double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, x);    // − k·hi(π/2)
    r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);           // − k·lo(π/2)
    const double r2 = r * r;
    double ps = 0x1.952c77030ad4ap-49;                         //  1/17!
    ps = __builtin_fma(ps, r2, sconst(-0x1.ae7f3e733b81fp-41));     // −1/15!
    ps = __builtin_fma(ps, r2, sconst(0x1.6124613a86d09p-33));      //  1/13!
    ps = __builtin_fma(ps, r2, sconst(-0x1.ae64567f544e4p-26));     // −1/11!
    ps = __builtin_fma(ps, r2, sconst(0x1.71de3a556c734p-19));      //  1/9!
    ps = __builtin_fma(ps, r2, sconst(-0x1.a01a01a01a01ap-13));     // −1/7!
    ps = __builtin_fma(ps, r2, sconst(0x1.1111111111111p-7));       //  1/5!
    ps = __builtin_fma(ps, r2, sconst(-0x1.5555555555555p-3));      // −1/3!
    const double sr = __builtin_fma(r * r2, ps, r);
    double pc = -0x1.6827863b97d97p-53;                        // −1/18!
    pc = __builtin_fma(pc, r2, sconst(0x1.ae7f3e733b81fp-45));      //  1/16!
    pc = __builtin_fma(pc, r2, sconst(-0x1.93974a8c07c9dp-37));     // −1/14!
    pc = __builtin_fma(pc, r2, sconst(0x1.1eed8eff8d898p-29));      //  1/12!
    pc = __builtin_fma(pc, r2, sconst(-0x1.27e4fb7789f5cp-22));     // −1/10!
    pc = __builtin_fma(pc, r2, sconst(0x1.a01a01a01a01ap-16));      //  1/8!
    pc = __builtin_fma(pc, r2, sconst(-0x1.6c16c16c16c17p-10));     // −1/6!
    pc = __builtin_fma(pc, r2, sconst(0x1.5555555555555p-5));       //  1/4!
    const double cr = __builtin_fma(r2 * r2, pc, __builtin_fma(-0.5, r2, 1.0));
    const int q = (int)k & 3;
    const double sv = (q & 1) ? cr : sr;
    const double cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

// sin x and cos x together to within about 0.75 ulp, for |x| < 2^20 (else
// the library sincos), each also as a two-word value s + sl, c + cl within
// about 0.2 ulp of s. The same Cody–Waite reduction and Taylor polynomials
// as sincos_cw, with corrections that make the result nearly correctly
// rounded: the reduced argument is kept as the exact two-word sum r + rl
// (the FMA yields the rounding error of r − k·lo(π/2)), so sin(r + rl) = r +
// (r³·S(r²) + rl·(1 − r²/2)) and cos(r + rl) = w + ((1 − w) − r²/2 − rl·r +
// r⁴·C(r²)) with w = 1 − r²/2 rounded; the rounding error of w is recovered
// exactly and that of r² by an FMA, so only the final sum rounds at the
// scale of the result, and that sum's rounding error is the low word (exact:
// the leading term dominates the correction). Against glibc's sin and cos
// the rounded values differ in the last bit on about 1.5 % of arguments in
// [0, π] (sincos_cw: 20–25 %; on gfx950 the library's sincos is further
// from glibc's than this).
GEON_D void sincos_acc_dw(double x, double* s, double* sl, double* c, double* cl) {
    if (!(fabs(x) < 0x1p20)) {
        const SinCos r = sincos_lib(x);
        *s = r.s;
        *c = r.c;
        *sl = 0.0;
        *cl = 0.0;
        return;
    }
    const double k = rint(x * 0x1.45f306dc9c883p-1);          // x·(2/π)
    const double r1 = __builtin_fma(-k, 0x1.921fb54442d18p+0, x);   // x − k·hi(π/2), exact
    const double r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r1);  // − k·lo(π/2)
    const double rl = __builtin_fma(-k, 0x1.1a62633145c07p-54, r1 - r);   // its rounding error
    const double z = r * r;
    const double zl = __builtin_fma(r, r, -z);                 // r² = z + zl exactly
    const double hz = 0.5 * z;
    double ps = 0x1.952c77030ad4ap-49;                         //  1/17!
    ps = __builtin_fma(ps, z, sconst(-0x1.ae7f3e733b81fp-41));      // −1/15!
    ps = __builtin_fma(ps, z, sconst(0x1.6124613a86d09p-33));       //  1/13!
    ps = __builtin_fma(ps, z, sconst(-0x1.ae64567f544e4p-26));      // −1/11!
    ps = __builtin_fma(ps, z, sconst(0x1.71de3a556c734p-19));       //  1/9!
    ps = __builtin_fma(ps, z, sconst(-0x1.a01a01a01a01ap-13));      // −1/7!
    ps = __builtin_fma(ps, z, sconst(0x1.1111111111111p-7));        //  1/5!
    ps = __builtin_fma(ps, z, sconst(-0x1.5555555555555p-3));       // −1/3!
    const double ts = __builtin_fma(z * r, ps, __builtin_fma(-hz, rl, rl));
    const double sr = r + ts;
    const double srl = (r - sr) + ts;
    double pc = -0x1.6827863b97d97p-53;                        // −1/18!
    pc = __builtin_fma(pc, z, sconst(0x1.ae7f3e733b81fp-45));       //  1/16!
    pc = __builtin_fma(pc, z, sconst(-0x1.93974a8c07c9dp-37));      // −1/14!
    pc = __builtin_fma(pc, z, sconst(0x1.1eed8eff8d898p-29));       //  1/12!
    pc = __builtin_fma(pc, z, sconst(-0x1.27e4fb7789f5cp-22));      // −1/10!
    pc = __builtin_fma(pc, z, sconst(0x1.a01a01a01a01ap-16));       //  1/8!
    pc = __builtin_fma(pc, z, sconst(-0x1.6c16c16c16c17p-10));      // −1/6!
    pc = __builtin_fma(pc, z, sconst(0x1.5555555555555p-5));        //  1/4!
    const double w = 1.0 - hz;                                 // hz ≤ 0.31: exact remainder
    const double e = __builtin_fma(-0.5, zl, (1.0 - w) - hz);
    const double tc = __builtin_fma(z * z, pc, __builtin_fma(-r, rl, e));
    const double cr = w + tc;
    const double crl = (w - cr) + tc;
    const int q = (int)k & 3;
    const bool sw = q & 1, ns = q & 2, nc = (q + 1) & 2;
    const double sv = sw ? cr : sr, svl = sw ? crl : srl;
    const double cv = sw ? sr : cr, cvl = sw ? srl : crl;
    *s = ns ? -sv : sv;
    *sl = ns ? -svl : svl;
    *c = nc ? -cv : cv;
    *cl = nc ? -cvl : cvl;
}
GEON_D void sincos_acc(double x, double* s, double* c) {
    double sl, cl;
    sincos_acc_dw(x, s, &sl, c, &cl);
}

// sin and cos of θ = θ0 + d from the two-word s0 + s0l = sin θ0 and c0 +
// c0l = cos θ0 (sincos_acc_dw), for |d| ≤ 1/16, by rotation as sincos_rot:
// sin θ = s0 + (s0l + s0·(cos d − 1) + (c0 + c0l)·sin d), cos θ = c0 + (c0l +
// c0·(cos d − 1) − (s0 + s0l)·sin d). The correction is at most about |d|
// times the larger of the two values, so its own rounding errors stay well
// below an ulp of that value: within about 0.6 ulp of max(|sin θ|, |sin θ0|)
// (the same for cos), i.e. about as close to the correctly rounded value as
// sincos_acc's except where θ passes within |d| of a zero of the function,
// where the error is that absolute amount. Larger |d| take sincos_acc.
// FB and ok as sincos_rot.
template <bool FB = true>
GEON_D void sincos_rot_dw(double th, double s0, double s0l, double c0, double c0l, double d,
                          double* s, double* c, bool* ok = nullptr) {
    if constexpr (FB) {
        if (!(fabs(d) <= 0.0625)) {
            sincos_acc(th, s, c);
            return;
        }
    } else {
        *ok = *ok && fabs(d) <= 0.0625;
    }
    const double d2 = d * d;
    double ps = __builtin_fma(0x1.71de3a556c734p-19, d2, sconst(-0x1.a01a01a01a01ap-13));  //  1/9!, −1/7!
    ps = __builtin_fma(ps, d2, sconst(0x1.1111111111111p-7));                             //  1/5!
    ps = __builtin_fma(ps, d2, sconst(-0x1.5555555555555p-3));                            // −1/3!
    const double sd = __builtin_fma(d * d2, ps, d);
    double pc = __builtin_fma(-0x1.27e4fb7789f5cp-22, d2, sconst(0x1.a01a01a01a01ap-16)); // −1/10!, 1/8!
    pc = __builtin_fma(pc, d2, sconst(-0x1.6c16c16c16c17p-10));                           // −1/6!
    pc = __builtin_fma(pc, d2, sconst(0x1.5555555555555p-5));                             //  1/4!
    pc = __builtin_fma(pc, d2, -0.5);                                                     // −1/2!
    const double cm = d2 * pc;
    *s = s0 + __builtin_fma(s0, cm, __builtin_fma(c0, sd, __builtin_fma(c0l, sd, s0l)));
    *c = c0 + __builtin_fma(c0, cm, __builtin_fma(-s0, sd, __builtin_fma(-s0l, sd, c0l)));
}

// tanh x in about a quarter of the instructions of the general library
// routine, within about an ulp (checked against 40-digit arithmetic by the
// test suite). With a = min(|x|, 22) (tanh rounds to 1 from 19.06 on):
// e^{2a} = 2^k·(1 + p) by the Cody–Waite reduction 2a = k·ln 2 + r (ln 2 in
// two parts; the first product is exact inside an FMA, and r = 2a when
// k = 0) and p = e^r − 1 from its Taylor polynomial to r^13 (truncation
// below 2^-56 relative on |r| ≤ ln2/2). The denominator D = 2^k(1 + p) + 1
// is kept as an exact two-word sum, q = 2/D comes from a Newton-refined
// reciprocal and the exact FMA remainder, and tanh a = 1 − q is summed with
// its rounding error carried, so the final rounding dominates the error.
// The rounding errors of r and of p = r + r²·P(r) enter the low word of D.
// NaN stays NaN; the sign of x is copied (tanh(−0) = −0).
GEON_D double tanh_cw(double x) {
    const double ax = fabs(x);
    // |x| ≥ 19.1: 1 − tanh|x| = 2/(e^{2|x|} + 1) < 2^-54, half the spacing
    // below 1, so tanh x rounds to ±1 (the wall arguments (r_s + R)/s of an
    // Alcubierre bubble are past this almost everywhere)
    if (ax >= 19.1) return __builtin_copysign(1.0, x);
    const double a = ax > 22.0 ? 22.0 : ax;
    const double t = a + a;
    const double k = rint(t * 0x1.71547652b82fep+0);            // t / ln 2
    const double r1 = __builtin_fma(-k, 0x1.62e42fefa39efp-1, t);   // − k·hi(ln 2), exact
    const double kl = k * 0x1.abc9e3b39803fp-56;                     // k·lo(ln 2)
    const double r = r1 - kl;
    const double rl = (r1 - r) - kl;        // the rounding error of r
    double P = 0x1.6124613a86d09p-33;                            //  1/13!
    P = __builtin_fma(P, r, sconst(0x1.1eed8eff8d898p-29));      //  1/12!
    P = __builtin_fma(P, r, sconst(0x1.ae64567f544e4p-26));      //  1/11!
    P = __builtin_fma(P, r, sconst(0x1.27e4fb7789f5cp-22));      //  1/10!
    P = __builtin_fma(P, r, sconst(0x1.71de3a556c734p-19));      //  1/9!
    P = __builtin_fma(P, r, sconst(0x1.a01a01a01a01ap-16));      //  1/8!
    P = __builtin_fma(P, r, sconst(0x1.a01a01a01a01ap-13));      //  1/7!
    P = __builtin_fma(P, r, sconst(0x1.6c16c16c16c17p-10));      //  1/6!
    P = __builtin_fma(P, r, sconst(0x1.1111111111111p-7));       //  1/5!
    P = __builtin_fma(P, r, sconst(0x1.5555555555555p-5));       //  1/4!
    P = __builtin_fma(P, r, sconst(0x1.5555555555555p-3));       //  1/3!
    P = __builtin_fma(P, r, 0.5);                                //  1/2!
    const double w = (r * r) * P;
    const double p = r + w;                 // e^r − 1 = p + pl (+ rl·e^r)
    const double pl = (r - p) + w;
    const double s = ldexp(1.0, (int)k);                         // 2^k, 0 ≤ k ≤ 63
    const double B = s * p;
    const double u = s + B;                 // s + B exactly as u + ul (|s| ≥ |B|)
    const double ul = (s - u) + B;
    const double D = u + 1.0;               // u + 1 exactly as D + Dl (u ≥ 1)
    const double Dl = (ul + ((u - D) + 1.0)) + s * __builtin_fma(rl, 1.0 + p, pl);
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(D);     // 1/D, refined twice
#else
    double y = 1.0 / D;                     // host build (tests)
#endif
    y = __builtin_fma(y, __builtin_fma(-D, y, 1.0), y);
    y = __builtin_fma(y, __builtin_fma(-D, y, 1.0), y);
    const double q = 2.0 * y;
    const double ql = __builtin_fma(-q, Dl, __builtin_fma(-q, D, 2.0)) * y;   // 2/(D + Dl) − q
    const double h = 1.0 - q;               // 1 − q exactly as h + (1 − h) − q (q ≤ 1)
    const double res = h + (((1.0 - h) - q) - ql);
    // below 2^-27, x³/3 is under half an ulp of x: tanh x rounds to x
    return ax < 0x1p-27 ? x : __builtin_copysign(res, x);
}
// The same as a call: the fourteen tanh of an Alcubierre Γ, inlined and
// interleaved, would push the ray kernels past 256 VGPRs.
__device__ __attribute__((noinline)) inline double tanh_cw_call(double x) { return tanh_cw(x); }

// sin and cos of θ = θ0 + d from s0 = sin θ0 and c0 = cos θ0, for
// |d| ≤ 1/16 by rotation: sin d = d + d³·P(d²) and cos d − 1 = d²·Q(d²)
// from their Taylor series to d^9 and d^10 (truncation below 1e-20), then
// sin θ = s0 + (s0·(cos d − 1) + c0·sin d) and cos θ = c0 + (c0·(cos d − 1)
// − s0·sin d): the corrections are small, so the result carries the error of
// (s0, c0) plus one rounding. Larger |d| take sincos_cw(θ).
// FB = false: no fallback; *ok is cleared for |d| > 1/16 (the caller then
// discards the result).
template <bool FB = true>
GEON_D void sincos_rot(double th, double s0, double c0, double d, double* s, double* c,
                       bool* ok = nullptr) {
    if constexpr (FB) {
        if (!(fabs(d) <= 0.0625)) {
            sincos_cw(th, s, c);
            return;
        }
    } else {
        *ok = *ok && fabs(d) <= 0.0625;
    }
    const double d2 = d * d;
    double ps = __builtin_fma(0x1.71de3a556c734p-19, d2, sconst(-0x1.a01a01a01a01ap-13));  //  1/9!, −1/7!
    ps = __builtin_fma(ps, d2, sconst(0x1.1111111111111p-7));                             //  1/5!
    ps = __builtin_fma(ps, d2, sconst(-0x1.5555555555555p-3));                            // −1/3!
    const double sd = __builtin_fma(d * d2, ps, d);
    double pc = __builtin_fma(-0x1.27e4fb7789f5cp-22, d2, sconst(0x1.a01a01a01a01ap-16)); // −1/10!, 1/8!
    pc = __builtin_fma(pc, d2, sconst(-0x1.6c16c16c16c17p-10));                           // −1/6!
    pc = __builtin_fma(pc, d2, sconst(0x1.5555555555555p-5));                             //  1/4!
    pc = __builtin_fma(pc, d2, -0.5);                                                     // −1/2!
    const double cm = d2 * pc;
    *s = s0 + __builtin_fma(s0, cm, c0 * sd);
    *c = c0 + __builtin_fma(c0, cm, -(s0 * sd));
}
#endif

// sin θ and cos θ of the spherical charts: sincos_cw on the GPU. The host
// build of this header (the parity tests compare it with the oracle bit for
// bit) takes the C library's, which the oracle uses too, so that it checks
// the formulation itself; sincos_cw has its own accuracy test.
GEON_D void theta_sincos(double x, double* s, double* c) {
#if defined(__HIP_DEVICE_COMPILE__)
    sincos_cw(x, s, c);
#else
    sincos(x, s, c);
#endif
}
// sin θ and cos θ of an RK4 stage whose θ differs from the step's θ0 by the
// stage increment: on the GPU by rotation from (sin θ0, cos θ0); the host
// build evaluates them directly like theta_sincos.
template <bool FB = true>
GEON_D void stage_sincos(double th, double th0, double s0, double c0, double* s, double* c,
                         bool* ok = nullptr) {
#if defined(__HIP_DEVICE_COMPILE__)
    sincos_rot<FB>(th, s0, c0, th - th0, s, c, ok);
#else
    (void)th0; (void)s0; (void)c0; (void)ok;
    sincos(th, s, c);
#endif
}
// Two-word sin θ and cos θ at a step's θ (sincos_acc_dw) and the rounded
// values at its stages by rotation from them (sincos_rot_dw): for metrics
// whose analytic Γ takes sin θ and cos θ at every stage and whose rays near
// a photon sphere amplify last-bit differences orbit by orbit
// (Schwarzschild). The host build evaluates the C library's at every stage,
// as theta_sincos, unless GEON_HOST_DEVICE_MATH asks for the GPU's routines
// (a test compares such traces with the oracle).
GEON_D void theta_sincos_dw(double x, double* s, double* sl, double* c, double* cl) {
#if defined(__HIP_DEVICE_COMPILE__) || defined(GEON_HOST_DEVICE_MATH)
    sincos_acc_dw(x, s, sl, c, cl);
#else
    sincos(x, s, c);
    *sl = 0.0;
    *cl = 0.0;
#endif
}
template <bool FB = true>
GEON_D void stage_sincos_dw(double th, double th0, double s0, double s0l, double c0, double c0l,
                            double* s, double* c, bool* ok = nullptr) {
#if defined(__HIP_DEVICE_COMPILE__) || defined(GEON_HOST_DEVICE_MATH)
    sincos_rot_dw<FB>(th, s0, s0l, c0, c0l, th - th0, s, c, ok);
#else
    (void)th0; (void)s0; (void)s0l; (void)c0; (void)c0l; (void)ok;
    sincos(th, s, c);
#endif
}

GEON_D bool finite8(const double* y) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 8; ++i) ok = ok && __builtin_isfinite(y[i]);
    return ok;
}

// ---- generic Christoffel from FD ∂g (metric/mod.rs:25-69) -----------------
// Policy MP provides: g_nz(μ,ν), dg_nz(α,μ,ν), ginv_nz(μ,σ) constexpr masks,
// g(P, x, g[16]) filling the g_nz entries, ginv(P, g, gi[16]) filling the
// ginv_nz entries.
template <class MP>
struct FD {
    static constexpr bool S_nz(int s, int a, int b) {
        return MP::dg_nz(a, s, b) || MP::dg_nz(b, s, a) || MP::dg_nz(s, a, b);
    }
    static constexpr bool gam_nz(int mu, int a, int b) {
        return (MP::ginv_nz(mu, 0) && S_nz(0, a, b)) || (MP::ginv_nz(mu, 1) && S_nz(1, a, b)) ||
               (MP::ginv_nz(mu, 2) && S_nz(2, a, b)) || (MP::ginv_nz(mu, 3) && S_nz(3, a, b));
    }
    static constexpr bool dir_nz(int a) {
        bool any = false;
        for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n) any = any || MP::dg_nz(a, m, n);
        return any;
    }

    // S(σ,α,β) = (∂_α g_σβ + ∂_β g_σα) − ∂_σ g_αβ with structural zeros dropped
    template <int s, int a, int b>
    GEON_D static double S(const double* dg) {
        constexpr bool A = MP::dg_nz(a, s, b), B = MP::dg_nz(b, s, a), C = MP::dg_nz(s, a, b);
        const double va = dg[a * 16 + s * 4 + b];
        const double vb = dg[b * 16 + s * 4 + a];
        const double vc = dg[s * 16 + a * 4 + b];
        if constexpr (A && B && C) return (va + vb) - vc;
        else if constexpr (A && B) return va + vb;
        else if constexpr (A && C) return va - vc;
        else if constexpr (B && C) return vb - vc;
        else if constexpr (A) return va;
        else if constexpr (B) return vb;
        else return -vc;
    }

    // ∂_a g from the metric at x + h e_a (gp) and x − h e_a (gm)
    template <int a>
    GEON_D static void dg_dir(const geon_dmetric& P, const double* gp, const double* gm,
                              double* dg) {
        sfor<4>([&](auto Mu) {
            sfor<4>([&](auto Nu) {
                constexpr int mu = CI(Mu), nu = CI(Nu);
                if constexpr (MP::dg_nz(a, mu, nu)) {
                    if constexpr (nu < mu && MP::dg_nz(a, nu, mu))
                        dg[a * 16 + mu * 4 + nu] = dg[a * 16 + nu * 4 + mu];  // symmetric g
                    else if (P.mk)   // a difference of two metric values is never −0
                        dg[a * 16 + mu * 4 + nu] =
                            div_rn_pz(gp[mu * 4 + nu] - gm[mu * 4 + nu], P.fd_two_h, P.fd_inv_2h);
                    else
                        dg[a * 16 + mu * 4 + nu] = (gp[mu * 4 + nu] - gm[mu * 4 + nu]) / P.fd_two_h;
                }
            });
        });
    }

    // ∂g by central differences over the coordinates g depends on.
    GEON_D static void dg(const geon_dmetric& P, const double* x, double* dg) {
        sfor<4>([&](auto A) {
            constexpr int a = CI(A);
            if constexpr (dir_nz(a)) {
                double xp[4] = {x[0], x[1], x[2], x[3]};
                double xm[4] = {x[0], x[1], x[2], x[3]};
                xp[a] += P.fd_h;
                xm[a] -= P.fd_h;
                double gp[16], gm[16];
                MP::g(P, xp, gp);
                MP::g(P, xm, gm);
                dg_dir<a>(P, gp, gm, dg);
            }
        });
    }

    // The first σ whose term enters Γ^μ_αβ's sum.
    static constexpr int first_sigma(int mu, int a, int b) {
        for (int sg = 0; sg < 4; ++sg)
            if (MP::ginv_nz(mu, sg) && S_nz(sg, a, b)) return sg;
        return 4;
    }

    // Γ^μ_αβ = 0.5 Σ_σ g^μσ S(σ,α,β) (metric/mod.rs:55-66); Γ^μ_βα equals
    // Γ^μ_αβ bit for bit (S is a commutative sum of the same two terms
    // minus a symmetric entry) and is copied.
    // SIGNED: each sum starts from 0.0 as the reference's does, which turns
    // a −0 sum into +0. Ray steps pass false and start from the first term
    // (one add less per entry): there Γ only enters contract(), whose own sum
    // from 0.0 gives the same value for a +0 and a −0 term, so dp does not
    // change. Curvature and the exported Γ keep the reference's zeros.
    template <bool SIGNED = true>
    GEON_D static void christoffel(const geon_dmetric& P, const double* x, double* gam) {
        double g0[16], gi[16], dgv[64];
        MP::g(P, x, g0);
        MP::ginv(P, g0, gi);
        dg(P, x, dgv);
        assemble<SIGNED>(gi, dgv, gam);
    }

    // The assembly from a g⁻¹ and ∂g (christoffel above, or a metric whose
    // ∂g is written out with shared subexpressions).
    template <bool SIGNED = true>
    GEON_D static void assemble(const double* gi, const double* dgv, double* gam) {
        sfor<4>([&](auto Mu) {
            sfor<4>([&](auto A) {
                sfor<4>([&](auto B) {
                    constexpr int mu = CI(Mu), a = CI(A), b = CI(B);
                    if constexpr (gam_nz(mu, a, b)) {
                        if constexpr (b < a && gam_nz(mu, b, a)) {
                            gam[mu * 16 + a * 4 + b] = gam[mu * 16 + b * 4 + a];
                        } else {
                            double s = 0.0;
                            sfor<4>([&](auto Sg) {
                                constexpr int sg = CI(Sg);
                                if constexpr (MP::ginv_nz(mu, sg) && S_nz(sg, a, b)) {
                                    const double t = gi[mu * 4 + sg] * S<sg, a, b>(dgv);
                                    if constexpr (!SIGNED && sg == first_sigma(mu, a, b)) s = t;
                                    else s += t;
                                }
                            });
                            gam[mu * 16 + a * 4 + b] = 0.5 * s;
                        }
                    }
                });
            });
        });
    }
};

// dp^μ = −Σ_αβ Γ^μ_αβ p^α p^β (geodesic.rs:49-57), masked.
template <class GP>
GEON_D void contract(const double* gam, const double* p, double* dp) {
    sfor<4>([&](auto Mu) {
        constexpr int mu = CI(Mu);
        double s = 0.0;
        sfor<4>([&](auto A) {
            sfor<4>([&](auto B) {
                constexpr int a = CI(A), b = CI(B);
                if constexpr (GP::gam_nz(mu, a, b)) s += gam[mu * 16 + a * 4 + b] * p[a] * p[b];
            });
        });
        dp[mu] = -s;
    });
}

// dot4 (types.rs:22-31) with g's structural zeros dropped.
template <class MP>
GEON_D double dot4(const double* g, const double* a, const double* b) {
    double s = 0.0;
    sfor<4>([&](auto I) {
        sfor<4>([&](auto J) {
            constexpr int i = CI(I), j = CI(J);
            if constexpr (MP::g_nz(i, j)) s += g[i * 4 + j] * a[i] * b[j];
        });
    });
    return s;
}

// ---- Minkowski (minkowski.rs) ---------------------------------------------
struct Minkowski {
    static constexpr int flat_test = 1;   // Γ ≡ 0: every step is a flat step
    static constexpr bool g_nz(int m, int n) { return m == n; }
    static constexpr bool ginv_nz(int m, int s) { return m == s; }
    static constexpr bool gam_nz(int, int, int) { return false; }
    GEON_D static void g(const geon_dmetric&, const double*, double* g) {
        g[0] = -1.0; g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
    }
    GEON_D static void accel(const geon_dmetric&, const double*, double* dp) {
        dp[0] = dp[1] = dp[2] = dp[3] = -0.0;
    }
    GEON_D static int terminate(const geon_dmetric& P, const double* x) {
        double r2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
        return (r2 > P.esc2) ? ESCAPED : RUNNING;
    }
    GEON_D static void christoffel(const geon_dmetric&, const double*, double*) {}
    static constexpr bool dep(int) { return false; }
    GEON_D static void ginv_at(const geon_dmetric& P, const double* x, double* gi) { g(P, x, gi); }
};

// ---- Schwarzschild (schwarzschild.rs) -------------------------------------
struct Schwarzschild {
    static constexpr int flat_test = 0;
    static constexpr bool g_nz(int m, int n) { return m == n; }
    // g depends on r (every diagonal entry) and θ (g_φφ)
    static constexpr bool dg_nz(int a, int m, int n) {
        return (a == 1 && m == n) || (a == 2 && m == 3 && n == 3);
    }
    static constexpr bool gam_nz(int mu, int a, int b) {
        return (mu == 0 && ((a == 0 && b == 1) || (a == 1 && b == 0))) ||
               (mu == 1 && a == b) ||
               (mu == 2 && ((a == 1 && b == 2) || (a == 2 && b == 1) || (a == 3 && b == 3))) ||
               (mu == 3 && ((a == 1 && b == 3) || (a == 3 && b == 1) || (a == 2 && b == 3) ||
                            (a == 3 && b == 2)));
    }
    GEON_D static void g(const geon_dmetric& P, const double* x, double* g) {
        const double r = x[1], m = P.M;
        const double f = 1.0 - 2.0 * m / r;
        const double st = sin(x[2]);
        g[0] = -f;
        g[5] = 1.0 / f;
        g[10] = r * r;
        g[15] = r * r * st * st;
    }
    // schwarzschild.rs:61-107
    // Shared subexpressions are computed once: 2M, r², r²f and r − 2M enter
    // several entries with the same operands, and −M/(r²f) = −(M/(r²f)).
    GEON_D static void christoffel(const geon_dmetric& P, const double* x, double* c) {
        double sin_t, cos_t;
        sincos(x[2], &sin_t, &cos_t);
        christoffel_sc(P, x, sin_t, cos_t, c);
    }
    // the same with sin θ and cos θ given (rk4_step hands each stage its
    // values: theta_sincos_dw at the step's θ, stage_sincos_dw at the stages')
    // FAST: the short sequences whatever the operands, *ok cleared where
    // their range tests fail (the caller then discards the result). IEEE:
    // the IEEE divisions whatever the operands.
    template <bool FAST = false, bool IEEE = false>
    GEON_D static void christoffel_sc(const geon_dmetric& P, const double* x, double sin_t,
                                      double cos_t, double* c, bool* ok = nullptr) {
        const double r = x[1], m = P.M;
        const double two_m = 2.0 * m;
        const double r2 = r * r;
        const double rs = r - two_m;
        // Every quotient is the IEEE one. Where 2^-100 < |r|, |f| < 2^100 and
        // the host found 2^-100 ≤ M ≤ 2^100 (P.mk), all operands and
        // quotients lie within 2^±400 and the shorter sequences give the same
        // bits: 1/r by rcp_rn, 2M/r from it by Markstein's correction (2M is
        // never −0), the rest by div_ns; elsewhere the divisions themselves.
        const double ar = fabs(r);
        const bool rok = P.mk && ar > 0x1p-100 && ar < 0x1p100;
        double inv_r, f;
        if (!IEEE && (FAST || __builtin_expect(rok, 1))) {
            inv_r = rcp_rn(r);
            f = 1.0 - div_rn_pz(two_m, r, inv_r);
        } else {
            inv_r = 1.0 / r;
            f = 1.0 - two_m / r;
        }
        const double af = fabs(f);
        const bool fok = rok && af > 0x1p-100 && af < 0x1p100;
        if constexpr (FAST) *ok = *ok && fok;
        double gtr, gtt, cot;
        if (!IEEE && (FAST || __builtin_expect(fok, 1))) {
            gtr = div_ns(m, r2 * f);
            gtt = div_ns(m * f, r2);
            // cos θ is never zero (π/2 is irrational) nor within 2^-60 of it
            cot = (fabs(sin_t) < 1e-8) ? 0.0 : div_ns(cos_t, sin_t);
        } else {
            gtr = m / (r2 * f);
            gtt = m * f / r2;
            cot = (fabs(sin_t) < 1e-8) ? 0.0 : cos_t / sin_t;
        }
        c[0 * 16 + 0 * 4 + 1] = gtr;
        c[0 * 16 + 1 * 4 + 0] = gtr;
        c[1 * 16 + 0 * 4 + 0] = gtt;
        c[1 * 16 + 1 * 4 + 1] = -gtr;
        c[1 * 16 + 2 * 4 + 2] = -rs;
        c[1 * 16 + 3 * 4 + 3] = -rs * sin_t * sin_t;
        c[2 * 16 + 1 * 4 + 2] = inv_r;
        c[2 * 16 + 2 * 4 + 1] = inv_r;
        c[2 * 16 + 3 * 4 + 3] = -sin_t * cos_t;
        c[3 * 16 + 1 * 4 + 3] = inv_r;
        c[3 * 16 + 3 * 4 + 1] = inv_r;
        c[3 * 16 + 2 * 4 + 3] = cot;
        c[3 * 16 + 3 * 4 + 2] = cot;
    }
    GEON_D static void accel(const geon_dmetric& P, const double* y, double* dp) {
        double c[64];
        christoffel(P, y, c);
        contract<Schwarzschild>(c, y + 4, dp);
    }
    static constexpr bool theta_stages = true;
    static constexpr bool theta_dw = true;
    GEON_D static void accel_sc(const geon_dmetric& P, const double* y, double* dp, double st,
                                double ct) {
        double c[64];
        christoffel_sc(P, y, st, ct, c);
        contract<Schwarzschild>(c, y + 4, dp);
    }
    // accel_sc with the range tests deferred to the end of the step
    // (rk4_step_full): ok cleared where one fails
    static constexpr bool fast_stages = true;
    GEON_D static void accel_sc_fast(const geon_dmetric& P, const double* y, double* dp,
                                     double st, double ct, bool& ok) {
        double c[64];
        christoffel_sc<true>(P, y, st, ct, c, &ok);
        contract<Schwarzschild>(c, y + 4, dp);
    }
    // accel_sc with every quotient an IEEE division (the step taken again)
    // accel_sc with every quotient an IEEE division (the step taken again)
    GEON_D static void accel_sc_ieee(const geon_dmetric& P, const double* y, double* dp,
                                     double st, double ct) {
        double c[64];
        christoffel_sc<false, true>(P, y, st, ct, c);
        contract<Schwarzschild>(c, y + 4, dp);
    }
    static constexpr bool dep(int d) { return d == 1 || d == 2; }
    static constexpr bool ginv_nz(int m, int s) { return m == s; }
    // schwarzschild.rs:47-59
    GEON_D static void ginv_at(const geon_dmetric& P, const double* x, double* gi) {
        const double r = x[1], m = P.M;
        const double f = 1.0 - 2.0 * m / r;
        const double st = sin(x[2]);
        gi[0] = -1.0 / f;
        gi[5] = f;
        gi[10] = 1.0 / (r * r);
        gi[15] = 1.0 / (r * r * st * st);
    }
    GEON_D static int terminate(const geon_dmetric& P, const double* x) {
        const double r = x[1];
        if (!__builtin_isfinite(r)) return DIVERGED;
        if (r < P.cap_r) return CAPTURED;
        if (r > P.esc_r) return ESCAPED;
        return RUNNING;
    }
};

// ---- Morris–Thorne (morris_thorne.rs), generic FD Γ -----------------------
struct MorrisThorne {
    static constexpr int flat_test = 0;
    static constexpr bool g_nz(int m, int n) { return m == n; }
    static constexpr bool dg_nz(int a, int m, int n) {
        return (a == 1 && m == n && (m == 2 || m == 3)) || (a == 2 && m == 3 && n == 3);
    }
    static constexpr bool ginv_nz(int m, int s) { return m == s; }
    static constexpr bool gam_nz(int mu, int a, int b) { return FD<MorrisThorne>::gam_nz(mu, a, b); }
    GEON_D static void g(const geon_dmetric& P, const double* x, double* g) {
        const double l = x[1], b0 = P.b0;
        const double r2 = b0 * b0 + l * l;
        const double st = sin(x[2]);
        g[0] = -1.0; g[5] = 1.0; g[10] = r2;
        g[15] = r2 * st * st;
    }
    // Cofactor inverse of a diagonal matrix: c_ii = Π_{j≠i} m_jj in the
    // cofactor scheme's order, det = m00·c00, g^ii = c_ii/det.
    GEON_D static void ginv(const geon_dmetric&, const double* g, double* gi) {
        if (__builtin_expect(!ginv_fast(g, gi), 0)) ginv_ieee(g, gi);
    }
    // The inverse when 2^-500 < |det| < 2^500 (false otherwise, gi then
    // meaningless).
    // g_tt = −1 and g_ll = 1 (g() sets them), so c00 = (1·m10)·m15 and
    // n5 = (−1·m10)·m15 both round the one product m10·m15 and det = −c00 =
    // n5 bit for bit: c00/det is exactly −1 and n5/det exactly 1 for every
    // finite nonzero det. One division serves the other two quotients; with
    // det ≠ 0 both numerators are nonzero (g_θθ = b0² + l² > 0, g_φφ ≠ 0).
    GEON_D static bool ginv_fast(const double* g, double* gi) {
        bool ok = true;
        ginv_nocheck(g, gi, ok);
        return ok;
    }
    // The short sequence without the branch: ok is cleared where |det| is out
    // of range, and gi is then meaningless.
    GEON_D static void ginv_nocheck(const double* g, double* gi, bool& ok) {
        const double m0 = g[0], m5 = g[5], m10 = g[10], m15 = g[15];
        const double det = m0 * (m5 * m10 * m15);
        const double ad = fabs(det);
        ok = ok && ad > 0x1p-500 && ad < 0x1p500;
        const double y = rcp_rn(det);
        gi[0] = -1.0;
        gi[5] = 1.0;
        gi[10] = div_rn_pz(m0 * m5 * m15, det, y);
        gi[15] = div_rn_pz(m0 * m5 * m10, det, y);
    }
    // The cofactor quotients as IEEE divisions.
    GEON_D static void ginv_ieee(const double* g, double* gi) {
        const double m0 = g[0], m5 = g[5], m10 = g[10], m15 = g[15];
        const double c00 = m5 * m10 * m15;
        const double det = m0 * c00;
        gi[0] = c00 / det;
        gi[5] = (m0 * m10 * m15) / det;
        gi[10] = (m0 * m5 * m15) / det;
        gi[15] = (m0 * m5 * m10) / det;
    }
    // The generic Γ of metric/mod.rs:51-69 with ∂g by central differences
    // (metric/mod.rs:25-43) written out for this metric: g_θθ = b0² + l² and
    // g_φφ = g_θθ sin²θ, so the centre and the l ± h evaluations share
    // sin θ and b0². sin θ and cos θ come from theta_sincos; sin(θ ± h) from
    // rotating them by the constants sin h, cos h (angle addition) instead
    // of two more sines:
    // each differs from a directly evaluated sine by at most an ulp, the
    // same size as one libm against another, so ∂_θ g_φφ moves by
    // ~1e-11 relative (its FD quotient amplifies by 1/h). Every other entry
    // is the value the generic loop produces.
    // SIGNED: as FD::assemble (false in ray steps)
    template <bool SIGNED = true>
    GEON_D static void christoffel(const geon_dmetric& P, const double* x, double* c) {
        double st, ct;
        theta_sincos(x[2], &st, &ct);
        christoffel_sc<SIGNED>(P, x, st, ct, c);
    }
    // the same with st = sin θ and ct = cos θ given. FAST: the inverse's
    // short sequence whatever det is, *ok cleared where its range test fails
    // (the caller then discards the result). IEEE: the inverse's IEEE
    // divisions whatever det is.
    template <bool SIGNED = true, bool FAST = false, bool IEEE = false>
    GEON_D static void christoffel_sc(const geon_dmetric& P, const double* x, double st, double ct,
                                      double* c, bool* ok = nullptr) {
        const double l = x[1], h = P.fd_h, two_h = P.fd_two_h;
        const double b2 = P.b0 * P.b0;
        const double r2 = b2 + l * l;
        double g0[16], gi[16], dgv[64];
        g0[0] = -1.0; g0[5] = 1.0; g0[10] = r2;
        g0[15] = r2 * st * st;
        const double lp = l + h, lm = l - h;
        const double r2p = b2 + lp * lp, r2m = b2 + lm * lm;
        // quotients by 2h from the exact host reciprocal 1/(2h); metric values
        // here are O(b0² + l²), far from overflow in every state that can
        // still escape, and each numerator is a difference of two
        // non-negative values, never −0
        const double i2h = P.fd_inv_2h;
        dgv[1 * 16 + 2 * 4 + 2] = div_rn_pz(r2p - r2m, two_h, i2h);
        dgv[1 * 16 + 3 * 4 + 3] = div_rn_pz(r2p * st * st - r2m * st * st, two_h, i2h);
        const double sp = st * P.fd_cos_h + ct * P.fd_sin_h;   // sin(θ + h)
        const double sm = st * P.fd_cos_h - ct * P.fd_sin_h;   // sin(θ − h)
        dgv[2 * 16 + 3 * 4 + 3] = div_rn_pz(r2 * sp * sp - r2 * sm * sm, two_h, i2h);
        if constexpr (FAST) {
            ginv_nocheck(g0, gi, *ok);
        } else if constexpr (IEEE) {
            ginv_ieee(g0, gi);
        } else {
            ginv(P, g0, gi);
        }
        FD<MorrisThorne>::assemble<SIGNED>(gi, dgv, c);
    }
    GEON_D static void accel(const geon_dmetric& P, const double* y, double* dp) {
        double c[64];
        christoffel<false>(P, y, c);
        contract<MorrisThorne>(c, y + 4, dp);
    }
    // rk4_step hands each stage sin θ and cos θ (stage_sincos)
    static constexpr bool theta_stages = true;
    static constexpr bool theta_dw = false;
    GEON_D static void accel_sc(const geon_dmetric& P, const double* y, double* dp, double st,
                                double ct) {
        double c[64];
        christoffel_sc<false>(P, y, st, ct, c);
        contract<MorrisThorne>(c, y + 4, dp);
    }
    // accel_sc with every range test deferred to the end of the step
    // (rk4_step_full): ok cleared where one fails
    static constexpr bool fast_stages = true;
    GEON_D static void accel_sc_fast(const geon_dmetric& P, const double* y, double* dp,
                                     double st, double ct, bool& ok) {
        double c[64];
        christoffel_sc<false, true>(P, y, st, ct, c, &ok);
        contract<MorrisThorne>(c, y + 4, dp);
    }
    // accel_sc with every quotient an IEEE division (the step taken again)
    GEON_D static void accel_sc_ieee(const geon_dmetric& P, const double* y, double* dp,
                                     double st, double ct) {
        double c[64];
        christoffel_sc<false, false, true>(P, y, st, ct, c);
        contract<MorrisThorne>(c, y + 4, dp);
    }
    static constexpr bool dep(int d) { return d == 1 || d == 2; }
    GEON_D static void ginv_at(const geon_dmetric& P, const double* x, double* gi) {
        double g0[16];
        g(P, x, g0);
        ginv(P, g0, gi);
    }
    GEON_D static int terminate(const geon_dmetric& P, const double* x) {
        const double l = x[1];
        if (!__builtin_isfinite(l)) return DIVERGED;
        return (fabs(l) > P.esc_l) ? ESCAPED : RUNNING;
    }
};

// ---- Alcubierre (alcubierre.rs), generic FD Γ -----------------------------
// tanh of the shape function: tanh_cw on the GPU. The host build of this
// header (the parity tests compare it with the oracle bit for bit) takes the
// C library's, which the oracle uses too, so that it checks the formulation
// itself; tanh_cw has its own accuracy test.
// INLINE: inline the routine (kernels with registers to spare) or call it.
template <bool INLINE = false>
GEON_D double shape_tanh(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (INLINE) return tanh_cw(x);
    else return tanh_cw_call(x);
#else
    return tanh(x);
#endif
}

struct Alcubierre {
    static constexpr int flat_test = 2;   // flat steps decided by flat_steps()
    static constexpr bool g_nz(int m, int n) {
        return m == n || (m == 0 && n == 1) || (m == 1 && n == 0);
    }
    static constexpr bool dg_nz(int a, int m, int n) {
        return a >= 1 && ((m == 0 && n == 0) || (m == 0 && n == 1) || (m == 1 && n == 0));
    }
    static constexpr bool ginv_nz(int m, int s) { return g_nz(m, s); }
    static constexpr bool gam_nz(int mu, int a, int b) { return FD<Alcubierre>::gam_nz(mu, a, b); }

    GEON_D static double rs2(const geon_dmetric& P, const double* x) {
        const double dx = x[1] - P.ship_x, dy = x[2], dz = x[3];
        return dx * dx + dy * dy + dz * dz;
    }
    GEON_D static double rs(const geon_dmetric& P, const double* x) { return sqrt(rs2(P, x)); }
    // alcubierre.rs:33-44 — denom = 2·tanh(R/s) precomputed on the host
    template <bool INLINE_TANH = false>
    GEON_D static double shape(const geon_dmetric& P, double r_s) {
        double tp, tm;
        if (P.s_pow2) {
            tp = shape_tanh<INLINE_TANH>((r_s + P.R) * P.inv_s);
            tm = shape_tanh<INLINE_TANH>((r_s - P.R) * P.inv_s);
        } else {
            tp = shape_tanh<INLINE_TANH>((r_s + P.R) / P.s);
            tm = shape_tanh<INLINE_TANH>((r_s - P.R) / P.s);
        }
        // tp ≥ tm, so the difference is never −0
        return P.mk ? div_rn_pz(tp - tm, P.denom, P.inv_denom) : (tp - tm) / P.denom;
    }
    // The shape function at points r_s close to a centre r0 (curvature
    // stencils, P.shape_exp): both tanh expanded about their values at r0,
    // tanh(u0 + δ) = t + Σ_{k=1..5} c_k δ^k with t = tanh(u0), s2 = 1 − t²,
    // c1 = s2, c2 = −t·s2, c3 = s2(t² − 1/3), c4 = t·s2(2/3 − t²),
    // c5 = s2(2/15 − t² + t⁴) (the derivatives of tanh in terms of tanh),
    // |δ| ≤ 1e-3: the remainder is below 1e-18. The centre's tanh carries
    // its own last-bit error into every stencil point alike, so the
    // differences that the central differences take are formed from the
    // small, accurately rounded corrections.
    struct ShapeExp {
        double S0, r0, inv_S0;  // |x − x_s|² at the centre, its square root, 1/S0
        double up, um;          // (r0 ± R)/s
        double tp, tm;          // their tanh
        double cp[5], cm[5];
    };
    GEON_D static double shape_arg(const geon_dmetric& P, double v) {
        return P.s_pow2 ? v * P.inv_s : v / P.s;
    }
    GEON_D static void tanh_coeffs(double t, double* c) {
        const double t2 = t * t;
        const double s2 = (1.0 - t) * (1.0 + t);
        c[0] = s2;
        c[1] = -t * s2;
        c[2] = s2 * (t2 - 0x1.5555555555555p-2);                    // 1/3
        c[3] = (t * s2) * (0x1.5555555555555p-1 - t2);              // 2/3
        c[4] = s2 * __builtin_fma(t2, t2, 0x1.1111111111111p-3 - t2);   // 2/15
    }
    GEON_D static double tanh_near(double t, const double* c, double d) {
        double q = __builtin_fma(c[4], d, c[3]);
        q = __builtin_fma(q, d, c[2]);
        q = __builtin_fma(q, d, c[1]);
        q = __builtin_fma(q, d, c[0]);
        return __builtin_fma(q, d, t);
    }
    // S0 = dx² + dy² + dz² at the centre, as rs() sums it
    GEON_D static void shape_exp(const geon_dmetric& P, double S0, ShapeExp& E) {
        E.S0 = S0;
        E.r0 = sqrt(S0);
        E.inv_S0 = 1.0 / S0;
        E.up = shape_arg(P, E.r0 + P.R);
        E.um = shape_arg(P, E.r0 - P.R);
        E.tp = shape_tanh<true>(E.up);
        E.tm = shape_tanh<true>(E.um);
        tanh_coeffs(E.tp, E.cp);
        tanh_coeffs(E.tm, E.cm);
    }
    GEON_D static double shape_of_tanh(const geon_dmetric& P, double tp, double tm) {
        return P.mk ? div_rn_pz(tp - tm, P.denom, P.inv_denom) : (tp - tm) / P.denom;
    }
    GEON_D static double shape_near(const geon_dmetric& P, const ShapeExp& E, double r_s) {
        const double tp = tanh_near(E.tp, E.cp, shape_arg(P, r_s + P.R) - E.up);
        const double tm = tanh_near(E.tm, E.cm, shape_arg(P, r_s - P.R) - E.um);
        return shape_of_tanh(P, tp, tm);
    }
    // The same from the stencil point's S = dx² + dy² + dz²: r_s − r0 =
    // r0·(√(1 + ε) − 1) with ε = (S − S0)/S0 from its Taylor series to ε⁵
    // (|ε| ≤ 1e-3: remainder below 2e-20 relative), so neither the square
    // root nor the tanh arguments are formed; both tanh move by (r_s − r0)/s.
    // Points with |ε| > 1e-3 (next to the bubble's centre) take the square
    // root.
    GEON_D static double shape_near_sq(const geon_dmetric& P, const ShapeExp& E, double S) {
        const double eps = (S - E.S0) * E.inv_S0;
        if (!(fabs(eps) <= 1e-3)) return shape_near(P, E, sqrt(S));
        double q = __builtin_fma(eps, 0x1.cp-6, -0x1.4p-5);      //  7/256, −5/128
        q = __builtin_fma(q, eps, 0.0625);                          //  1/16
        q = __builtin_fma(q, eps, -0.125);                          // −1/8
        q = __builtin_fma(q, eps, 0.5);                             //  1/2
        const double d = (E.r0 * (eps * q)) * P.inv_s;
        return shape_of_tanh(P, tanh_near(E.tp, E.cp, d), tanh_near(E.tm, E.cm, d));
    }
    // g from the shape value f at the point (alcubierre.rs:48-63)
    GEON_D static void g_of_f(const geon_dmetric& P, double f, double* g) {
        const double v = P.v;
        g[0] = -1.0 + v * v * f * f;
        g[1] = -v * f;
        g[4] = -v * f;
        g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
    }
    GEON_D static void g(const geon_dmetric& P, const double* x, double* g) {
        g_of_f(P, shape(P, rs(P, x)), g);
    }
    // Cofactor inverse of [[a,b],[b,1]] ⊕ I: det = a − b², g^00 = 1/det,
    // g^01 = g^10 = −b/det, g^11 = a/det, g^22 = g^33 = det/det = 1. The
    // reciprocal serves the other two quotients (a = −1 + v²f² is never −0).
    GEON_D static void ginv(const geon_dmetric& P, const double* g, double* gi) {
        const double a = g[0], b = g[1];
        const double det = a - b * b;
        const double ad = fabs(det);
        if (P.mk && ad > 0x1p-500 && ad < 0x1p500) {
            const double y = rcp_rn(det);
            gi[0] = y;
            gi[1] = div_rn(-b, det, y);
            gi[5] = div_rn_pz(a, det, y);
        } else {
            gi[0] = 1.0 / det;
            gi[1] = -b / det;
            gi[5] = a / det;
        }
        gi[4] = gi[1];
        gi[10] = 1.0;
        gi[15] = 1.0;
    }
    template <bool SIGNED = true>
    GEON_D static void christoffel(const geon_dmetric& P, const double* x, double* c) {
        FD<Alcubierre>::christoffel<SIGNED>(P, x, c);
    }
    static constexpr bool dep(int d) { return d >= 1; }
    GEON_D static void ginv_at(const geon_dmetric& P, const double* x, double* gi) {
        double g0[16];
        g(P, x, g0);
        ginv(P, g0, gi);
    }
    GEON_D static void accel(const geon_dmetric& P, const double* y, double* dp) {
        // Beyond flat_rs both tanh round to exactly 1 at every stencil point,
        // so f and every ∂g vanish and Γ is exactly zero.
        if (rs(P, y) > P.flat_rs) {
            dp[0] = dp[1] = dp[2] = dp[3] = -0.0;
            return;
        }
        double c[64];
        christoffel<false>(P, y, c);
        contract<Alcubierre>(c, y + 4, dp);
    }
    GEON_D static int terminate(const geon_dmetric& P, const double* x) {
        const double r2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
        if (!__builtin_isfinite(r2)) return DIVERGED;
        return (r2 > P.esc2) ? ESCAPED : RUNNING;
    }
    // The number of consecutive steps from y, this one included, whose four
    // RK4 stages all see the exact flat region of accel's early-out (0: not
    // this step). While Γ = 0 the stage momenta equal p, which the folded
    // step leaves unchanged, and every step moves the position by h·p up to
    // rounding; so every stage of K steps lies within K·|h|·|p| of y and its
    // r_s is at least r_s(y) − K·|h|·|p|. The 1e-9·(1 + r) margin on r_s and
    // the relative 1e-6 on the distance cover the rounding of the stage
    // positions and of r_s over K ≤ 65536 steps. NaN or infinite states
    // give 0.
    GEON_D static uint32_t flat_steps(const geon_dmetric& P, const double* y, double h) {
        const double r = rs(P, y);
        const double d = r - P.flat_rs - 1e-9 * (1.0 + r);
        if (!(d > 0.0)) return 0;
        const double p2 = y[5] * y[5] + y[6] * y[6] + y[7] * y[7];
        if (!(d * d > h * h * p2)) return 0;   // one step
        const double k = d / (fabs(h) * sqrt(p2) * (1.0 + 1e-6));
        return k >= 65536.0 ? 65536u : (k >= 1.0 ? (uint32_t)k : 1u);
    }
};

// ---- integrator (integrator.rs:16-40) -------------------------------------
// The step below with dp = −0 at every stage (what accel returns where Γ is
// exactly zero), the zero terms folded: the same operations on the same
// values, so the same bits. Stage momenta are p + hh·(−0) and p + h·(−0);
// momenta advance by (h/6)·(−0).
GEON_D void rk4_step_flat(const double* y, double h, double* out) {
    const double hh = 0.5 * h;
    const double zh = hh * -0.0, zf = h * -0.0;
    const double sixth = h / 6.0;
    const double zs = sixth * -0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double p = y[4 + i];
        const double k23 = p + zh;
        const double t = 2.0 * k23;
        const double acc = ((p + t) + t) + (p + zf);
        out[i] = y[i] + sixth * acc;
        out[4 + i] = p + zs;
    }
}

// k_i = (p_i, dp_i); the combination ((k1 + 2k2) + 2k3) + k4 is accumulated
// in that order. Metrics with flat regions take the folded step there.
// Metrics whose Γ takes sin θ and cos θ as arguments (theta_stages) get
// them from rk4_step: theta_sincos at the step's θ, stage_sincos at the
// stages' (theta_dw: their two-word forms).
template <class MP, class = void>
struct ThetaStages : std::false_type {};
template <class MP>
struct ThetaStages<MP, std::void_t<decltype(MP::theta_stages)>>
    : std::integral_constant<bool, MP::theta_stages> {};
// Metrics whose stages can defer their quotients' range tests (accel_sc_fast)
template <class MP, class = void>
struct FastStages : std::false_type {};
template <class MP>
struct FastStages<MP, std::void_t<decltype(MP::fast_stages)>>
    : std::integral_constant<bool, MP::fast_stages> {};

// Steps of a ray still known to be flat from its last flatness test
// (metrics whose flat regions are tested per step); 0: test again.
struct FlatRun {
    uint32_t left = 0;
    bool nf = false;   // the flatness test of this step already ran and failed
};

// The full step (Γ evaluated at the four stages). FAST (metrics with
// accel_sc_fast): the stages take the short quotient sequences without their
// range tests, and the stage rotation of (sin θ, cos θ) without its fallback
// for increments beyond 1/16, collecting the tests' outcome; a lane where
// one failed takes the step again (IEEE: every quotient an IEEE division,
// the rotation with its fallback). The short sequences equal the IEEE
// quotients wherever their tests pass, so the step's value is the one the
// tested stages give either way. A branch around each quotient in every
// stage made the compiler keep copies of the state and of the merged values
// for the cold side; one branch per step leaves the stages straight-line
// code, and without the rotation's fallback inside them the Morris–Thorne
// step loop fits four waves per SIMD.
template <class MP, bool FAST = FastStages<MP>::value, bool IEEE = false>
GEON_D void rk4_step_full(const geon_dmetric& P, const double* y, double h, double* out) {
    double acc[8], yt[8], dp[4];
    const double hh = 0.5 * h;
    double s0 = 0.0, c0 = 0.0, s0l = 0.0, c0l = 0.0;
    bool ok = true;
    auto accel = [&](const double* ys, double st, double ct) {
        if constexpr (FAST) MP::accel_sc_fast(P, ys, dp, st, ct, ok);
        else if constexpr (IEEE) MP::accel_sc_ieee(P, ys, dp, st, ct);
        else if constexpr (ThetaStages<MP>::value) MP::accel_sc(P, ys, dp, st, ct);
    };
    auto stage = [&](const double* ys) {
        if constexpr (ThetaStages<MP>::value) {
            double st, ct;
            if constexpr (MP::theta_dw)
                stage_sincos_dw<!FAST>(ys[2], y[2], s0, s0l, c0, c0l, &st, &ct, &ok);
            else
                stage_sincos<!FAST>(ys[2], y[2], s0, c0, &st, &ct, &ok);
            accel(ys, st, ct);
        } else {
            MP::accel(P, ys, dp);
        }
    };
    if constexpr (ThetaStages<MP>::value) {
        if constexpr (MP::theta_dw)
            theta_sincos_dw(y[2], &s0, &s0l, &c0, &c0l);
        else
            theta_sincos(y[2], &s0, &c0);
        accel(y, s0, c0);
    } else {
        MP::accel(P, y, dp);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[i] = y[4 + i];
        acc[4 + i] = dp[i];
        yt[i] = y[i] + hh * y[4 + i];
        yt[4 + i] = y[4 + i] + hh * dp[i];
    }
    stage(yt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double kx = yt[4 + i];
        acc[i] = acc[i] + 2.0 * kx;
        acc[4 + i] = acc[4 + i] + 2.0 * dp[i];
        yt[i] = y[i] + hh * kx;
        yt[4 + i] = y[4 + i] + hh * dp[i];
    }
    stage(yt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double kx = yt[4 + i];
        acc[i] = acc[i] + 2.0 * kx;
        acc[4 + i] = acc[4 + i] + 2.0 * dp[i];
        yt[i] = y[i] + h * kx;
        yt[4 + i] = y[4 + i] + h * dp[i];
    }
    stage(yt);
    const double sixth = h / 6.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[i] = acc[i] + yt[4 + i];
        acc[4 + i] = acc[4 + i] + dp[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = y[i] + sixth * acc[i];
    if constexpr (FAST) {
        if (__builtin_expect(!ok, 0)) rk4_step_full<MP, false, true>(P, y, h, out);
    }
}

// Returns true when the step evaluated Γ at its four stages, false when it
// was a flat step.
template <class MP>
GEON_D bool rk4_step(const geon_dmetric& P, const double* y, double h, double* out, FlatRun& fr) {
    if constexpr (MP::flat_test == 1) {
        rk4_step_flat(y, h, out);
        return false;
    } else if constexpr (MP::flat_test == 2) {
        if (fr.left > 0) {
            --fr.left;
            rk4_step_flat(y, h, out);
            return false;
        }
        if (fr.nf) {
            fr.nf = false;
        } else {
            const uint32_t k = MP::flat_steps(P, y, h);
            if (k > 0) {
                fr.left = k - 1;
                rk4_step_flat(y, h, out);
                return false;
            }
        }
    }
    rk4_step_full<MP>(P, y, h, out);
    return true;
}

template <class MP>
GEON_D void rk4_step(const geon_dmetric& P, const double* y, double h, double* out) {
    FlatRun fr;
    rk4_step<MP>(P, y, h, out, fr);
}

// |g(p,p)| at the final position (geodesic.rs:75-77).
template <class MP>
GEON_D double null_residual(const geon_dmetric& P, const double* y) {
    double g[16];
    MP::g(P, y, g);
    return fabs(dot4<MP>(g, y + 4, y + 4));
}

// asymptotic_direction (metric/mod.rs:91-108; alcubierre.rs:81-88;
// morris_thorne.rs:54-77)
template <class MP>
GEON_D void direction(const double* y, double* d) {
    const double* x = y;
    const double* p = y + 4;
    if constexpr (std::is_same<MP, Alcubierre>::value) {
        const double vx = p[1], vy = p[2], vz = p[3];
        const double n = fmax(sqrt(vx * vx + vy * vy + vz * vz), 1e-30);
        d[0] = vx / n; d[1] = vy / n; d[2] = vz / n;
    } else {
        double r, sign;
        if constexpr (std::is_same<MP, MorrisThorne>::value) {
            r = fmax(fabs(x[1]), 1e-30);
            sign = (x[1] < 0.0) ? -1.0 : 1.0;
        } else {
            r = x[1];
            sign = 1.0;
        }
        double sin_t, cos_t, sin_p, cos_p;
        sincos(x[2], &sin_t, &cos_t);
        sincos(x[3], &sin_p, &cos_p);
        const double pr = p[1], pt = p[2], pp = p[3];
        const double vx = sin_t * cos_p * pr + r * cos_t * cos_p * pt - r * sin_t * sin_p * pp;
        const double vy = sin_t * sin_p * pr + r * cos_t * sin_p * pt + r * sin_t * cos_p * pp;
        const double vz = cos_t * pr - r * sin_t * pt;
        const double n = fmax(sqrt(vx * vx + vy * vy + vz * vz), 1e-30);
        d[0] = sign * vx / n; d[1] = vy / n; d[2] = vz / n;
    }
}

// ---- flat spans ---------------------------------------------------------------
// Along a run of flat steps (Γ exactly zero at every stage) the folded step
// leaves the momentum as it is — p + (h/6)·(−0) is p, except that with h < 0
// a −0 component becomes +0 once, which changes no increment below — and
// adds the same increment δ_i = (h/6)·acc_i, acc_i as in rk4_step_flat, to
// every position coordinate: x_{j+1} = x_j + δ with the rounding of the full
// step. Where the positions of the next m steps provably stay inside the
// escape sphere, neither the escape test before each of those steps nor the
// finiteness guard after it can fire, so the m steps are m additions per
// coordinate with the same bits as m checked steps.
//
// The bound: each step moves the spatial position by |δ_s|(1 + 2^-52) plus
// at most 2^-53 of |x| of rounding, so after j ≤ 65536 steps |x_j| ≤ r0 +
// j·|δ_s|(1 + 1e-15) + 1.5e-11·R. With j·|δ_s|(1 + 1e-9) ≤ R(1 − 1e-9) −
// r0(1 + 1e-9) (R = √esc², r0 = |x_0|, both as computed, each within ulps)
// that stays below R(1 − 9e-10), and the computed r² below esc²; t grows by
// |δ_0| a step and stays far from overflow. Returns the number of steps taken
// (0: none guaranteed, take a checked step).
GEON_D uint32_t flat_span_advance(double* y, double h, double esc2, uint32_t cap) {
    const double hh = 0.5 * h;
    const double zh = hh * -0.0, zf = h * -0.0;
    const double sixth = h / 6.0;
    const double zs = sixth * -0.0;
    double d[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double p = y[4 + i];
        const double k23 = p + zh;
        const double t = 2.0 * k23;
        const double acc = ((p + t) + t) + (p + zf);
        d[i] = sixth * acc;
    }
    const double r0 = sqrt(y[1] * y[1] + y[2] * y[2] + y[3] * y[3]);
    const double ds = sqrt(d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
    const double avail = sqrt(esc2) * (1.0 - 1e-9) - r0 * (1.0 + 1e-9);
    const double mj = avail / (ds * (1.0 + 1e-9));   // +inf for δ_s = 0
    if (!(avail > 0.0 && avail < 1e300) || !(mj >= 1.0) ||
        !(fabs(y[0]) + 131072.0 * fabs(d[0]) < 1e300))
        return 0;
    const uint32_t m = mj >= (double)cap ? cap : (uint32_t)mj;
#pragma unroll
    for (int i = 0; i < 4; ++i) y[4 + i] = y[4 + i] + zs;
    double x0 = y[0], x1 = y[1], x2 = y[2], x3 = y[3];
    for (uint32_t j = 0; j < m; ++j) {
        x0 = x0 + d[0];
        x1 = x1 + d[1];
        x2 = x2 + d[2];
        x3 = x3 + d[3];
    }
    y[0] = x0; y[1] = x1; y[2] = x2; y[3] = x3;
    return m;
}

// Spans advance at most this many steps per call. A span is a loop of one
// addition per coordinate and step, and the lanes of a wave run it together
// for as many steps as the longest of them: rays that left a bubble's wall
// at different times test flatness at different distances and would take
// runs of very different lengths in the same pass (the 1024² Alcubierre
// render from x = +10 spent 34 of its 49 ms in the flight to the escape
// sphere). Cut into chunks, long runs advance in step with each other; the
// rest of a run is kept (FlatRun::left) and taken by the next span without
// a new test, so the steps and their bits are unchanged. Measured on that
// render (and x = −10, all flat): 128 → 19.3 + 4.4 ms, 256 → 20.0 + 4.0,
// 1024 → 23.8 + 3.8, 4096 → 34.3 + 3.8, unchunked 49 + 3.8.
#ifndef GEON_SPAN_CHUNK
#define GEON_SPAN_CHUNK 256u
#endif

// The flat span for a state whose escape test just passed, at most `cap`
// steps (and GEON_SPAN_CHUNK): advances y and the run, returns the steps
// taken (0: take an ordinary step, which then knows the flatness test's
// outcome). Metrics with flat regions escape on x² + y² + z² > esc².
template <class MP>
GEON_D uint32_t flat_span(const geon_dmetric& P, double* y, double h, FlatRun& fr, uint32_t cap) {
    if (cap > GEON_SPAN_CHUNK) cap = GEON_SPAN_CHUNK;
    uint32_t run = 65536u;   // flat steps known from this one on (flat_test 1: all)
    if constexpr (MP::flat_test == 2) {
        run = fr.left;
        if (run == 0) {
            run = MP::flat_steps(P, y, h);
            if (run == 0) {
                fr.nf = true;
                return 0;
            }
        }
        fr.left = run;   // an ordinary step takes a flat step while > 0
    }
    const uint32_t m = flat_span_advance(y, h, P.esc2, run < cap ? run : cap);
    if constexpr (MP::flat_test == 2) fr.left = run - m;
    return m;
}

// trace (geodesic.rs:62-117): terminate first, step, NaN guard keeps the
// pre-step state, max_steps → MAXSTEPS. Runs of flat steps go as spans.
// gamma_steps (optional) receives the number of steps that evaluated Γ
// (every step of a metric without flat regions).
template <class MP>
GEON_D int trace(const geon_dmetric& P, double* y, double h, uint32_t max_steps, uint32_t* steps,
                 uint32_t* gamma_steps = nullptr) {
    FlatRun fr;
    uint32_t n = 0, ng = 0;
    int cls = MAXSTEPS;
    if constexpr (MP::flat_test == 0) {
        // Two steps per iteration, the state alternating between ya and yb:
        // the same steps as the loop below, without copying the state back
        // each step (the copy was a register shuffle of the whole state).
        double ya[8], yb[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ya[i] = y[i];
        bool in_b = false;   // the current state is yb
        for (;;) {
            if (n >= max_steps) break;
            int t = MP::terminate(P, ya);
            if (t != RUNNING) {
                cls = t;
                break;
            }
            rk4_step_full<MP>(P, ya, h, yb);
            if (!finite8(yb)) {
                cls = DIVERGED;
                break;
            }
            ++n;
            in_b = true;
            if (n >= max_steps) break;
            t = MP::terminate(P, yb);
            if (t != RUNNING) {
                cls = t;
                break;
            }
            rk4_step_full<MP>(P, yb, h, ya);
            if (!finite8(ya)) {
                cls = DIVERGED;
                break;
            }
            ++n;
            in_b = false;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) y[i] = in_b ? yb[i] : ya[i];
        *steps = n;
        if (gamma_steps) *gamma_steps = n;
        return cls;
    }
    while (n < max_steps) {
        const int t = MP::terminate(P, y);
        if (t != RUNNING) {
            cls = t;
            break;
        }
        if constexpr (MP::flat_test != 0) {
            const uint32_t m = flat_span<MP>(P, y, h, fr, max_steps - n);
            if (m > 0) {
                n += m;
                continue;
            }
        }
        double yn[8];
        const bool full = rk4_step<MP>(P, y, h, yn, fr);
        if (!finite8(yn)) {
            cls = DIVERGED;
            break;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) y[i] = yn[i];
        ++n;
        if constexpr (MP::flat_test == 2) ng += full ? 1u : 0u;
    }
    *steps = n;
    if (gamma_steps) {
        if constexpr (MP::flat_test == 0) *gamma_steps = n;
        else if constexpr (MP::flat_test == 1) *gamma_steps = 0;
        else *gamma_steps = ng;
    }
    return cls;
}

// ---- adaptive stepping -------------------------------------------------------
// The reference integrates with fixed steps and names adaptive stepping as its
// next step (integrator.rs:5-7); this is that step, not a reference method.
// Cash–Karp embedded RK4(5) (J. R. Cash, A. H. Karp, ACM TOMS 16 (1990)
// 201-222): six derivative evaluations, the fifth-order solution advances and
// its difference to the fourth-order one estimates the local error. Every sum
// runs left to right in the tableau's order. (Not part of the op-counting
// build, whose scalar stands in for double.)
#ifndef GEON_OPCOUNT
namespace ck {
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 3.0 / 10.0, a42 = -9.0 / 10.0, a43 = 6.0 / 5.0;
constexpr double a51 = -11.0 / 54.0, a52 = 5.0 / 2.0, a53 = -70.0 / 27.0, a54 = 35.0 / 27.0;
constexpr double a61 = 1631.0 / 55296.0, a62 = 175.0 / 512.0, a63 = 575.0 / 13824.0,
                 a64 = 44275.0 / 110592.0, a65 = 253.0 / 4096.0;
constexpr double b1 = 37.0 / 378.0, b3 = 250.0 / 621.0, b4 = 125.0 / 594.0, b6 = 512.0 / 1771.0;
// fifth- minus fourth-order weights
constexpr double e1 = b1 - 2825.0 / 27648.0, e3 = b3 - 18575.0 / 48384.0,
                 e4 = b4 - 13525.0 / 55296.0, e5 = -277.0 / 14336.0, e6 = b6 - 1.0 / 4.0;
}  // namespace ck

// dy/dλ = (p, dp/dλ) (geodesic.rs:38-59)
template <class MP>
GEON_D void geodesic_rhs(const geon_dmetric& P, const double* y, double* k) {
#pragma unroll
    for (int i = 0; i < 4; ++i) k[i] = y[4 + i];
    MP::accel(P, y, k + 4);
}

template <class MP>
GEON_D void ck_step(const geon_dmetric& P, const double* y, double h, double* yn, double* err) {
    using namespace ck;
    double k1[8], k2[8], k3[8], k4[8], k5[8], k6[8], yt[8];
    geodesic_rhs<MP>(P, y, k1);
#pragma unroll
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a21 * k1[i]);
    geodesic_rhs<MP>(P, yt, k2);
#pragma unroll
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
    geodesic_rhs<MP>(P, yt, k3);
#pragma unroll
    for (int i = 0; i < 8; ++i) yt[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
    geodesic_rhs<MP>(P, yt, k4);
#pragma unroll
    for (int i = 0; i < 8; ++i)
        yt[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    geodesic_rhs<MP>(P, yt, k5);
#pragma unroll
    for (int i = 0; i < 8; ++i)
        yt[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    geodesic_rhs<MP>(P, yt, k6);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        yn[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b6 * k6[i]);
        err[i] = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i]);
    }
}

// trace with adaptive steps: the termination test before each step, the NaN
// guard after it and max_steps (counting accepted steps) as in trace. A step
// is accepted when e = max_i |err_i| / (atol + rtol·max(|y_i|, |y_new_i|)) ≤ 1;
// the next step is h·0.9·e^(−1/4), at most 5h (e below 1.0498e-3) and at
// least h/10, and |h| stays within [1e-8, 8]·|h0| (the upper bound keeps a
// step from jumping a feature the stages would miss) — square roots and
// divisions only, which host and device round alike. A step falling below
// the lower bound ends the trace as DIVERGED.
template <class MP>
GEON_D int trace_adaptive(const geon_dmetric& P, double* y, double h0, double rtol, double atol,
                          uint32_t max_steps, uint32_t* steps, uint32_t* rejected) {
    double h = h0;
    const double h_min = fabs(h0) * 1e-8, h_max = fabs(h0) * 8.0;
    uint32_t n = 0, rej = 0;
    int cls = MAXSTEPS;
    while (n < max_steps) {
        const int t = MP::terminate(P, y);
        if (t != RUNNING) {
            cls = t;
            break;
        }
        double yn[8], err[8];
        ck_step<MP>(P, y, h, yn, err);
        double e = 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double sc = atol + rtol * fmax(fabs(y[i]), fabs(yn[i]));
            e = fmax(e, fabs(err[i]) / sc);
        }
        // a trial step that is not finite (fmax above drops a NaN from e) is
        // rejected like one that is too long, with the smallest factor 0.1;
        // the ray diverges only when h falls below h_min
        const bool bad = !finite8(yn) || !finite8(err);
        if (bad || e > 1.0) {
            ++rej;
            h = h * (bad ? 0.1 : fmax(0.9 / sqrt(sqrt(e)), 0.1));
            if (!(fabs(h) >= h_min)) {
                cls = DIVERGED;
                break;
            }
            continue;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) y[i] = yn[i];
        ++n;
        h = h * (e > 1.0498e-3 ? 0.9 / sqrt(sqrt(e)) : 5.0);
        if (fabs(h) > h_max) h = copysign(h_max, h);
    }
    *steps = n;
    *rejected = rej;
    return cls;
}
#endif  // GEON_OPCOUNT

// camera.rs:113-160 — pixel (i, j) + (du, dv) → initial state.
GEON_D void camera_ray(const geon_dframe& F, uint32_t i, uint32_t j, double du, double dv,
                       double* y) {
    const double u = (2.0 * ((double)i + du) / (double)F.W - 1.0) * F.w_im;
    const double v = -(2.0 * ((double)j + dv) / (double)F.H - 1.0) * F.h_im;
    const double dx = -1.0, dy = -v, dz = u;
    const double norm = sqrt(dx * dx + dy * dy + dz * dz);
    const double n1 = dx / norm, n2 = dy / norm, n3 = dz / norm;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        y[mu] = F.obs[mu];
        y[4 + mu] = F.tet[mu] + n1 * F.tet[4 + mu] + n2 * F.tet[8 + mu] + n3 * F.tet[12 + mu];
    }
}

// ---- sky (sky.rs) -----------------------------------------------------------
#define GEON_TAU 6.283185307179586
#define GEON_PI 3.141592653589793

GEON_D long long f2i64(double v) {
    if (__builtin_isnan(v)) return 0;
    if (v >= 9223372036854775807.0) return 9223372036854775807LL;
    if (v <= -9223372036854775808.0) return (-9223372036854775807LL - 1);
    return (long long)v;
}
GEON_D double rem_euclid(double a, double b) {
    const double r = fmod(a, b);
    return (r < 0.0) ? r + fabs(b) : r;
}
GEON_D long long rem_euclid_i(long long a, long long b) {
    const long long r = a % b;
    return (r < 0) ? r + b : r;
}
GEON_D double clampd(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
GEON_D unsigned long long sky_hash1(unsigned long long x) {
    x *= 0x9E3779B97F4A7C15ull;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}
GEON_D void sky_sample(const double* dir, bool show_stars, float* rgb) {
    const double x = dir[0], y = dir[1], z = dir[2];
    const double r = fmax(sqrt(x * x + y * y + z * z), 1e-30);
    const double theta = acos(clampd(z / r, -1.0, 1.0));
    const double phi = atan2(y, x);
    const double phi_shift = rem_euclid(phi + 0.37, GEON_TAU);
    const double theta_shift = clampd(theta + 0.21, 0.0, GEON_PI);
    // sectors (sky.rs:72-104)
    const double pal[6][3] = {{0.15, 0.08, 0.20}, {0.08, 0.15, 0.22}, {0.20, 0.15, 0.08},
                              {0.10, 0.18, 0.10}, {0.22, 0.10, 0.10}, {0.10, 0.10, 0.22}};
    const long long phi_mod = f2i64(floor(phi_shift / GEON_TAU * 6.0));
    const long long band = f2i64(floor(theta_shift / GEON_PI * 3.0));
    const int sec = (int)rem_euclid_i(phi_mod, 6);
    const double band_mod = (rem_euclid_i(band, 2) == 0) ? 1.0 : 0.6;
    const long long ic = f2i64(floor(phi_shift / GEON_TAU * 24.0));
    const long long jc = f2i64(floor(theta_shift / GEON_PI * 12.0));
    const double check = (rem_euclid_i(ic + jc, 2) == 0) ? 1.0 : 0.55;
    float s[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c] = (float)(pal[sec][c] * band_mod * check);
    // stars (sky.rs:43-68)
    float st[3] = {0.0f, 0.0f, 0.0f};
    if (show_stars) {
        const long long i = f2i64(floor(phi_shift / GEON_TAU * 512.0));
        const long long j = f2i64(floor(theta_shift / GEON_PI * 256.0));
        const unsigned long long h =
            sky_hash1(((unsigned long long)i * 73856093ull) ^ ((unsigned long long)j * 19349663ull));
        const double prob = (double)(h & 0xFFFF) / 65535.0;
        if (prob < 0.004) {
            const double br = 0.4 + 0.6 * ((double)((h >> 16) & 0xFF) / 255.0);
            const double temp = (double)((h >> 24) & 0xFF) / 255.0;
            st[0] = (float)(br * (1.0 - 0.3 * temp));
            st[1] = (float)(br * (1.0 - 0.1 * fabs(temp - 0.5)));
            st[2] = (float)(br * (0.7 + 0.3 * temp));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = fminf(s[c] + st[c], 1.0f);
}

}  // namespace geon
