// geon_curv.hpp — curvature (curvature.rs:25-141) of Alcubierre points by a
// team of lanes whose shared metric stencil and Γ stencil are staged
// through LDS: 8-lane teams for Riemann (CurvStencil, below), quads for the
// Ricci family (AlcQuad, at the end).
//
// A team of 8 lanes serves one point; a 64-lane block serves 8 points. The
// central differences of Γ need Γ at the point and at x ± h·e_d for every
// coordinate d that Γ depends on (curvature.rs:34-44; along the others the
// two evaluations are identical and the difference is exactly zero): 7
// evaluations of the generic FD Γ (metric/mod.rs:25-69), each of which
// evaluates the metric at its point and at ± h along x, y and z — 49 metric
// evaluations, each a square root and two tanh. Only 31 of them are
// distinct, bit for bit: the point x + s1·h e_d + s2·h e_a (d ≠ a) is formed
// by the same two roundings whether it is reached from x + s1·h e_d or from
// x + s2·h e_a, and x ± h e_d serves both as a shift of x and as the centre
// of stencil point ± d. The team therefore
//   A. evaluates the shape function f (alcubierre.rs:33-44) at the 31
//      distinct points, 4 rounds of 8 lanes, into LDS (on the GPU from the
//      expansion of its two tanh about the point, Alcubierre::shape_exp);
//   B. forms, on lane t < 7, Γ at stencil point t from those f values with
//      the generic formula (g from f, g⁻¹, ∂g by central differences, Γ) and
//      writes its 30 nonzero entries to LDS;
//   C. forms on each lane its share of the outputs from LDS — two Ricci
//      components, or 32 Riemann components — then R, G and T:
//   ∂_d Γ = (Γ(x + h e_d) − Γ(x − h e_d)) · (1/(2h))
//   R^ρ_σμν = ((∂_μΓ^ρ_νσ − ∂_νΓ^ρ_μσ) + Σ_λ Γ^ρ_μλΓ^λ_νσ) − Σ_λ Γ^ρ_νλΓ^λ_μσ
//   R_σν = Σ_α R^α_σαν,  R = Σ g^μν R_μν,  G = R_μν − ½ g_μν R,  T = G/(8π)
// Every value is formed by the same operations in the same order as one
// lane evaluating the reference formulas point by point, so the results are
// identical to it. Structural zeros are skipped with the metric's masks;
// adding an exact zero never changes an IEEE sum.
//
// The phases are plain functions of the lane index and the team's LDS rows,
// so the same code runs on the host (one lane after another) for checking;
// the kernel in geon_kernels.hip places the barriers between them.
#pragma once
#include "geon_device.hpp"

namespace geon {

#define GEON_CURV_TEAM 8     // lanes per point
#define GEON_CURV_PTS 8      // points per 64-lane block

template <class MP>
struct CurvStencil {
    // 64-bit mask of the structurally nonzero Γ^μ_αβ, index μ·16 + α·4 + β
    static constexpr unsigned long long nz_mask() {
        unsigned long long m = 0;
        for (int i = 0; i < 64; ++i)
            if (MP::gam_nz(i / 16, (i / 4) % 4, i % 4)) m |= 1ull << i;
        return m;
    }
    static constexpr unsigned long long kNZ = nz_mask();
    static constexpr int kNNZ = __builtin_popcountll(kNZ);   // LDS slots per Γ
    static constexpr int n_dirs() {
        int c = 0;
        for (int d = 0; d < 4; ++d) c += MP::dep(d) ? 1 : 0;
        return c;
    }
    static constexpr int kPoints = 1 + 2 * n_dirs();   // Γ stencil points
    // LDS row of one point: kPoints Γ of kNNZ doubles, plus one so that the
    // rows of the 8 points of a block start in different banks
    static constexpr int kRow = kPoints * kNNZ + 1;
    // j-th dependent coordinate, and the position of coordinate d among them
    static constexpr int dir(int j) {
        for (int d = 0; d < 4; ++d)
            if (MP::dep(d) && j-- == 0) return d;
        return -1;
    }
    static constexpr int dir_index(int d) {
        int j = 0;
        for (int e = 0; e < d; ++e) j += MP::dep(e) ? 1 : 0;
        return j;
    }
    GEON_D static bool nz(int i) { return (kNZ >> i) & 1ull; }
    GEON_D static int slot(int i) { return __builtin_popcountll(kNZ & ((1ull << i) - 1ull)); }
};

// ---- the Alcubierre metric stencil -----------------------------------------
// Slots of the 31 distinct metric points of one curvature point x (m = 0 for
// +h, 1 for −h; coordinates d, a ∈ {1, 2, 3}):
//    0                           x
//    1 + 2(d−1) + m              x_d ± h                 (= Γ stencil point t)
//    7 + 4(d−1) + 2·m1 + m2      (x_d ± h) ± h
//   19 + 4p + 2·m_lo + m_hi      x_lo ± h and x_hi ± h   (p = 0, 1, 2 for the
//                                                         pairs (1,2), (1,3), (2,3))
struct AlcStencil {
    static constexpr int kUsed = 31;
    static constexpr int kRow = 33;   // LDS row per point (bank spread)

    // slot of the metric point that Γ stencil point t (0 = x, 1 + 2(d−1) + m
    // = x_d ± h) evaluates at its own ± h along coordinate a
    GEON_D static int shift_slot(int t, int a, int m) {
        if (t == 0) return 1 + 2 * (a - 1) + m;
        const int d = 1 + (t - 1) / 2, m1 = (t - 1) & 1;
        if (a == d) return 7 + 4 * (d - 1) + 2 * m1 + m;
        const int lo = d < a ? d : a, hi = d < a ? a : d;
        const int mlo = d < a ? m1 : m, mhi = d < a ? m : m1;
        return 19 + 4 * (lo + hi - 3) + 2 * mlo + mhi;
    }

    // x with the shifts of slot j applied in order: the first to coordinate
    // c1, then the second to c2 (c = 0 means none; x_t is never shifted).
    // x − h is x + (−h) in IEEE arithmetic.
    GEON_D static void point(const geon_dmetric& P, const double* x, int j, double* xe) {
        int c1 = 0, c2 = 0, m1 = 0, m2 = 0;
        if (j >= 1 && j <= 6) {
            c1 = 1 + (j - 1) / 2;
            m1 = (j - 1) & 1;
        } else if (j >= 7 && j <= 18) {
            const int k = j - 7;
            c1 = c2 = 1 + k / 4;
            m1 = (k >> 1) & 1;
            m2 = k & 1;
        } else if (j >= 19 && j <= 30) {
            const int k = j - 19, p = k / 4;
            c1 = p == 2 ? 2 : 1;
            c2 = p == 0 ? 2 : 3;
            m1 = (k >> 1) & 1;
            m2 = k & 1;
        }
        const double h1 = m1 ? -P.fd_h : P.fd_h, h2 = m2 ? -P.fd_h : P.fd_h;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = x[c];
            if (c == c1 && c1 != 0) v = v + h1;
            if (c == c2 && c2 != 0) v = v + h2;
            xe[c] = v;
        }
    }
};

// phase A: the shape value f at metric-stencil slot j of point x; with E
// (the expansion about x, Alcubierre::shape_exp), from the expansion
GEON_D double curv_alc_point_f(const geon_dmetric& P, const double* x, int j,
                               const Alcubierre::ShapeExp* E = nullptr) {
    double xe[4];
    AlcStencil::point(P, x, j, xe);
    if (E) return Alcubierre::shape_near_sq(P, *E, Alcubierre::rs2(P, xe));
    return Alcubierre::shape<true>(P, Alcubierre::rs(P, xe));
}
// Whether this build's phase A expands the shape function about the point:
// on the GPU where the stencil allows it (P.shape_exp); the host build
// evaluates every tanh, as the reference, unless GEON_HOST_DEVICE_MATH asks
// for the GPU's arithmetic.
GEON_D bool curv_shape_exp(const geon_dmetric& P) {
#if defined(__HIP_DEVICE_COMPILE__) || defined(GEON_HOST_DEVICE_MATH)
    return P.shape_exp != 0;
#else
    (void)P;
    return false;
#endif
}

// phase B: Γ at Γ-stencil point t (< 7) from the staged f values, as
// FD<Alcubierre>::christoffel forms it at that point; the nonzero entries
// are written to gam_row[0 .. kNNZ).
GEON_D void curv_alc_gamma(const geon_dmetric& P, const double* sf, int t, double* gam_row) {
    using CS = CurvStencil<Alcubierre>;
    double g0[16], gi[16], dgv[64], gam[64];
    Alcubierre::g_of_f(P, sf[t], g0);   // the centre of stencil point t is slot t
    Alcubierre::ginv(P, g0, gi);
    sfor<4>([&](auto A) {
        constexpr int a = CI(A);
        if constexpr (FD<Alcubierre>::dir_nz(a)) {
            double gp[16], gm[16];
            Alcubierre::g_of_f(P, sf[AlcStencil::shift_slot(t, a, 0)], gp);
            Alcubierre::g_of_f(P, sf[AlcStencil::shift_slot(t, a, 1)], gm);
            FD<Alcubierre>::dg_dir<a>(P, gp, gm, dgv);
        }
    });
    FD<Alcubierre>::assemble(gi, dgv, gam);
    sfor<64>([&](auto I) {
        constexpr int i = CI(I);
        if constexpr ((CS::kNZ >> i) & 1ull)
            gam_row[__builtin_popcountll(CS::kNZ & ((1ull << i) - 1ull))] = gam[i];
    });
}

// ∂_d Γ[i] and Γ[i] at x from a point's LDS row
template <class MP>
GEON_D double curv_dgam(const geon_dmetric& P, const double* row, int d, int i) {
    using CS = CurvStencil<MP>;
    const int j = CS::dir_index(d), s = CS::slot(i);
    return (row[(1 + 2 * j) * CS::kNNZ + s] - row[(2 + 2 * j) * CS::kNNZ + s]) * P.fd_inv_2h;
}
template <class MP>
GEON_D double curv_gam0(const double* row, int i) {
    return row[CurvStencil<MP>::slot(i)];
}

// phase C, Riemann: components t, t + 8, … (32 of the 256) of lane t
template <class MP>
GEON_D void curv_riemann_lane(const geon_dmetric& P, const double* row, int t, double* out256) {
    using CS = CurvStencil<MP>;
    for (int c = t; c < 256; c += GEON_CURV_TEAM) {
        const int rho = c / 64, sg = (c / 16) % 4, mu = (c / 4) % 4, nu = c % 4;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        if (MP::dep(mu) && CS::nz(rho * 16 + nu * 4 + sg)) t1 = curv_dgam<MP>(P, row, mu, rho * 16 + nu * 4 + sg);
        if (MP::dep(nu) && CS::nz(rho * 16 + mu * 4 + sg)) t2 = curv_dgam<MP>(P, row, nu, rho * 16 + mu * 4 + sg);
        for (int l = 0; l < 4; ++l) {
            if (CS::nz(rho * 16 + mu * 4 + l) && CS::nz(l * 16 + nu * 4 + sg))
                t3 += curv_gam0<MP>(row, rho * 16 + mu * 4 + l) * curv_gam0<MP>(row, l * 16 + nu * 4 + sg);
            if (CS::nz(rho * 16 + nu * 4 + l) && CS::nz(l * 16 + mu * 4 + sg))
                t4 += curv_gam0<MP>(row, rho * 16 + nu * 4 + l) * curv_gam0<MP>(row, l * 16 + mu * 4 + sg);
        }
        out256[c] = t1 - t2 + t3 - t4;
    }
}

// phase C, Ricci R_σν = Σ_α R^α_σαν: components t and t + 8 of lane t
template <class MP>
GEON_D void curv_ricci_lane(const geon_dmetric& P, const double* row, int t, double* ric2) {
    using CS = CurvStencil<MP>;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = t + GEON_CURV_TEAM * h;
        const int sg = i / 4, nu = i % 4;
        double s = 0.0;
        for (int a = 0; a < 4; ++a) {
            double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
            if (MP::dep(a) && CS::nz(a * 16 + nu * 4 + sg)) t1 = curv_dgam<MP>(P, row, a, a * 16 + nu * 4 + sg);
            if (MP::dep(nu) && CS::nz(a * 16 + a * 4 + sg)) t2 = curv_dgam<MP>(P, row, nu, a * 16 + a * 4 + sg);
            for (int l = 0; l < 4; ++l) {
                if (CS::nz(a * 16 + a * 4 + l) && CS::nz(l * 16 + nu * 4 + sg))
                    t3 += curv_gam0<MP>(row, a * 16 + a * 4 + l) * curv_gam0<MP>(row, l * 16 + nu * 4 + sg);
                if (CS::nz(a * 16 + nu * 4 + l) && CS::nz(l * 16 + a * 4 + sg))
                    t4 += curv_gam0<MP>(row, a * 16 + nu * 4 + l) * curv_gam0<MP>(row, l * 16 + a * 4 + sg);
            }
            s += t1 - t2 + t3 - t4;
        }
        ric2[h] = s;
    }
}

// phase C, after all 16 Ricci components are in ric16: R = Σ g^μν R_μν in
// index order (the same in every lane), then lane t's two components of G
// (what = 2) or T = G·factor (what = 3); what = 4 gives R in e[0].
// g and g⁻¹ at x come from the staged f(x).
GEON_D void curv_alc_outputs_lane(const geon_dmetric& P, double f0, const double* ric16, int t,
                                  int32_t what, double factor, double* e) {
    double g[16], gi[16];
    Alcubierre::g_of_f(P, f0, g);
    Alcubierre::ginv(P, g, gi);
    double R = 0.0;
    sfor<16>([&](auto I) {
        constexpr int i = CI(I);
        if constexpr (Alcubierre::ginv_nz(i / 4, i % 4)) R += gi[i] * ric16[i];
    });
    if (what == 4) {
        e[0] = R;
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = t + GEON_CURV_TEAM * h;
        double v = ric16[i];
        if (Alcubierre::g_nz(i / 4, i % 4)) v = ric16[i] - 0.5 * g[i] * R;
        e[h] = (what == 3) ? v * factor : v;
    }
}


// ---- Ricci, R, G and T of Alcubierre points by quads ----------------------
// The Ricci family needs of ∂Γ only ∂_aΓ^a_νσ (term 1 of R^a_σaν) and
// ∂_νΓ^a_aσ (term 2), and of Γ only the groups Γ^μ_α· that have a nonzero
// entry. A quad of 4 lanes serves one point and a 64-lane block 16 points:
//   A. the 31 shape values f of the metric stencil in 8 rounds of the 4
//      lanes, each round one family of stencil points whose shifted axes
//      are fixed and whose lanes differ only in the signs of the shifts
//      (±x, ±y; ±z and x; (x_d ± h) ± h for d = x, y, z; x_a ± h and
//      x_b ± h for the pairs (x,y), (x,z), (y,z)), so the squares of the
//      unshifted coordinates are formed once per point;
//   B. Γ at the seven Γ-stencil points, two rounds: ±x, ±y on lanes 0-3,
//      then ±z on lanes 0-1 and x on lanes 2-3. The two lanes of a ± pair
//      swap the entries the rows need, and ∂_dΓ = (Γ(x + h e_d) −
//      Γ(x − h e_d))·(1/(2h)) is formed once and stored, with Γ(x), over
//      the shape values (read by then);
//   C. lane σ forms R_σν for ν = 0..3 with every index but σ known at
//      compile time. The nonzero entries of each stored group form a
//      prefix Γ^μ_α0 .. Γ^μ_α(L−1) and are stored alone; lane σ reads the
//      group at offset σ and takes +0 where σ ≥ L;
//   D. R = Σ g^μν R_μν from the quad's R_00, R_01, R_10, R_11, R_22, R_33,
//      then lane σ's four components of G or T.
// Every value is formed by the operations of the team kernel above in the
// same order. Where the team kernel skips a product whose Γ^l_νσ is a
// structural zero, the quad multiplies by +0 and adds ±0 to a sum that
// started at +0 and so is never −0: the sum is unchanged. The outputs are
// therefore the team kernel's bit for bit.
struct AlcQuad {
    static constexpr bool nz(int mu, int a, int b) { return Alcubierre::gam_nz(mu, a, b); }
    // number of nonzero entries of the group Γ^μ_α·
    static constexpr int len(int mu, int a) {
        return (nz(mu, a, 0) ? 1 : 0) + (nz(mu, a, 1) ? 1 : 0) + (nz(mu, a, 2) ? 1 : 0) +
               (nz(mu, a, 3) ? 1 : 0);
    }
    static constexpr bool prefixes() {
        for (int i = 0; i < 64; ++i)
            if (nz(i / 16, (i / 4) % 4, i % 4) != (i % 4 < len(i / 16, (i / 4) % 4))) return false;
        return true;
    }
    static constexpr bool grp(int mu, int a) { return len(mu, a) > 0; }
    // groups kept in row d: 0 = Γ(x) (all groups), d = 1..3 = ∂_dΓ (the
    // groups Γ^d_ν· of term 1 and Γ^a_a· of term 2)
    static constexpr bool stored(int d, int mu, int a) {
        return grp(mu, a) && (d == 0 || mu == d || mu == a);
    }
    static constexpr int row_size(int d) {
        int c = 0;
        for (int i = 0; i < 16; ++i) c += stored(d, i / 4, i % 4) ? len(i / 4, i % 4) : 0;
        return c;
    }
    static constexpr int row_off(int d) {
        int o = 0;
        for (int e = 0; e < d; ++e) o += row_size(e);
        return o;
    }
    // first double of group (μ, α) in row d
    static constexpr int off(int d, int mu, int a) {
        int o = row_off(d);
        for (int i = 0; i < mu * 4 + a; ++i) o += stored(d, i / 4, i % 4) ? len(i / 4, i % 4) : 0;
        return o;
    }
    // doubles per point in LDS: the rows, room for a read 3 past the last
    // group, rounded to an odd count so that the 16 points of a block start
    // in different banks
    static constexpr int stride() { return (row_off(4) + 3) | 1; }
    // compact slot of Γ^μ_αβ in the 30-entry rows curv_alc_gamma writes
    static constexpr int slot(int mu, int a, int b) {
        int c = 0;
        for (int i = 0; i < mu * 16 + a * 4 + b; ++i) c += nz(i / 16, (i / 4) % 4, i % 4) ? 1 : 0;
        return c;
    }
    // Γ^μ_αβ is needed in one of the ∂ rows of the set `rows` (bit d)
    static constexpr bool wanted(int rows, int mu, int a, int b) {
        return nz(mu, a, b) && ((((rows >> 1) & 1) && stored(1, mu, a)) ||
                                (((rows >> 2) & 1) && stored(2, mu, a)) ||
                                (((rows >> 3) & 1) && stored(3, mu, a)));
    }

    // phase A, round R (0..7), lane t: the shape value at the lane's metric
    // stencil point into pt[slot] (AlcStencil numbering). sq = (x − x_s)²,
    // y², z² at the point; lane 3 of round 1 repeats x (slot 0).
    template <int R>
    GEON_D static void shape_round(const geon_dmetric& P, const double* x, const double* sq, int t,
                                   double* pt, const Alcubierre::ShapeExp* E = nullptr) {
        const double h = P.fd_h;
        const double hA = (t & 2) ? -h : h, hB = (t & 1) ? -h : h;
        double s[3] = {sq[0], sq[1], sq[2]};
        int j;
        auto shifted = [&](int c, double v) {   // the square of coordinate c at v
            const double d = (c == 1) ? v - P.ship_x : v;
            return d * d;
        };
        if constexpr (R == 0) {          // x_1 ± h (t = 0, 1), x_2 ± h (t = 2, 3)
            const double v1 = shifted(1, x[1] + hB), v2 = shifted(2, x[2] + hB);
            s[0] = t < 2 ? v1 : s[0];
            s[1] = t < 2 ? s[1] : v2;
            j = 1 + t;
        } else if constexpr (R == 1) {   // x_3 ± h (t = 0, 1), x (t = 2, 3)
            const double v3 = shifted(3, x[3] + hB);
            s[2] = t < 2 ? v3 : s[2];
            j = t < 2 ? 5 + t : 0;
        } else if constexpr (R <= 4) {   // (x_d ± h) ± h
            constexpr int d = R - 1;
            s[d - 1] = shifted(d, (x[d] + hA) + hB);
            j = 7 + 4 * (d - 1) + t;
        } else {                         // x_lo ± h, x_hi ± h
            constexpr int p = R - 5, lo = p == 2 ? 2 : 1, hi = p == 0 ? 2 : 3;
            s[lo - 1] = shifted(lo, x[lo] + hA);
            s[hi - 1] = shifted(hi, x[hi] + hB);
            j = 19 + 4 * p + t;
        }
        const double S = s[0] + s[1] + s[2];
        pt[j] = E ? Alcubierre::shape_near_sq(P, *E, S) : Alcubierre::shape<true>(P, sqrt(S));
    }
    // the squares of the unshifted coordinates (Alcubierre::rs)
    GEON_D static void base_squares(const geon_dmetric& P, const double* x, double* sq) {
        const double dx = x[1] - P.ship_x, dy = x[2], dz = x[3];
        sq[0] = dx * dx;
        sq[1] = dy * dy;
        sq[2] = dz * dz;
    }
    // row 0: Γ(x) from its compact entries
    GEON_D static void row0(const double* g, double* pt) {
        sfor<64>([&](auto I) {
            constexpr int i = CI(I), mu = i / 16, a = (i / 4) % 4, b = i % 4;
            if constexpr (nz(mu, a, b)) {
                constexpr int o = off(0, mu, a) + b, sl = slot(mu, a, b);
                pt[o] = g[sl];
            }
        });
    }
    // row d: ∂_dΓ = (Γ(x + h e_d) − Γ(x − h e_d))·(1/(2h)) (curvature.rs:37-43)
    template <int d>
    GEON_D static void drow(const double* gplus, const double* gminus, double inv_2h, double* pt) {
        sfor<64>([&](auto I) {
            constexpr int i = CI(I), mu = i / 16, a = (i / 4) % 4, b = i % 4;
            if constexpr (nz(mu, a, b) && stored(d, mu, a)) {
                constexpr int o = off(d, mu, a) + b, sl = slot(mu, a, b);
                pt[o] = (gplus[sl] - gminus[sl]) * inv_2h;
            }
        });
    }
    // entry σ of the group at o with L nonzero entries: +0 past them
    template <int L>
    GEON_D static double at(const double* pt, int o, int sg) {
        const double v = pt[o + sg];
        if constexpr (L == 4) return v;
        else return sg < L ? v : 0.0;
    }
    // lane σ: R_σν, ν = 0..3 (curvature.rs:48-85 through R^α_σαν)
    GEON_D static void ricci(const double* pt, int sg, double* ric) {
        sfor<4>([&](auto Nu) {
            constexpr int nu = CI(Nu);
            double s = 0.0;
            sfor<4>([&](auto A) {
                constexpr int a = CI(A);
                double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
                if constexpr (Alcubierre::dep(a) && grp(a, nu)) {
                    constexpr int o = off(a, a, nu);
                    t1 = at<len(a, nu)>(pt, o, sg);
                }
                if constexpr (Alcubierre::dep(nu) && grp(a, a)) {
                    constexpr int o = off(nu, a, a);
                    t2 = at<len(a, a)>(pt, o, sg);
                }
                sfor<4>([&](auto L) {
                    constexpr int l = CI(L);
                    if constexpr (nz(a, a, l) && grp(l, nu)) {
                        constexpr int o1 = off(0, a, a) + l, o2 = off(0, l, nu);
                        t3 += pt[o1] * at<len(l, nu)>(pt, o2, sg);
                    }
                    if constexpr (nz(a, nu, l) && grp(l, a)) {
                        constexpr int o1 = off(0, a, nu) + l, o2 = off(0, l, a);
                        t4 += pt[o1] * at<len(l, a)>(pt, o2, sg);
                    }
                });
                s += t1 - t2 + t3 - t4;
            });
            ric[nu] = s;
        });
    }
    // lane σ's outputs from f(x), the quad's R_00, R_01, R_10, R_11, R_22,
    // R_33 (r6) and its own R_σν: e[ν] = G_σν (what 2) or G_σν·factor
    // (what 3); *R_out = R = Σ g^μν R_μν in index order.
    GEON_D static void outputs(const geon_dmetric& P, double f0, const double* r6,
                               const double* ric, int sg, int32_t what, double factor, double* e,
                               double* R_out) {
        double g[16], gi[16];
        Alcubierre::g_of_f(P, f0, g);
        Alcubierre::ginv(P, g, gi);
        double R = 0.0;
        R += gi[0] * r6[0];
        R += gi[1] * r6[1];
        R += gi[4] * r6[2];
        R += gi[5] * r6[3];
        R += gi[10] * r6[4];
        R += gi[15] * r6[5];
        *R_out = R;
        // g_σν for the lane's σ: g_00, g_01 = g_10, and 1 on the rest of the diagonal
        sfor<4>([&](auto Nu) {
            constexpr int nu = CI(Nu);
            double gv;
            bool gnz;
            if constexpr (nu == 0) {
                gv = sg == 0 ? g[0] : g[4];
                gnz = sg <= 1;
            } else if constexpr (nu == 1) {
                gv = sg == 0 ? g[1] : g[5];
                gnz = sg <= 1;
            } else {
                gv = g[nu * 5];
                gnz = sg == nu;
            }
            double v = ric[nu];
            if (gnz) v = ric[nu] - 0.5 * gv * R;
            e[nu] = (what == 3) ? v * factor : v;
        });
    }
};
static_assert(!Alcubierre::dep(0), "the ∂ rows are indexed by spatial directions 1..3");
static_assert(AlcQuad::prefixes(), "every Γ group's nonzero entries form a prefix");
static_assert(AlcQuad::row_off(4) >= AlcStencil::kUsed, "the shape values fit under the rows");

}  // namespace geon
