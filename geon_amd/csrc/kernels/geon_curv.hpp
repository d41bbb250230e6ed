// geon_curv.hpp — curvature (curvature.rs:25-141) of Alcubierre points by a
// team of lanes whose shared metric stencil and Γ stencil are staged
// through LDS.
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
//      distinct points, 4 rounds of 8 lanes, into LDS;
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

// phase A: the shape value f at metric-stencil slot j of point x
GEON_D double curv_alc_point_f(const geon_dmetric& P, const double* x, int j) {
    double xe[4];
    AlcStencil::point(P, x, j, xe);
    return Alcubierre::shape<true>(P, Alcubierre::rs(P, xe));
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

}  // namespace geon
