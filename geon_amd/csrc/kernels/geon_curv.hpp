// geon_curv.hpp — curvature (curvature.rs:25-141) with the Γ stencil of each
// point staged through LDS.
//
// A team of 8 lanes serves one point; a 64-lane block serves 8 points. The
// central differences of Γ need Γ at the point and at x ± h·e_d for every
// coordinate d that Γ depends on (curvature.rs:34-44; along the others the
// two evaluations are identical and the difference is exactly zero): 7
// evaluations for Alcubierre (5 for Schwarzschild and Morris–Thorne, whose
// cheap Γ is better left to one lane per point). Lane t of the team
// evaluates stencil point t and writes the nonzero entries of Γ to LDS
// (compact: 30 entries for Alcubierre). After one barrier each lane forms
// its share of the outputs from LDS — two Ricci components, or 32 Riemann
// components — with the same operations in the same order as a single lane
// evaluating the reference formulas:
//   ∂_d Γ = (Γ(x + h e_d) − Γ(x − h e_d)) · (1/(2h))
//   R^ρ_σμν = ((∂_μΓ^ρ_νσ − ∂_νΓ^ρ_μσ) + Σ_λ Γ^ρ_μλΓ^λ_νσ) − Σ_λ Γ^ρ_νλΓ^λ_μσ
//   R_σν = Σ_α R^α_σαν,  R = Σ g^μν R_μν,  G = R_μν − ½ g_μν R,  T = G/(8π)
// Structural zeros are skipped with the metric's masks; adding an exact
// zero never changes an IEEE sum.
#pragma once
#include "geon_device.hpp"

namespace geon {

#define GEON_CURV_TEAM 8     // lanes per point
#define GEON_CURV_PTS 8      // points per 64-lane block
#define GEON_CURV_NZ 32      // LDS slots per Γ (the largest metric needs 30)

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
    static constexpr int n_dirs() {
        int c = 0;
        for (int d = 0; d < 4; ++d) c += MP::dep(d) ? 1 : 0;
        return c;
    }
    static constexpr int kPoints = 1 + 2 * n_dirs();   // stencil points
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

template <class MP>
GEON_D void curvature_team(const geon_dmetric& P, uint64_t n, const double* __restrict__ xs,
                           int32_t what, double* __restrict__ out, double factor) {
    using CS = CurvStencil<MP>;
    static_assert(CS::kPoints <= GEON_CURV_TEAM, "stencil larger than the team");
    static_assert(__builtin_popcountll(CS::kNZ) <= GEON_CURV_NZ, "too many nonzero Γ");
    __shared__ double s_gam[GEON_CURV_PTS][CS::kPoints][GEON_CURV_NZ];
    __shared__ double s_ric[GEON_CURV_PTS][16];

    const uint32_t t = threadIdx.x % GEON_CURV_TEAM;
    const uint32_t q = threadIdx.x / GEON_CURV_TEAM;
    const uint64_t k = (uint64_t)blockIdx.x * GEON_CURV_PTS + q;
    const bool valid = k < n;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
        x[0] = xs[k];
        x[1] = xs[n + k];
        x[2] = xs[2 * n + k];
        x[3] = xs[3 * n + k];
    }

    // ---- stencil: lane t evaluates Γ at stencil point t ----
    if (valid && t < (uint32_t)CS::kPoints) {
        // x_plus[d] += h (odd t) or x_minus[d] -= h (even t > 0); the other
        // coordinates are copied untouched
        const int d = t > 0 ? CS::dir((int)(t - 1) / 2) : -1;
        double xe[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            xe[c] = (c != d) ? x[c] : ((t & 1u) ? x[c] + P.fd_h : x[c] - P.fd_h);
        double gam[64];
        MP::christoffel(P, xe, gam);
        sfor<64>([&](auto I) {
            constexpr int i = CI(I);
            if constexpr ((CS::kNZ >> i) & 1ull)
                s_gam[q][t][__builtin_popcountll(CS::kNZ & ((1ull << i) - 1ull))] = gam[i];
        });
    }
    __syncthreads();

    const double* G0 = s_gam[q][0];
    auto gam0 = [&](int i) { return G0[CS::slot(i)]; };
    // ∂_d Γ[i] for a dependent coordinate d
    auto dgam = [&](int d, int i) {
        const int j = CS::dir_index(d);
        return (s_gam[q][1 + 2 * j][CS::slot(i)] - s_gam[q][2 + 2 * j][CS::slot(i)]) * P.fd_inv_2h;
    };

    if (what == 0) {   // full Riemann: 32 components per lane
        if (!valid) return;
        for (int c = (int)t; c < 256; c += GEON_CURV_TEAM) {
            const int rho = c / 64, sg = (c / 16) % 4, mu = (c / 4) % 4, nu = c % 4;
            double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
            if (MP::dep(mu) && CS::nz(rho * 16 + nu * 4 + sg)) t1 = dgam(mu, rho * 16 + nu * 4 + sg);
            if (MP::dep(nu) && CS::nz(rho * 16 + mu * 4 + sg)) t2 = dgam(nu, rho * 16 + mu * 4 + sg);
            for (int l = 0; l < 4; ++l) {
                if (CS::nz(rho * 16 + mu * 4 + l) && CS::nz(l * 16 + nu * 4 + sg))
                    t3 += gam0(rho * 16 + mu * 4 + l) * gam0(l * 16 + nu * 4 + sg);
                if (CS::nz(rho * 16 + nu * 4 + l) && CS::nz(l * 16 + mu * 4 + sg))
                    t4 += gam0(rho * 16 + nu * 4 + l) * gam0(l * 16 + mu * 4 + sg);
            }
            out[k * 256 + c] = t1 - t2 + t3 - t4;
        }
        return;
    }

    // ---- Ricci R_σν = Σ_α R^α_σαν: components t and t + 8 of this lane ----
    double ric[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = (int)t + GEON_CURV_TEAM * h;
        const int sg = i / 4, nu = i % 4;
        double s = 0.0;
        for (int a = 0; a < 4; ++a) {
            double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
            if (MP::dep(a) && CS::nz(a * 16 + nu * 4 + sg)) t1 = dgam(a, a * 16 + nu * 4 + sg);
            if (MP::dep(nu) && CS::nz(a * 16 + a * 4 + sg)) t2 = dgam(nu, a * 16 + a * 4 + sg);
            for (int l = 0; l < 4; ++l) {
                if (CS::nz(a * 16 + a * 4 + l) && CS::nz(l * 16 + nu * 4 + sg))
                    t3 += gam0(a * 16 + a * 4 + l) * gam0(l * 16 + nu * 4 + sg);
                if (CS::nz(a * 16 + nu * 4 + l) && CS::nz(l * 16 + a * 4 + sg))
                    t4 += gam0(a * 16 + nu * 4 + l) * gam0(l * 16 + a * 4 + sg);
            }
            s += t1 - t2 + t3 - t4;
        }
        ric[h] = s;
        s_ric[q][i] = s;
    }
    if (what == 1) {
        if (valid) {
            out[k * 16 + t] = ric[0];
            out[k * 16 + t + GEON_CURV_TEAM] = ric[1];
        }
        return;
    }
    __syncthreads();   // all 16 components of every point are in LDS
    if (!valid) return;
    double gi[16];
    MP::ginv_at(P, x, gi);
    double R = 0.0;   // R = Σ g^μν R_μν in index order, the same in every lane
    sfor<16>([&](auto I) {
        constexpr int i = CI(I);
        if constexpr (MP::ginv_nz(i / 4, i % 4)) R += gi[i] * s_ric[q][i];
    });
    if (what == 4) {
        if (t == 0) out[k] = R;
        return;
    }
    double g[16];
    MP::g(P, x, g);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = (int)t + GEON_CURV_TEAM * h;
        double e = ric[h];
        if (MP::g_nz(i / 4, i % 4)) e = ric[h] - 0.5 * g[i] * R;
        out[k * 16 + i] = (what == 3) ? e * factor : e;
    }
}

}  // namespace geon
