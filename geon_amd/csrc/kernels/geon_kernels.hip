// geon_kernels.hip — gfx950 kernels of the geodesic path, built as a code
// object (hipcc --genco) and loaded by the host runtime with hipModuleLoad.
//
// Frame parameters (metric constants, camera tetrad, step schedule) live in
// a ring of __constant__ slots written by the host before each launch; the
// slot index is a kernel argument. Ray data is structure-of-arrays: plane k
// of an n-ray array starts at base + k·n, so consecutive lanes touch
// consecutive addresses.
#include "geon_device.hpp"
#include "geon_curv.hpp"
#include "geon_sched.hpp"

using namespace geon;

extern "C" {
__constant__ geon_dframe c_frames[GEON_CONST_SLOTS];
}

// Σ over the wave's lanes of v, added to *p by one lane (every lane of the
// wave must be here).
__device__ void wave_add_u64(unsigned long long* p, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v != 0ull) atomicAdd(p, v);
}

// ---- render: one lane per (pixel, sub-sample) ray -------------------------
// ray r → pixel r / spa², sub-sample r % spa² (render.rs:89-92), so the
// sub-rays of one pixel share a wave and their trajectories stay coherent.
// stats->gamma_steps receives the launch's steps that evaluated Γ.
template <class MP>
__device__ void render_rays(uint32_t slot, uint64_t n_rays, int32_t* __restrict__ cls,
                            uint32_t* __restrict__ steps, double* __restrict__ resid,
                            double* __restrict__ dir, geon_dstats* __restrict__ stats) {
    const geon_dframe& F = c_frames[slot];
    const uint64_t ray = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ng = 0;
    if (ray < n_rays) {
        const uint32_t spa = F.spa, spa2 = spa * spa;
        const uint64_t pix = ray / spa2;
        const uint32_t sub = (uint32_t)(ray - pix * spa2);
        const uint32_t i = (uint32_t)(pix % F.W), j = (uint32_t)(pix / F.W);
        const double du = ((double)(sub % spa) + 0.5) * F.inv_spa;
        const double dv = ((double)(sub / spa) + 0.5) * F.inv_spa;
        double y[8];
        camera_ray(F, i, j, du, dv, y);
        uint32_t n;
        const int c = trace<MP>(F.m, y, F.step, F.max_steps, &n, &ng);
        cls[ray] = c;
        steps[ray] = n;
        resid[ray] = null_residual<MP>(F.m, y);
        double d[3] = {0.0, 0.0, 0.0};
        if (c == ESCAPED) direction<MP>(y, d);
        dir[ray] = d[0];
        dir[n_rays + ray] = d[1];
        dir[2 * n_rays + ray] = d[2];
    }
    if (stats) wave_add_u64(&stats->gamma_steps, ng);
}

// ---- trace from caller-supplied initial states ----------------------------
template <class MP>
__device__ void trace_states(uint32_t slot, uint64_t n, const double* __restrict__ y0,
                             int32_t* __restrict__ cls, uint32_t* __restrict__ steps,
                             double* __restrict__ yf, double* __restrict__ resid,
                             double* __restrict__ dir) {
    const geon_dframe& F = c_frames[slot];
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double y[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) y[c] = y0[c * n + k];
    uint32_t st;
    const int c = trace<MP>(F.m, y, F.step, F.max_steps, &st);
    if (cls) cls[k] = c;
    if (steps) steps[k] = st;
    if (yf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) yf[q * n + k] = y[q];
    }
    if (resid) resid[k] = null_residual<MP>(F.m, y);
    if (dir) {
        double d[3] = {0.0, 0.0, 0.0};
        if (c == ESCAPED) direction<MP>(y, d);
        dir[k] = d[0];
        dir[n + k] = d[1];
        dir[2 * n + k] = d[2];
    }
}

// trace_states with adaptive steps (trace_adaptive): h0 = the frame's step,
// max_steps counts accepted steps; `rejected` receives the rejected attempts.
template <class MP>
__device__ void trace_states_adaptive(uint32_t slot, uint64_t n, const double* __restrict__ y0,
                                      double rtol, double atol, int32_t* __restrict__ cls,
                                      uint32_t* __restrict__ steps, uint32_t* __restrict__ rejected,
                                      double* __restrict__ yf, double* __restrict__ dir) {
    const geon_dframe& F = c_frames[slot];
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double y[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) y[c] = y0[c * n + k];
    uint32_t st, rj;
    const int c = trace_adaptive<MP>(F.m, y, F.step, rtol, atol, F.max_steps, &st, &rj);
    if (cls) cls[k] = c;
    if (steps) steps[k] = st;
    if (rejected) rejected[k] = rj;
    if (yf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) yf[q * n + k] = y[q];
    }
    if (dir) {
        double d[3] = {0.0, 0.0, 0.0};
        if (c == ESCAPED) direction<MP>(y, d);
        dir[k] = d[0];
        dir[n + k] = d[1];
        dir[2 * n + k] = d[2];
    }
}

// WAVES: the minimum waves per SIMD the register allocation must allow.
// Alcubierre's full step (seven metric evaluations per stage) otherwise takes
// over 256 VGPRs, one wave per SIMD; at four it spills four.
#define GEON_KERNELS(NAME, MP, WAVES)                                                           \
    extern "C" __global__ __launch_bounds__(256, WAVES) void geon_render_rays_##NAME(           \
        uint32_t slot, uint64_t n_rays, int32_t* cls, uint32_t* steps, double* resid,          \
        double* dir, geon_dstats* stats) {                                                      \
        render_rays<MP>(slot, n_rays, cls, steps, resid, dir, stats);                           \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256, WAVES) void geon_trace_states_##NAME(          \
        uint32_t slot, uint64_t n, const double* y0, int32_t* cls, uint32_t* steps,            \
        double* yf, double* resid, double* dir) {                                               \
        trace_states<MP>(slot, n, y0, cls, steps, yf, resid, dir);                              \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256, WAVES) void geon_trace_adaptive_##NAME(        \
        uint32_t slot, uint64_t n, const double* y0, double rtol, double atol, int32_t* cls,   \
        uint32_t* steps, uint32_t* rejected, double* yf, double* dir) {                         \
        trace_states_adaptive<MP>(slot, n, y0, rtol, atol, cls, steps, rejected, yf, dir);       \
    }

GEON_KERNELS(mink, Minkowski, 1)
GEON_KERNELS(schw, Schwarzschild, 1)
GEON_KERNELS(alc, Alcubierre, 4)
GEON_KERNELS(mt, MorrisThorne, 1)

// ---- per-pixel reduction, sky colour and frame statistics ------------------
// render.rs:96-163: colour per sub-ray, f32 average, majority class with
// ties captured > diverged > escaped, integer per-pixel step average, max
// finite residual; statistics summed over pixels.
__device__ void finalize_pixels(const geon_dframe& F, uint64_t n_pix, const int32_t* __restrict__ cls,
                                const uint32_t* __restrict__ steps, const double* __restrict__ resid,
                                const double* __restrict__ dir, float* __restrict__ rgb,
                                int32_t* __restrict__ cls_pix, double* __restrict__ dir_pix,
                                geon_dstats* __restrict__ stats) {
    const uint64_t pix = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t spa2 = F.spa * F.spa;
    const uint64_t n_rays = n_pix * spa2;
    unsigned long long l_cap = 0, l_esc = 0, l_div = 0, l_max = 0, l_rs = 0, l_sp = 0;
    unsigned long long l_resid = 0;
    if (pix < n_pix) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
        uint32_t nc = 0, ne = 0, nd = 0, nmax = 0;
        unsigned long long total = 0;
        double mres = 0.0;
        double fd[3] = {0.0, 0.0, 0.0};
        bool have = false;
        for (uint32_t s = 0; s < spa2; ++s) {
            const uint64_t r = pix * spa2 + s;
            const int c = cls[r];
            float col[3];
            if (c == CAPTURED) {
                ++nc;
                col[0] = F.captured_rgb[0]; col[1] = F.captured_rgb[1]; col[2] = F.captured_rgb[2];
            } else if (c == ESCAPED) {
                ++ne;
                const double d[3] = {dir[r], dir[n_rays + r], dir[2 * n_rays + r]};
                if (!have) { fd[0] = d[0]; fd[1] = d[1]; fd[2] = d[2]; have = true; }
                sky_sample(d, F.show_stars != 0, col);
            } else {
                ++nd;
                if (c == MAXSTEPS) ++nmax;
                col[0] = F.diverged_rgb[0]; col[1] = F.diverged_rgb[1]; col[2] = F.diverged_rgb[2];
            }
            acc[0] += col[0]; acc[1] += col[1]; acc[2] += col[2];
            total += steps[r];
            const double rr = resid[r];
            if (__builtin_isfinite(rr) && rr > mres) mres = rr;
        }
        const float nf = (float)spa2;
        if (rgb) {
            rgb[3 * pix] = acc[0] / nf;
            rgb[3 * pix + 1] = acc[1] / nf;
            rgb[3 * pix + 2] = acc[2] / nf;
        }
        int term;
        if (nd > 0 && nd >= nc && nd >= ne) term = (nmax == nd) ? MAXSTEPS : DIVERGED;
        else if (nc >= ne) term = CAPTURED;
        else term = ESCAPED;
        if (cls_pix) cls_pix[pix] = term;
        if (dir_pix) {
            dir_pix[pix] = fd[0];
            dir_pix[n_pix + pix] = fd[1];
            dir_pix[2 * n_pix + pix] = fd[2];
        }
        l_cap = term == CAPTURED;
        l_esc = term == ESCAPED;
        l_div = (term == DIVERGED || term == MAXSTEPS);
        l_max = nmax;
        l_rs = total;
        l_sp = total / spa2;
        l_resid = (unsigned long long)__double_as_longlong(mres);
    }
    // block reduction, then one atomic per counter per block
    __shared__ unsigned long long red[7][256];
    const int t = threadIdx.x;
    red[0][t] = l_cap; red[1][t] = l_esc; red[2][t] = l_div; red[3][t] = l_max;
    red[4][t] = l_rs; red[5][t] = l_sp; red[6][t] = l_resid;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int q = 0; q < 6; ++q) red[q][t] += red[q][t + w];
            red[6][t] = red[6][t] > red[6][t + w] ? red[6][t] : red[6][t + w];
        }
        __syncthreads();
    }
    if (t == 0 && stats) {
        atomicAdd(&stats->captured, red[0][0]);
        atomicAdd(&stats->escaped, red[1][0]);
        atomicAdd(&stats->diverged, red[2][0]);
        atomicAdd(&stats->max_steps_hit, red[3][0]);
        atomicAdd(&stats->ray_steps, red[4][0]);
        atomicAdd(&stats->steps_pix_sum, red[5][0]);
        atomicMax(&stats->max_resid_bits, red[6][0]);
    }
}

extern "C" __global__ __launch_bounds__(256) void geon_render_finalize(
    uint32_t slot, uint64_t n_pix, const int32_t* __restrict__ cls, const uint32_t* __restrict__ steps,
    const double* __restrict__ resid, const double* __restrict__ dir, float* __restrict__ rgb,
    int32_t* __restrict__ cls_pix, double* __restrict__ dir_pix, geon_dstats* __restrict__ stats) {
    finalize_pixels(c_frames[slot], n_pix, cls, steps, resid, dir, rgb, cls_pix, dir_pix, stats);
}

// the same reduction for a scheduled camera job (launched with 256 threads)
extern "C" __global__ __launch_bounds__(256) void geon_job_finalize(const geon_djob* __restrict__ J) {
    const uint64_t n_pix = (uint64_t)J->F.W * J->F.H;
    finalize_pixels(J->F, n_pix, J->cls, J->steps, J->resid, J->dir, J->rgb, J->cls_pix, J->dir_pix,
                    J->stats);
}

// T_μν k^μ k^ν at each of n points from their T (n × 16, row-major per
// point), summed over μ then ν from 0.0 as (T_μν·k^μ)·k^ν
// (curvature.rs:127-136).
extern "C" __global__ void geon_nec_kk(const double* __restrict__ T, uint64_t n, double k0, double k1,
                                       double k2, double k3, double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double k[4] = {k0, k1, k2, k3};
    const double* t = T + i * 16;
    double s = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) s += t[mu * 4 + nu] * k[mu] * k[nu];
    out[i] = s;
}

// Zero n 64-bit words (counters, cursors, statistics) without a memset node.
extern "C" __global__ void geon_zero_u64(unsigned long long* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Fill n 64-bit words with v (empty slot markers).
extern "C" __global__ void geon_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Between serving quanta: a queue cursor that ran past the claimable end is
// pulled back to it (bound read from `bound_ptr` when given, else `bound`).
extern "C" __global__ void geon_clamp_cursor(unsigned long long* cursor, unsigned long long bound,
                                             const unsigned long long* bound_ptr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const unsigned long long b = bound_ptr ? *bound_ptr : bound;
        if (*cursor > b) *cursor = b;
    }
}

// Zero a statistics block (used instead of a memset node so the sequence
// stays valid inside captured graphs).
extern "C" __global__ void geon_stats_clear(geon_dstats* stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        stats->captured = stats->escaped = stats->diverged = 0;
        stats->max_resid_bits = stats->steps_pix_sum = stats->max_steps_hit = 0;
        stats->ray_steps = stats->fp64_rays = stats->gamma_steps = 0;
    }
}

// ---- curvature (curvature.rs:25-141) ----------------------------------------
// One lane per point. ∂_d Γ by central differences of Γ (times 1/(2h),
// curvature.rs:34-44) only along coordinates Γ depends on; the other
// directions difference identical evaluations and are exactly zero.
// Riemann R^ρ_σμν = ((∂_μΓ^ρ_νσ − ∂_νΓ^ρ_μσ) + Σ_λ Γ^ρ_μλΓ^λ_νσ) − Σ_λ Γ^ρ_νλΓ^λ_μσ
// (curvature.rs:48-70); Ricci R_σν = Σ_α R^α_σαν; R = Σ g^μν R_μν;
// G = R_μν − ½ g_μν R; T = G·(1/(8π)). Structural zeros are skipped with
// the masks of the metric policy, which leaves every sum unchanged.
template <class MP>
__device__ void curvature_point(uint32_t slot, uint64_t n, const double* __restrict__ xs,
                                int32_t what, double* __restrict__ out,
                                double* __restrict__ scratch, double factor) {
    const geon_dframe& F = c_frames[slot];
    const geon_dmetric& P = F.m;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double x[4] = {xs[k], xs[n + k], xs[2 * n + k], xs[3 * n + k]};
    double gam[64];
    MP::christoffel(P, x, gam);
    auto nz = [](int r, int a, int b) constexpr { return MP::gam_nz(r, a, b); };
    (void)nz;

    if (what == 0) {  // full Riemann: ∂Γ through a per-point scratch area
        sfor<4>([&](auto D) {
            constexpr int d = CI(D);
            if constexpr (MP::dep(d)) {
                double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
                xp[d] += P.fd_h;
                xm[d] -= P.fd_h;
                double gp[64], gm[64];
                MP::christoffel(P, xp, gp);
                MP::christoffel(P, xm, gm);
                sfor<64>([&](auto I) {
                    constexpr int i = CI(I);
                    if constexpr (MP::gam_nz(i / 16, (i / 4) % 4, i % 4))
                        scratch[(uint64_t)(d * 64 + i) * n + k] = (gp[i] - gm[i]) * P.fd_inv_2h;
                });
            }
        });
        for (int rho = 0; rho < 4; ++rho)
            for (int sg = 0; sg < 4; ++sg)
                for (int mu = 0; mu < 4; ++mu)
                    for (int nu = 0; nu < 4; ++nu) {
                        double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
                        if (MP::dep(mu) && MP::gam_nz(rho, nu, sg))
                            t1 = scratch[(uint64_t)(mu * 64 + rho * 16 + nu * 4 + sg) * n + k];
                        if (MP::dep(nu) && MP::gam_nz(rho, mu, sg))
                            t2 = scratch[(uint64_t)(nu * 64 + rho * 16 + mu * 4 + sg) * n + k];
                        for (int l = 0; l < 4; ++l) {
                            if (MP::gam_nz(rho, mu, l) && MP::gam_nz(l, nu, sg))
                                t3 += gam[rho * 16 + mu * 4 + l] * gam[l * 16 + nu * 4 + sg];
                            if (MP::gam_nz(rho, nu, l) && MP::gam_nz(l, mu, sg))
                                t4 += gam[rho * 16 + nu * 4 + l] * gam[l * 16 + mu * 4 + sg];
                        }
                        out[k * 256 + rho * 64 + sg * 16 + mu * 4 + nu] = t1 - t2 + t3 - t4;
                    }
        return;
    }

    // Ricci needs D1[d][ν][σ] = ∂_d Γ^d_νσ and D2[d][α][σ] = ∂_d Γ^α_ασ only.
    double D1[64], D2[64];
    sfor<4>([&](auto D) {
        constexpr int d = CI(D);
        if constexpr (MP::dep(d)) {
            double xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
            xp[d] += P.fd_h;
            xm[d] -= P.fd_h;
            double gp[64], gm[64];
            MP::christoffel(P, xp, gp);
            MP::christoffel(P, xm, gm);
            sfor<4>([&](auto B) {
                sfor<4>([&](auto C) {
                    constexpr int b = CI(B), c = CI(C);
                    if constexpr (MP::gam_nz(d, b, c))
                        D1[d * 16 + b * 4 + c] =
                            (gp[d * 16 + b * 4 + c] - gm[d * 16 + b * 4 + c]) * P.fd_inv_2h;
                    if constexpr (MP::gam_nz(b, b, c))
                        D2[d * 16 + b * 4 + c] =
                            (gp[b * 16 + b * 4 + c] - gm[b * 16 + b * 4 + c]) * P.fd_inv_2h;
                });
            });
        }
    });
    double ric[16];
    sfor<4>([&](auto Sg) {
        sfor<4>([&](auto Nu) {
            constexpr int sg = CI(Sg), nu = CI(Nu);
            double s = 0.0;
            sfor<4>([&](auto A) {
                constexpr int a = CI(A);
                double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
                if constexpr (MP::dep(a) && MP::gam_nz(a, nu, sg)) t1 = D1[a * 16 + nu * 4 + sg];
                if constexpr (MP::dep(nu) && MP::gam_nz(a, a, sg)) t2 = D2[nu * 16 + a * 4 + sg];
                sfor<4>([&](auto L) {
                    constexpr int l = CI(L);
                    if constexpr (MP::gam_nz(a, a, l) && MP::gam_nz(l, nu, sg))
                        t3 += gam[a * 16 + a * 4 + l] * gam[l * 16 + nu * 4 + sg];
                    if constexpr (MP::gam_nz(a, nu, l) && MP::gam_nz(l, a, sg))
                        t4 += gam[a * 16 + nu * 4 + l] * gam[l * 16 + a * 4 + sg];
                });
                s += t1 - t2 + t3 - t4;
            });
            ric[sg * 4 + nu] = s;
        });
    });
    if (what == 1) {
#pragma unroll
        for (int i = 0; i < 16; ++i) out[k * 16 + i] = ric[i];
        return;
    }
    double gi[16];
    MP::ginv_at(P, x, gi);
    double R = 0.0;
    sfor<16>([&](auto I) {
        constexpr int i = CI(I);
        if constexpr (MP::ginv_nz(i / 4, i % 4)) R += gi[i] * ric[i];
    });
    if (what == 4) {
        out[k] = R;
        return;
    }
    double g[16];
    MP::g(P, x, g);
    sfor<16>([&](auto I) {
        constexpr int i = CI(I);
        double e = ric[i];
        if constexpr (MP::g_nz(i / 4, i % 4)) e = ric[i] - 0.5 * g[i] * R;
        out[k * 16 + i] = (what == 3) ? e * factor : e;
    });
}

#define GEON_CURV_KERNEL(NAME, MP)                                                              \
    extern "C" __global__ __launch_bounds__(256) void geon_curvature_##NAME(                    \
        uint32_t slot, uint64_t n, const double* x, int32_t what, double* out, double* scratch,\
        double factor) {                                                                        \
        curvature_point<MP>(slot, n, x, what, out, scratch, factor);                            \
    }

// Metrics with analytic or cheap Γ keep the whole point in one lane.
GEON_CURV_KERNEL(mink, Minkowski)
GEON_CURV_KERNEL(schw, Schwarzschild)
GEON_CURV_KERNEL(mt, MorrisThorne)

// Alcubierre's Γ is expensive (seven metric evaluations, each a square root
// and two tanh), and the seven Γ of a point's stencil share most of their
// metric points: a team of 8 lanes per point evaluates the 31 distinct
// metric points once, forms the seven Γ on seven lanes and the outputs on
// all eight, exchanging through LDS (geon_curv.hpp); 64-lane blocks,
// 8 points per block. Same outputs bit for bit as one lane per point. The
// host launches it for Riemann (what 0); the Ricci family goes to
// geon_curvature_alc_ricci below, which gives the same values.
extern "C" __global__ __launch_bounds__(64) void geon_curvature_alc(
    uint32_t slot, uint64_t n, const double* xs, int32_t what, double* out, double* scratch,
    double factor) {
    (void)scratch;
    using CS = CurvStencil<Alcubierre>;
    static_assert(CS::kPoints <= GEON_CURV_TEAM, "stencil larger than the team");
    __shared__ double s_f[GEON_CURV_PTS][AlcStencil::kRow];
    __shared__ double s_gam[GEON_CURV_PTS][CS::kRow];
    __shared__ double s_ric[GEON_CURV_PTS][17];
    const geon_dmetric& P = c_frames[slot].m;
    const int t = (int)(threadIdx.x % GEON_CURV_TEAM);
    const int q = (int)(threadIdx.x / GEON_CURV_TEAM);
    const uint64_t k = (uint64_t)blockIdx.x * GEON_CURV_PTS + q;
    const bool valid = k < n;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
        x[0] = xs[k];
        x[1] = xs[n + k];
        x[2] = xs[2 * n + k];
        x[3] = xs[3 * n + k];
    }
    // A: the 31 distinct metric points, 4 rounds of the 8 lanes
    Alcubierre::ShapeExp E;
    const bool ex = curv_shape_exp(P);
    if (ex) Alcubierre::shape_exp(P, Alcubierre::rs2(P, x), E);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = t + GEON_CURV_TEAM * r;
        if (valid && j < AlcStencil::kUsed) s_f[q][j] = curv_alc_point_f(P, x, j, ex ? &E : nullptr);
    }
    __syncthreads();
    // B: Γ at the seven stencil points
    if (valid && t < CS::kPoints) curv_alc_gamma(P, s_f[q], t, &s_gam[q][t * CS::kNNZ]);
    __syncthreads();
    // C: outputs
    if (what == 0) {
        if (!valid) return;
        curv_riemann_lane<Alcubierre>(P, s_gam[q], t, out + k * 256);
        return;
    }
    double ric[2];
    curv_ricci_lane<Alcubierre>(P, s_gam[q], t, ric);
    if (what == 1) {
        if (valid) {
            out[k * 16 + t] = ric[0];
            out[k * 16 + t + GEON_CURV_TEAM] = ric[1];
        }
        return;
    }
    s_ric[q][t] = ric[0];
    s_ric[q][t + GEON_CURV_TEAM] = ric[1];
    __syncthreads();   // all 16 components of every point are in LDS
    if (!valid) return;
    double e[2];
    curv_alc_outputs_lane(P, s_f[q][0], s_ric[q], t, what, factor, e);
    if (what == 4) {
        if (t == 0) out[k] = e[0];
        return;
    }
    out[k * 16 + t] = e[0];
    out[k * 16 + t + GEON_CURV_TEAM] = e[1];
}

// A double moved between the lanes of each quad by DPP quad_perm: CTRL =
// 0xB1 swaps lanes 0↔1 and 2↔3, k·0x55 broadcasts lane k. Every lane of
// the wave must be active.
template <int CTRL>
__device__ double quad_dpp(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// Ricci (what 1), Einstein (2), stress-energy (3) and the Ricci scalar (4) of
// Alcubierre points by quads (AlcQuad in geon_curv.hpp): 4 lanes per point,
// 16 points per 64-lane block; the same outputs bit for bit as
// geon_curvature_alc, which keeps the Riemann output (what 0). Every lane
// runs every phase (a point past n computes at x = 0 and stores nothing), so
// that the DPP exchanges see the whole wave.
#define GEON_CURV_QPTS 16
extern "C" __global__ __launch_bounds__(64) void geon_curvature_alc_ricci(
    uint32_t slot, uint64_t n, const double* xs, int32_t what, double* out, double* scratch,
    double factor) {
    (void)scratch;
    __shared__ double s_pt[GEON_CURV_QPTS][AlcQuad::stride()];
    const geon_dmetric& P = c_frames[slot].m;
    const int t = (int)(threadIdx.x & 3u);
    const int q = (int)(threadIdx.x >> 2);
    const uint64_t k = (uint64_t)blockIdx.x * GEON_CURV_QPTS + q;
    const bool valid = k < n;
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
        x[0] = xs[k];
        x[1] = xs[n + k];
        x[2] = xs[2 * n + k];
        x[3] = xs[3 * n + k];
    }
    double* pt = s_pt[q];
    // A: the 31 distinct metric points, 8 rounds of the 4 lanes
    double sq[3];
    AlcQuad::base_squares(P, x, sq);
    Alcubierre::ShapeExp E;
    const bool ex = curv_shape_exp(P);
    if (ex) Alcubierre::shape_exp(P, sq[0] + sq[1] + sq[2], E);
    sfor<8>([&](auto R) { AlcQuad::shape_round<CI(R)>(P, x, sq, t, pt, ex ? &E : nullptr); });
    __syncthreads();
    // B: Γ at x ± h e_x, x ± h e_y (Γ-stencil points 1-4), then x ± h e_z
    // (5, 6) on lanes 0-1 and x (0) on lanes 2-3
    double ga[30], gb[30], oa[30], ob[30];
    curv_alc_gamma(P, pt, 1 + t, ga);
    curv_alc_gamma(P, pt, t < 2 ? 5 + t : 0, gb);
    const double f0 = pt[0];
    sfor<64>([&](auto I) {
        constexpr int i = CI(I), mu = i / 16, a = (i / 4) % 4, b = i % 4;
        if constexpr (AlcQuad::wanted(0x6, mu, a, b)) {
            constexpr int sl = AlcQuad::slot(mu, a, b);
            oa[sl] = quad_dpp<0xB1>(ga[sl]);
        }
        if constexpr (AlcQuad::wanted(0x8, mu, a, b)) {
            constexpr int sl = AlcQuad::slot(mu, a, b);
            ob[sl] = quad_dpp<0xB1>(gb[sl]);
        }
    });
    __syncthreads();   // every shape value has been read: the rows go over them
    const double ih = P.fd_inv_2h;
    if (t == 0) AlcQuad::drow<1>(ga, oa, ih, pt);   // lane 0 holds x + h e_x
    if (t == 1) AlcQuad::drow<3>(ob, gb, ih, pt);   // lane 1 holds x − h e_z
    if (t == 2) AlcQuad::drow<2>(ga, oa, ih, pt);   // lane 2 holds x + h e_y
    if (t == 3) AlcQuad::row0(gb, pt);              // lane 3 holds Γ(x)
    __syncthreads();
    // C: R_σν of lane σ
    double ric[4];
    AlcQuad::ricci(pt, t, ric);
    if (what == 1) {
        if (valid) {
#pragma unroll
            for (int nu = 0; nu < 4; ++nu) out[k * 16 + t * 4 + nu] = ric[nu];
        }
        return;
    }
    // D: R from the quad, then G or T
    const double r6[6] = {quad_dpp<0x00>(ric[0]), quad_dpp<0x00>(ric[1]), quad_dpp<0x55>(ric[0]),
                          quad_dpp<0x55>(ric[1]), quad_dpp<0xAA>(ric[2]), quad_dpp<0xFF>(ric[3])};
    double e[4], R;
    AlcQuad::outputs(P, f0, r6, ric, t, what, factor, e, &R);
    if (!valid) return;
    if (what == 4) {
        if (t == 0) out[k] = R;
        return;
    }
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) out[k * 16 + t * 4 + nu] = e[nu];
}

// ---- shadow bisection (validate_shadow.rs:20-95) ----------------------------
// One lane per problem k: Schwarzschild::new(mass[k]) (pad 1.001, escape
// 1000 M), equatorial launch at angle α from the inward radial direction,
// trace(step, max_steps) capture as the predicate, `iters` halvings of
// [0, π/2 − 1e-3].
__device__ void equatorial_ray(double m, double r_obs, double alpha, double* y) {
    const double f = 1.0 - 2.0 * m / r_obs;
    double sin_a, cos_a;
    sincos(alpha, &sin_a, &cos_a);
    y[0] = 0.0; y[1] = r_obs; y[2] = 1.5707963267948966; y[3] = 0.0;
    y[4] = 1.0 / sqrt(f);
    y[5] = -cos_a * sqrt(f);
    y[6] = 0.0;
    y[7] = sin_a / r_obs;
}

extern "C" __global__ __launch_bounds__(64) void geon_bisect_serial(
    uint64_t n, const double* __restrict__ mass, const double* __restrict__ r_obs, double step,
    uint32_t max_steps, uint32_t iters, double* __restrict__ alpha_c, double* __restrict__ b_c) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    geon_dmetric P = {};
    P.kind = 1;
    P.M = mass[k];
    P.cap_r = 1.001 * 2.0 * P.M;
    P.esc_r = 1000.0 * P.M;
    const double r = r_obs[k];
    double lo = 0.0, hi = 1.5707963267948966 - 1e-3;
    for (uint32_t it = 0; it < iters; ++it) {
        const double mid = 0.5 * (lo + hi);
        double y[8];
        equatorial_ray(P.M, r, mid, y);
        uint32_t s;
        const int c = trace<Schwarzschild>(P, y, step, max_steps, &s);
        if (c == CAPTURED) lo = mid;
        else hi = mid;
    }
    const double a = 0.5 * (lo + hi);
    const double f = 1.0 - 2.0 * P.M / r;
    alpha_c[k] = a;
    b_c[k] = r * sin(a) / sqrt(f);
}

// ---- persistent scheduler instantiations -------------------------------------
// WAVES as for the ray kernels: one ray per lane in FP32 fits four waves per
// SIMD with four spilled VGPRs (5 % faster than three); the FP64 kernels
// would spill dozens and keep the allocation the compiler picks.
#define GEON_SCHED(NAME, MP, T, TN, RPL, WAVES)                                                  \
    extern "C" __global__ __launch_bounds__(256, WAVES) void geon_sched_##NAME##_##TN##_r##RPL( \
        geon_sched_args A) {                                                                    \
        sched_run<MP, T, RPL>(A);                                                               \
    }
#define GEON_SCHED_ALL(NAME, MP)                                                                \
    GEON_SCHED(NAME, MP, double, f64, 1, 1)                                                     \
    GEON_SCHED(NAME, MP, double, f64, 2, 1)                                                     \
    GEON_SCHED(NAME, MP, float, f32, 1, 4)                                                      \
    GEON_SCHED(NAME, MP, float, f32, 2, 1)

GEON_SCHED_ALL(mink, Minkowski)
GEON_SCHED_ALL(schw, Schwarzschild)
GEON_SCHED_ALL(alc, Alcubierre)
GEON_SCHED_ALL(mt, MorrisThorne)
