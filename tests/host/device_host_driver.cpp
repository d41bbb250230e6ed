// Host build of the kernels' FP64 device code (geon_device.hpp,
// geon_curv.hpp) for tests/test_device_host.py: traces, Γ, and the
// Alcubierre curvature team run lane after lane with its LDS rows in
// ordinary arrays. Metric constants come from the runtime's own
// make_dmetric.
#include <string.h>

#include "../../geon_amd/csrc/host/geon_runtime.hpp"
#include "../../geon_amd/csrc/kernels/geon_curv.hpp"

using namespace geon;
using geon_host::make_dmetric;

extern "C" int hd_trace(const geon_metric_t* m, double* y, double h, uint32_t max_steps,
                        uint32_t* steps) {
    geon_dmetric P;
    if (make_dmetric(m, &P) != GEON_OK) return -100;
    switch (m->kind) {
        case GEON_METRIC_MINKOWSKI: return trace<Minkowski>(P, y, h, max_steps, steps);
        case GEON_METRIC_SCHWARZSCHILD: return trace<Schwarzschild>(P, y, h, max_steps, steps);
        case GEON_METRIC_ALCUBIERRE: return trace<Alcubierre>(P, y, h, max_steps, steps);
        default: return trace<MorrisThorne>(P, y, h, max_steps, steps);
    }
}

extern "C" int hd_trace_adaptive(const geon_metric_t* m, double* y, double h0, double rtol,
                                 double atol, uint32_t max_steps, uint32_t* steps,
                                 uint32_t* rejected) {
    geon_dmetric P;
    if (make_dmetric(m, &P) != GEON_OK) return -100;
    switch (m->kind) {
        case GEON_METRIC_MINKOWSKI:
            return trace_adaptive<Minkowski>(P, y, h0, rtol, atol, max_steps, steps, rejected);
        case GEON_METRIC_SCHWARZSCHILD:
            return trace_adaptive<Schwarzschild>(P, y, h0, rtol, atol, max_steps, steps, rejected);
        case GEON_METRIC_ALCUBIERRE:
            return trace_adaptive<Alcubierre>(P, y, h0, rtol, atol, max_steps, steps, rejected);
        default:
            return trace_adaptive<MorrisThorne>(P, y, h0, rtol, atol, max_steps, steps, rejected);
    }
}

extern "C" int hd_christoffel(const geon_metric_t* m, const double* x, double* c) {
    geon_dmetric P;
    if (make_dmetric(m, &P) != GEON_OK) return -100;
    memset(c, 0, 64 * sizeof(double));
    switch (m->kind) {
        case GEON_METRIC_SCHWARZSCHILD: Schwarzschild::christoffel(P, x, c); break;
        case GEON_METRIC_ALCUBIERRE: Alcubierre::christoffel(P, x, c); break;
        case GEON_METRIC_MORRIS_THORNE: MorrisThorne::christoffel(P, x, c); break;
        default: break;
    }
    return 0;
}

// what: 0 Riemann (256 per point), 1 Ricci, 2 Einstein, 3 stress-energy
// (16 per point), 4 Ricci scalar (1 per point); x is n×4.
extern "C" int hd_curvature_alc_team(const geon_metric_t* m, const double* x, long n, int what,
                                     double factor, double* out) {
    using CS = CurvStencil<Alcubierre>;
    geon_dmetric P;
    if (m->kind != GEON_METRIC_ALCUBIERRE || make_dmetric(m, &P) != GEON_OK) return -100;
    for (long k = 0; k < n; ++k) {
        const double* xk = x + 4 * k;
        double sf[AlcStencil::kRow], row[CS::kRow], ric[17];
        Alcubierre::ShapeExp E;
        const bool ex = curv_shape_exp(P);   // the GPU's expansion (GEON_HOST_DEVICE_MATH)
        if (ex) Alcubierre::shape_exp(P, Alcubierre::rs2(P, xk), E);
        for (int t = 0; t < GEON_CURV_TEAM; ++t)   // phase A
            for (int r = 0; r < 4; ++r) {
                const int j = t + GEON_CURV_TEAM * r;
                if (j < AlcStencil::kUsed) sf[j] = curv_alc_point_f(P, xk, j, ex ? &E : nullptr);
            }
        for (int t = 0; t < CS::kPoints; ++t) curv_alc_gamma(P, sf, t, &row[t * CS::kNNZ]);
        if (what == 0) {
            for (int t = 0; t < GEON_CURV_TEAM; ++t) curv_riemann_lane<Alcubierre>(P, row, t, out + 256 * k);
            continue;
        }
        for (int t = 0; t < GEON_CURV_TEAM; ++t) {
            double r2[2];
            curv_ricci_lane<Alcubierre>(P, row, t, r2);
            ric[t] = r2[0];
            ric[t + GEON_CURV_TEAM] = r2[1];
        }
        if (what == 1) {
            memcpy(out + 16 * k, ric, 16 * sizeof(double));
            continue;
        }
        for (int t = 0; t < GEON_CURV_TEAM; ++t) {
            double e[2];
            curv_alc_outputs_lane(P, sf[0], ric, t, what, factor, e);
            if (what == 4) {
                if (t == 0) out[k] = e[0];
                continue;
            }
            out[16 * k + t] = e[0];
            out[16 * k + t + GEON_CURV_TEAM] = e[1];
        }
    }
    return 0;
}

// The quad form of the Ricci family (AlcQuad): the four lanes of a point
// run one after another with the point's LDS row; what as above, 1-4 only.
extern "C" int hd_curvature_alc_quad(const geon_metric_t* m, const double* x, long n, int what,
                                     double factor, double* out) {
    geon_dmetric P;
    if (m->kind != GEON_METRIC_ALCUBIERRE || what < 1 || what > 4 ||
        make_dmetric(m, &P) != GEON_OK)
        return -100;
    for (long k = 0; k < n; ++k) {
        const double* xk = x + 4 * k;
        double pt[AlcQuad::stride()];
        for (int i = 0; i < AlcQuad::stride(); ++i) pt[i] = -1.0;   // unread slots show up
        double sq[3];
        AlcQuad::base_squares(P, xk, sq);
        Alcubierre::ShapeExp E;
        const bool ex = curv_shape_exp(P);   // the GPU's expansion (GEON_HOST_DEVICE_MATH)
        if (ex) Alcubierre::shape_exp(P, sq[0] + sq[1] + sq[2], E);
        const Alcubierre::ShapeExp* Ep = ex ? &E : nullptr;
        for (int t = 0; t < 4; ++t) {
            AlcQuad::shape_round<0>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<1>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<2>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<3>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<4>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<5>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<6>(P, xk, sq, t, pt, Ep);
            AlcQuad::shape_round<7>(P, xk, sq, t, pt, Ep);
        }
        double gs[7][30];
        for (int t = 0; t < 7; ++t) curv_alc_gamma(P, pt, t, gs[t]);
        const double f0 = pt[0];
        AlcQuad::drow<1>(gs[1], gs[2], P.fd_inv_2h, pt);
        AlcQuad::drow<2>(gs[3], gs[4], P.fd_inv_2h, pt);
        AlcQuad::drow<3>(gs[5], gs[6], P.fd_inv_2h, pt);
        AlcQuad::row0(gs[0], pt);
        double ric[4][4];
        for (int sg = 0; sg < 4; ++sg) AlcQuad::ricci(pt, sg, ric[sg]);
        if (what == 1) {
            for (int i = 0; i < 16; ++i) out[16 * k + i] = ric[i / 4][i % 4];
            continue;
        }
        const double r6[6] = {ric[0][0], ric[0][1], ric[1][0], ric[1][1], ric[2][2], ric[3][3]};
        for (int sg = 0; sg < 4; ++sg) {
            double e[4], R;
            AlcQuad::outputs(P, f0, r6, ric[sg], sg, what, factor, e, &R);
            if (what == 4) {
                if (sg == 0) out[k] = R;
                continue;
            }
            for (int nu = 0; nu < 4; ++nu) out[16 * k + 4 * sg + nu] = e[nu];
        }
    }
    return 0;
}
