// geon_metric_host.cpp — the Metric trait's methods on the host
// (src/metric/mod.rs:19-109), for bindings that implement the reference
// trait for the metrics this library runs on the GPU.
//
// These evaluate the kernels' own metric policies (geon_device.hpp) compiled
// for the host, where sin, cos and tanh are the C library's, as the
// reference's f64 methods are. Each method takes the definition the
// reference uses for that metric:
//   g              every metric's own g (schwarzschild.rs:33-45,
//                  minkowski.rs:11-18, alcubierre.rs:48-63,
//                  morris_thorne.rs:22-36)
//   dg             zero for Minkowski (minkowski.rs:20-22); otherwise the
//                  trait default, central differences of g (mod.rs:25-43)
//   g_inv          analytic for Schwarzschild (schwarzschild.rs:47-59) and
//                  Minkowski (minkowski.rs:24-26); otherwise the cofactor
//                  inverse of g (mod.rs:46-48, types.rs:34-46)
//   christoffel    analytic for Schwarzschild (schwarzschild.rs:61-107),
//                  zero for Minkowski (minkowski.rs:28-30); otherwise the
//                  trait default from g⁻¹ and ∂g (mod.rs:51-69)
//   terminate, asymptotic_direction   each metric's own rule
// and sky::sample (sky.rs:107-120), the kernels' sky_sample.
#include <string.h>

#include "geon_runtime.hpp"
#include "../kernels/geon_device.hpp"

using namespace geon_host;

namespace {

template <class F>
int with_metric(const geon_metric_t* m, F&& f) {
    if (!metric_valid(m)) return fail(GEON_E_INVALID, "invalid metric description");
    geon_dmetric P;
    GEON_TRY(make_dmetric(m, &P));
    switch (m->kind) {
    case GEON_METRIC_MINKOWSKI: return f(P, (geon::Minkowski*)nullptr);
    case GEON_METRIC_SCHWARZSCHILD: return f(P, (geon::Schwarzschild*)nullptr);
    case GEON_METRIC_ALCUBIERRE: return f(P, (geon::Alcubierre*)nullptr);
    default: return f(P, (geon::MorrisThorne*)nullptr);
    }
}

// nalgebra's try_inverse fails (the reference then panics, types.rs:39)
// exactly when the determinant is zero; the metrics' inverses are the
// cofactor scheme specialised to their sparsity, whose determinant is the
// product (Morris–Thorne: m00·(m11·m22·m33)) or a − b² (Alcubierre).
int cofactor_inverse(const geon_dmetric& P, const geon::MorrisThorne*, const double* x,
                     double* gi) {
    double g0[16] = {};
    geon::MorrisThorne::g(P, x, g0);
    const double det = g0[0] * (g0[5] * g0[10] * g0[15]);
    if (det == 0.0) return fail(GEON_E_SINGULAR, "metric became singular");
    geon::MorrisThorne::ginv(P, g0, gi);
    return GEON_OK;
}
int cofactor_inverse(const geon_dmetric& P, const geon::Alcubierre*, const double* x, double* gi) {
    double g0[16] = {};
    geon::Alcubierre::g(P, x, g0);
    const double det = g0[0] - g0[1] * g0[1];
    if (det == 0.0) return fail(GEON_E_SINGULAR, "metric became singular");
    geon::Alcubierre::ginv(P, g0, gi);
    return GEON_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int geon_metric_g(const geon_metric_t* metric, const double x[4], double g[16]) {
    if (!x || !g) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric& P, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        memset(g, 0, 16 * sizeof(double));
        MP::g(P, x, g);
        return GEON_OK;
    });
}

int geon_metric_dg(const geon_metric_t* metric, const double x[4], double dg[64]) {
    if (!x || !dg) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric& P, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        memset(dg, 0, 64 * sizeof(double));
        if constexpr (!std::is_same<MP, geon::Minkowski>::value) geon::FD<MP>::dg(P, x, dg);
        return GEON_OK;
    });
}

int geon_metric_g_inv(const geon_metric_t* metric, const double x[4], double gi[16]) {
    if (!x || !gi) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric& P, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        memset(gi, 0, 16 * sizeof(double));
        if constexpr (std::is_same<MP, geon::Minkowski>::value ||
                      std::is_same<MP, geon::Schwarzschild>::value) {
            MP::ginv_at(P, x, gi);
            return GEON_OK;
        } else {
            return cofactor_inverse(P, mp, x, gi);
        }
    });
}

int geon_metric_christoffel(const geon_metric_t* metric, const double x[4], double gam[64]) {
    if (!x || !gam) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric& P, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        memset(gam, 0, 64 * sizeof(double));
        if constexpr (std::is_same<MP, geon::Schwarzschild>::value) {
            MP::christoffel(P, x, gam);
        } else if constexpr (!std::is_same<MP, geon::Minkowski>::value) {
            double gi[16] = {};
            GEON_TRY(cofactor_inverse(P, mp, x, gi));
            // the trait default: g⁻¹ and central-difference ∂g, each sum
            // from 0.0 as the reference's loop
            geon::FD<MP>::christoffel(P, x, gam);
        }
        return GEON_OK;
    });
}

int geon_metric_terminate(const geon_metric_t* metric, const double x[4], int32_t* cls) {
    if (!x || !cls) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric& P, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        const int t = MP::terminate(P, x);
        *cls = t == geon::RUNNING ? -1 : t;
        return GEON_OK;
    });
}

int geon_metric_asymptotic_direction(const geon_metric_t* metric, const double x[4],
                                     const double p[4], double dir[3]) {
    if (!x || !p || !dir) return fail(GEON_E_INVALID, "bad argument");
    return with_metric(metric, [&](const geon_dmetric&, auto* mp) {
        using MP = std::remove_pointer_t<decltype(mp)>;
        const double y[8] = {x[0], x[1], x[2], x[3], p[0], p[1], p[2], p[3]};
        geon::direction<MP>(y, dir);
        return GEON_OK;
    });
}

int geon_sky_sample(const double dir[3], int32_t show_stars, float rgb[3]) {
    if (!dir || !rgb) return fail(GEON_E_INVALID, "bad argument");
    geon::sky_sample(dir, show_stars != 0, rgb);
    return GEON_OK;
}

}  // extern "C"
#pragma GCC visibility pop
