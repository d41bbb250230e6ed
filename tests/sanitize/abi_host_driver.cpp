// Host-side exercise of the C ABI entry points that need no GPU, built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize_cpu.py.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "geon_hip.h"

static int expect(int got, int want, const char* what) {
    if (got != want) {
        fprintf(stderr, "%s: status %d, expected %d (%s)\n", what, got, want, geon_last_error());
        return 1;
    }
    return 0;
}

int main() {
    int bad = 0;
    for (int k = 0; k < 4; ++k) {
        geon_metric_t m;
        bad += expect(geon_metric_default(k, &m), GEON_OK, "metric_default");
        const double obs[4] = {0.0, k == 2 ? 10.0 : 30.0, k == 2 ? 0.0 : 1.2, 0.3};
        double tet[16];
        bad += expect(geon_tetrad(&m, obs, tet), GEON_OK, "tetrad");
        geon_camera_t cam;
        bad += expect(geon_camera_new(obs, 60.0, 7, 5, &cam), GEON_OK, "camera_new");
        // the Metric trait on the host, at the observer
        double g[16], dg[64], gi[16], gam[64], dir[3];
        int32_t cls;
        bad += expect(geon_metric_g(&m, obs, g), GEON_OK, "metric_g");
        bad += expect(geon_metric_dg(&m, obs, dg), GEON_OK, "metric_dg");
        bad += expect(geon_metric_g_inv(&m, obs, gi), GEON_OK, "metric_g_inv");
        bad += expect(geon_metric_christoffel(&m, obs, gam), GEON_OK, "metric_christoffel");
        bad += expect(geon_metric_terminate(&m, obs, &cls), GEON_OK, "metric_terminate");
        bad += expect(geon_metric_asymptotic_direction(&m, obs, tet, dir), GEON_OK, "direction");
        for (int c = 0; c < 64; ++c) bad += isfinite(gam[c]) ? 0 : 1;
        for (uint32_t j = 0; j < cam.height; ++j)
            for (uint32_t i = 0; i < cam.width; ++i) {
                double y[8];
                bad += expect(geon_ray_state(&cam, tet, i, j, 0.5, 0.5, y), GEON_OK, "ray_state");
                for (int c = 0; c < 8; ++c) bad += isfinite(y[c]) ? 0 : 1;
            }
    }
    geon_metric_t s;
    geon_metric_default(GEON_METRIC_SCHWARZSCHILD, &s);
    const double inside[4] = {0.0, 1.5, 1.0, 0.0};
    double tet[16];
    bad += expect(geon_tetrad(&s, inside, tet), GEON_E_NOT_TIMELIKE, "tetrad inside horizon");
    bad += expect(geon_metric_default(9, &s), GEON_E_INVALID, "unknown kind");
    float px[9] = {-1.0f, 0.0f, 0.25f, 0.5f, 1.0f, 2.0f, NAN, INFINITY, 1e-9f};
    uint8_t out[9];
    bad += expect(geon_tonemap_u8(px, 9, out), GEON_OK, "tonemap");
    for (int c = -12; c <= 0; ++c) bad += geon_strerror(c) ? 0 : 1;
    geon_ctx_t* ctx = nullptr;
    const int st = geon_ctx_create(0, nullptr, &ctx);   // no GPU here: an error, not a crash
    if (st == GEON_OK) geon_ctx_destroy(ctx);
    geon_batch_t* b = nullptr;
    bad += expect(geon_batch_create(nullptr, &b), GEON_E_INVALID, "batch without context");
    geon_server_t* srv = nullptr;
    bad += expect(geon_serve_create(nullptr, nullptr, &srv), GEON_E_INVALID, "server without context");
    geon_metric_t mt;
    geon_metric_default(GEON_METRIC_MORRIS_THORNE, &mt);
    const double xs[4] = {0.0, 0.0, 1.5707963267948966, 0.0}, kv[4] = {1.0, 1.0, 0.0, 0.0};
    double nec = 0.0, vmin = 0.0, integral = 0.0;
    uint64_t imin = 0;
    bad += expect(geon_nec_contraction(nullptr, &mt, xs, 1, kv, &nec), GEON_E_INVALID,
                  "nec contraction without context");
    bad += expect(geon_nec_scan(nullptr, &mt, xs, 1, kv, 0.05, nullptr, &vmin, &imin, &integral),
                  GEON_E_INVALID, "nec scan without context");
    const double y8[8] = {0.0, 15.0, 1.5707963267948966, 0.0, 1.0, -1.0, 0.0, 0.0};
    bad += expect(geon_trace_rays_adaptive(nullptr, &mt, y8, 1, 0.05, 1e-9, 1e-12, 10, nullptr,
                                           nullptr, nullptr, nullptr, nullptr),
                  GEON_E_INVALID, "adaptive trace without context");
    bad += expect(geon_trace_rays_adaptive_dev(nullptr, &mt, y8, 1, 0.0, 1e-9, 1e-12, 10, nullptr,
                                               nullptr, nullptr, nullptr, nullptr, nullptr),
                  GEON_E_INVALID, "adaptive trace with h0 = 0");
    printf("driver done, failures %d\n", bad);
    return bad ? 1 : 0;
}
