/* abi_c_driver.c — a plain C99 client of include/geon_hip.h (test driver).
 *
 *   abi_c_driver exports            host-only entry points and error paths;
 *                                   prints "OK <checks>" (no GPU needed)
 *   abi_c_driver render W H FILE    geon_render of the Schwarzschild frame
 *                                   (M=1, pad 1.005, escape 100M, camera
 *                                   r=30M on the equator, 60°, step 0.05,
 *                                   max 4000) on device 0; writes per-pixel
 *                                   class (int32), direction (3 planes of
 *                                   f64), RGB (f32 ×3) and the statistics
 *                                   (geon_render_stats_t) to FILE
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/geon_hip.h"

static int fails = 0, checks = 0;
#define EXPECT(cond, what)                                              \
    do {                                                                \
        ++checks;                                                       \
        if (!(cond)) {                                                  \
            ++fails;                                                    \
            fprintf(stderr, "FAIL %s (%s:%d): %s\n", what, __FILE__, __LINE__, geon_last_error()); \
        }                                                               \
    } while (0)

static int exports(void) {
    int k;
    EXPECT(geon_version() > 0, "version");
    EXPECT(strcmp(geon_strerror(GEON_OK), "ok") == 0, "strerror ok");
    EXPECT(strlen(geon_strerror(GEON_E_SINGULAR)) > 0, "strerror singular");
    for (k = GEON_METRIC_MINKOWSKI; k <= GEON_METRIC_MORRIS_THORNE; ++k) {
        geon_metric_t m;
        double obs[4] = {0.0, 30.0, 1.5707963267948966, 0.0};
        double tet[16], g[16], dg[64], gi[16], gam[64], y[8], dir[3];
        int32_t cls;
        geon_camera_t cam;
        EXPECT(geon_metric_default(k, &m) == GEON_OK && m.kind == k, "metric_default");
        EXPECT(m.horizon_pad == 1.001 && m.escape_radius == 1000.0 && m.fd_step == 1e-5, "defaults");
        if (k == GEON_METRIC_ALCUBIERRE || k == GEON_METRIC_MINKOWSKI) obs[2] = 0.0;
        EXPECT(geon_camera_new(obs, 60.0, 16, 8, &cam) == GEON_OK, "camera_new");
        EXPECT(cam.fov_y_rad == 60.0 * (3.141592653589793 / 180.0), "fov radians");
        EXPECT(geon_tetrad(&m, obs, tet) == GEON_OK, "tetrad");
        EXPECT(geon_ray_state(&cam, tet, 3, 2, 0.5, 0.5, y) == GEON_OK, "ray_state");
        EXPECT(geon_metric_g(&m, obs, g) == GEON_OK, "metric_g");
        /* the camera ray is null: g(p, p) = 0 to rounding */
        {
            double s = 0.0;
            int i, j;
            for (i = 0; i < 4; ++i)
                for (j = 0; j < 4; ++j) s += g[4 * i + j] * y[4 + i] * y[4 + j];
            EXPECT(fabs(s) < 1e-12, "camera ray null");
        }
        EXPECT(geon_metric_dg(&m, obs, dg) == GEON_OK, "metric_dg");
        EXPECT(geon_metric_g_inv(&m, obs, gi) == GEON_OK, "metric_g_inv");
        EXPECT(geon_metric_christoffel(&m, obs, gam) == GEON_OK, "metric_christoffel");
        EXPECT(geon_metric_terminate(&m, obs, &cls) == GEON_OK && cls == -1, "terminate");
        EXPECT(geon_metric_asymptotic_direction(&m, obs, y + 4, dir) == GEON_OK, "direction");
        EXPECT(fabs(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] - 1.0) < 1e-14, "unit dir");
    }
    {   /* error paths: the reference's panics become status codes */
        geon_metric_t s;
        double inside[4] = {0.0, 1.5, 1.0, 0.0}, tet[16], g[16];
        geon_metric_default(GEON_METRIC_SCHWARZSCHILD, &s);
        EXPECT(geon_tetrad(&s, inside, tet) == GEON_E_NOT_TIMELIKE, "not timelike");
        EXPECT(geon_metric_g(NULL, inside, g) == GEON_E_INVALID, "null metric");
        EXPECT(geon_metric_default(7, &s) == GEON_E_INVALID, "bad kind");
        EXPECT(geon_camera_new(NULL, 60.0, 1, 1, NULL) == GEON_E_INVALID, "null camera");
    }
    {
        float px[4] = {-1.0f, 0.0f, 0.5f, 2.0f}, rgb[3];
        unsigned char out[4];
        double d[3] = {0.0, 0.0, 1.0};
        EXPECT(geon_tonemap_u8(px, 4, out) == GEON_OK && out[0] == 0 && out[3] == 255, "tonemap");
        EXPECT(geon_sky_sample(d, 1, rgb) == GEON_OK && rgb[0] >= 0.0f && rgb[0] <= 1.0f, "sky");
    }
    printf("%s %d\n", fails ? "FAIL" : "OK", checks);
    return fails ? 1 : 0;
}

static int render(unsigned w, unsigned h, const char* path) {
    geon_ctx_t* ctx = NULL;
    geon_metric_t m;
    geon_camera_t cam;
    geon_render_cfg_t cfg;
    geon_render_stats_t st;
    const double obs[4] = {0.0, 30.0, 1.5707963267948966, 0.0};
    size_t n = (size_t)w * h;
    int32_t* cls = malloc(n * sizeof *cls);
    double* dir = malloc(3 * n * sizeof *dir);
    float* rgb = malloc(3 * n * sizeof *rgb);
    FILE* f;
    int rc;
    if (!cls || !dir || !rgb) return 3;
    geon_metric_default(GEON_METRIC_SCHWARZSCHILD, &m);
    m.horizon_pad = 1.005;
    m.escape_radius = 100.0;
    geon_camera_new(obs, 60.0, w, h, &cam);
    memset(&cfg, 0, sizeof cfg);
    cfg.step = 0.05;
    cfg.max_steps = 4000;
    cfg.samples_per_axis = 1;
    cfg.diverged_rgb[0] = 1.0f;
    cfg.diverged_rgb[2] = 1.0f;
    cfg.show_stars = 1;
    cfg.precision = GEON_PREC_FP64;
    rc = geon_ctx_create(0, NULL, &ctx);
    if (rc != GEON_OK) {
        fprintf(stderr, "geon_ctx_create: %s (%s)\n", geon_strerror(rc), geon_last_error());
        return 1;
    }
    rc = geon_render(ctx, &m, &cam, &cfg, rgb, cls, dir, &st);
    geon_ctx_destroy(ctx);
    if (rc != GEON_OK) {
        fprintf(stderr, "geon_render: %s (%s)\n", geon_strerror(rc), geon_last_error());
        return 1;
    }
    f = fopen(path, "wb");
    if (!f) return 1;
    fwrite(cls, sizeof *cls, n, f);
    fwrite(dir, sizeof *dir, 3 * n, f);
    fwrite(rgb, sizeof *rgb, 3 * n, f);
    fwrite(&st, sizeof st, 1, f);
    fclose(f);
    printf("rendered %ux%u: captured %llu escaped %llu diverged %llu ray-steps %llu\n", w, h,
           (unsigned long long)st.captured, (unsigned long long)st.escaped,
           (unsigned long long)st.diverged, (unsigned long long)st.ray_steps);
    free(cls);
    free(dir);
    free(rgb);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && strcmp(argv[1], "exports") == 0) return exports();
    if (argc == 5 && strcmp(argv[1], "render") == 0)
        return render((unsigned)atoi(argv[2]), (unsigned)atoi(argv[3]), argv[4]);
    fprintf(stderr, "usage: abi_c_driver exports | render W H FILE\n");
    return 2;
}
