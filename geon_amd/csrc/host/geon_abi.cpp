// geon_abi.cpp — the extern "C" entry points of include/geon_hip.h.
#include <cmath>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "geon_runtime.hpp"

using namespace geon_host;

namespace {

const double kPi = 3.141592653589793;

hipStream_t pick_stream(geon_ctx* c, void* stream) {
    return stream ? (hipStream_t)stream : c->stream;
}

int fill_frame(const geon_metric_t* metric, geon_dframe* F) {
    memset(F, 0, sizeof *F);
    return make_dmetric(metric, &F->m);
}

// Camera::new + ray_state_sub constants (camera.rs:29-36,113-121).
int fill_camera(const geon_metric_t* metric, const geon_camera_t* cam, geon_dframe* F) {
    if (!cam || cam->width == 0 || cam->height == 0)
        return fail(GEON_E_INVALID, "camera has zero size");
    GEON_TRY(tetrad(metric, cam->observer, F->tet));
    memcpy(F->obs, cam->observer, sizeof F->obs);
    const double aspect = (double)cam->width / (double)cam->height;
    F->h_im = tan(cam->fov_y_rad * 0.5);
    F->w_im = F->h_im * aspect;
    F->W = cam->width;
    F->H = cam->height;
    return GEON_OK;
}

}  // namespace

namespace geon_host {

int fp64_frame(const geon_metric_t* metric, const geon_camera_t* cam, const geon_render_cfg_t* cfg,
               geon_dframe* F, uint64_t* n_pix, uint64_t* n_rays) {
    GEON_TRY(fill_frame(metric, F));
    GEON_TRY(fill_camera(metric, cam, F));
    const uint32_t spa = cfg->samples_per_axis ? cfg->samples_per_axis : 1;
    F->spa = spa;
    F->inv_spa = 1.0 / (double)spa;
    F->step = cfg->step;
    F->max_steps = cfg->max_steps;
    memcpy(F->captured_rgb, cfg->captured_rgb, sizeof F->captured_rgb);
    memcpy(F->diverged_rgb, cfg->diverged_rgb, sizeof F->diverged_rgb);
    F->show_stars = cfg->show_stars;
    F->precision = GEON_PREC_FP64;
    *n_pix = (uint64_t)cam->width * cam->height;
    *n_rays = *n_pix * spa * spa;
    return GEON_OK;
}

int fp64_run(geon_ctx* ctx, const geon_dframe* F, uint64_t n_pix, uint64_t n_rays,
             const fp64_planes& P, float* rgb, int32_t* cls_per_pixel, double* dir_per_pixel,
             geon_render_stats_t* stats, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
    uint32_t slot;
    GEON_TRY(ctx_slot(ctx, F, s, &slot));
    geon_dstats* d_stats = P.stats;
    int32_t* d_cls = P.cls;
    uint32_t* d_steps = P.steps;
    double* d_res = P.res;
    double* d_dir = P.dir;
    void* cargs[] = {&d_stats};
    GEON_HIP(hipModuleLaunchKernel(ctx->fn_stats_clear, 1, 1, 1, 64, 1, 1, 0, s, cargs, nullptr),
             "stats clear");
    if (t0) GEON_HIP(hipEventRecord(t0, s), "event");
    void* rargs[] = {&slot, (void*)&n_rays, &d_cls, &d_steps, &d_res, &d_dir, &d_stats};
    GEON_TRY(launch(ctx->fn_render[F->m.kind], n_rays, 256, s, rargs));
    if (t1) GEON_HIP(hipEventRecord(t1, s), "event");
    void* fargs[] = {&slot, (void*)&n_pix, &d_cls, &d_steps, &d_res, &d_dir,
                     &rgb, &cls_per_pixel, &dir_per_pixel, &d_stats};
    GEON_TRY(launch(ctx->fn_finalize, n_pix, 256, s, fargs));
    GEON_TRY(ctx_slot_done(ctx, slot, s));
    if (stats) {
        geon_dstats ds;
        GEON_HIP(hipMemcpyDtoHAsync(&ds, (hipDeviceptr_t)d_stats, sizeof ds, s), "D2H stats");
        GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
        float ms = 0.0f;
        if (t0 && t1) hipEventElapsedTime(&ms, t0, t1);
        memset(stats, 0, sizeof *stats);
        stats->captured = ds.captured;
        stats->escaped = ds.escaped;
        stats->diverged = ds.diverged;
        double mr;
        memcpy(&mr, &ds.max_resid_bits, sizeof mr);
        stats->max_null_residual = mr;
        stats->avg_steps = (double)ds.steps_pix_sum / (double)n_pix;
        stats->max_steps_hit = ds.max_steps_hit;
        stats->ray_steps = ds.ray_steps;
        stats->fp64_rays = n_rays;
        stats->kernel_ms = ms;
        stats->gamma_steps = ds.gamma_steps;
    }
    return GEON_OK;
}

}  // namespace geon_host

// Only the C ABI is exported; the library is built with -fvisibility=hidden.
#pragma GCC visibility push(default)
extern "C" {

const char* geon_strerror(int status) {
    switch (status) {
    case GEON_OK: return "ok";
    case GEON_E_INVALID: return "invalid argument";
    case GEON_E_NOT_TIMELIKE: return "observer is not in a timelike region (g_tt >= 0)";
    case GEON_E_GRAM_SCHMIDT: return "Gram-Schmidt produced a non-spacelike vector";
    case GEON_E_SINGULAR: return "metric became singular";
    case GEON_E_HIP: return "HIP runtime error";
    case GEON_E_NOMEM: return "device allocation failed";
    case GEON_E_ADMISSION: return "request does not fit the HBM budget";
    case GEON_E_NO_MODULE: return "gfx950 code object missing or not loadable";
    case GEON_E_UNSUPPORTED: return "unsupported combination";
    case GEON_E_NOT_FOUND: return "unknown ticket";
    case GEON_E_PENDING: return "request still running";
    default: return "unknown status";
    }
}

const char* geon_last_error(void) { return g_last_error.c_str(); }

int geon_version(void) { return 10000; }

int geon_metric_default(int32_t kind, geon_metric_t* out) {
    if (!out || kind < GEON_METRIC_MINKOWSKI || kind > GEON_METRIC_MORRIS_THORNE)
        return fail(GEON_E_INVALID, "unknown metric kind %d", kind);
    memset(out, 0, sizeof *out);
    out->kind = kind;
    out->horizon_pad = 1.001;
    out->escape_radius = 1000.0;
    out->fd_step = 1e-5;
    if (kind == GEON_METRIC_SCHWARZSCHILD) out->p[0] = 1.0;
    if (kind == GEON_METRIC_ALCUBIERRE) {
        out->p[0] = 2.0; out->p[1] = 0.5; out->p[2] = 2.0; out->p[3] = 0.0;
    }
    if (kind == GEON_METRIC_MORRIS_THORNE) out->p[0] = 1.0;
    return GEON_OK;
}

// Camera::new: f64::to_radians is deg·(π/180) (camera.rs:29-36).
int geon_camera_new(const double observer[4], double fov_y_deg, uint32_t width, uint32_t height,
                    geon_camera_t* out) {
    if (!observer || !out) return fail(GEON_E_INVALID, "bad argument");
    memcpy(out->observer, observer, sizeof out->observer);
    out->fov_y_rad = fov_y_deg * (kPi / 180.0);
    out->width = width;
    out->height = height;
    return GEON_OK;
}

int geon_ctx_create(int device, const char* code_object, geon_ctx_t** out) {
    return ctx_create(device, code_object, out);
}

void geon_ctx_destroy(geon_ctx_t* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    for (auto& kv : ctx->lanes) lane_reap(kv.second, true);
    ctx_destroy(ctx);
}

int geon_tetrad(const geon_metric_t* metric, const double observer[4], double tetrad_out[16]) {
    if (!metric_valid(metric) || !observer || !tetrad_out) return fail(GEON_E_INVALID, "bad argument");
    return tetrad(metric, observer, tetrad_out);
}

int geon_ray_state(const geon_camera_t* cam, const double tet[16], uint32_t i, uint32_t j,
                   double du, double dv, double y[8]) {
    if (!cam || !tet || !y || cam->width == 0 || cam->height == 0)
        return fail(GEON_E_INVALID, "bad argument");
    const double aspect = (double)cam->width / (double)cam->height;
    const double h = tan(cam->fov_y_rad * 0.5);
    const double w = h * aspect;
    const double u = (2.0 * ((double)i + du) / (double)cam->width - 1.0) * w;
    const double v = -(2.0 * ((double)j + dv) / (double)cam->height - 1.0) * h;
    const double dx = -1.0, dy = -v, dz = u;
    const double norm = sqrt(dx * dx + dy * dy + dz * dz);
    const double n1 = dx / norm, n2 = dy / norm, n3 = dz / norm;
    for (int mu = 0; mu < 4; ++mu) {
        y[mu] = cam->observer[mu];
        y[4 + mu] = tet[mu] + n1 * tet[4 + mu] + n2 * tet[8 + mu] + n3 * tet[12 + mu];
    }
    return GEON_OK;
}

int geon_trace_rays_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* y0_soa,
                        size_t n, double step, uint32_t max_steps, int32_t* cls, uint32_t* steps,
                        double* y_final_soa, double* null_resid, double* dir_soa, void* stream) {
    if (!ctx || !y0_soa) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    geon_dframe F;
    GEON_TRY(fill_frame(metric, &F));
    F.step = step;
    F.max_steps = max_steps;
    hipStream_t s = pick_stream(ctx, stream);
    uint32_t slot;
    GEON_TRY(ctx_slot(ctx, &F, s, &slot));
    uint64_t nn = n;
    void* args[] = {&slot, &nn, (void*)&y0_soa, &cls, &steps, &y_final_soa, &null_resid, &dir_soa};
    GEON_TRY(launch(ctx->fn_trace[metric->kind], nn, 256, s, args));
    return ctx_slot_done(ctx, slot, s);
}

int geon_trace_rays(geon_ctx_t* ctx, const geon_metric_t* metric, const double* y0_soa, size_t n,
                    double step, uint32_t max_steps, int32_t* cls, uint32_t* steps,
                    double* y_final_soa, double* null_resid, double* dir_soa) {
    if (!ctx || !y0_soa) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    void *d_y0, *d_cls = nullptr, *d_steps = nullptr, *d_yf = nullptr, *d_res = nullptr, *d_dir = nullptr;
    GEON_TRY(ctx_buffer(ctx, "tr_y0", n * 8 * sizeof(double), &d_y0));
    if (cls) GEON_TRY(ctx_buffer(ctx, "tr_cls", n * sizeof(int32_t), &d_cls));
    if (steps) GEON_TRY(ctx_buffer(ctx, "tr_steps", n * sizeof(uint32_t), &d_steps));
    if (y_final_soa) GEON_TRY(ctx_buffer(ctx, "tr_yf", n * 8 * sizeof(double), &d_yf));
    if (null_resid) GEON_TRY(ctx_buffer(ctx, "tr_res", n * sizeof(double), &d_res));
    if (dir_soa) GEON_TRY(ctx_buffer(ctx, "tr_dir", n * 3 * sizeof(double), &d_dir));
    hipStream_t s = ctx->stream;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_y0, (void*)y0_soa, n * 8 * sizeof(double), s), "H2D");
    GEON_TRY(geon_trace_rays_dev(ctx, metric, (const double*)d_y0, n, step, max_steps, (int32_t*)d_cls,
                                 (uint32_t*)d_steps, (double*)d_yf, (double*)d_res, (double*)d_dir,
                                 nullptr));
    if (cls) GEON_HIP(hipMemcpyDtoHAsync(cls, (hipDeviceptr_t)d_cls, n * 4, s), "D2H");
    if (steps) GEON_HIP(hipMemcpyDtoHAsync(steps, (hipDeviceptr_t)d_steps, n * 4, s), "D2H");
    if (y_final_soa) GEON_HIP(hipMemcpyDtoHAsync(y_final_soa, (hipDeviceptr_t)d_yf, n * 64, s), "D2H");
    if (null_resid) GEON_HIP(hipMemcpyDtoHAsync(null_resid, (hipDeviceptr_t)d_res, n * 8, s), "D2H");
    if (dir_soa) GEON_HIP(hipMemcpyDtoHAsync(dir_soa, (hipDeviceptr_t)d_dir, n * 24, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return GEON_OK;
}

static bool adaptive_args_ok(double h0, double rtol, double atol) {
    return h0 != 0.0 && isfinite(h0) && rtol >= 0.0 && isfinite(rtol) && atol > 0.0 && isfinite(atol);
}

int geon_trace_rays_adaptive_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* y0_soa,
                                 size_t n, double h0, double rtol, double atol, uint32_t max_steps,
                                 int32_t* cls, uint32_t* steps, uint32_t* rejected,
                                 double* y_final_soa, double* dir_soa, void* stream) {
    if (!ctx || !metric || !y0_soa || !adaptive_args_ok(h0, rtol, atol))
        return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    geon_dframe F;
    GEON_TRY(fill_frame(metric, &F));
    F.step = h0;
    F.max_steps = max_steps;
    hipStream_t s = pick_stream(ctx, stream);
    uint32_t slot;
    GEON_TRY(ctx_slot(ctx, &F, s, &slot));
    uint64_t nn = n;
    void* args[] = {&slot, &nn, (void*)&y0_soa, &rtol, &atol, &cls, &steps, &rejected, &y_final_soa,
                    &dir_soa};
    GEON_TRY(launch(ctx->fn_trace_adaptive[metric->kind], nn, 256, s, args));
    return ctx_slot_done(ctx, slot, s);
}

int geon_trace_rays_adaptive(geon_ctx_t* ctx, const geon_metric_t* metric, const double* y0_soa,
                             size_t n, double h0, double rtol, double atol, uint32_t max_steps,
                             int32_t* cls, uint32_t* steps, uint32_t* rejected, double* y_final_soa,
                             double* dir_soa) {
    if (!ctx || !metric || !y0_soa || !adaptive_args_ok(h0, rtol, atol))
        return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    void *d_y0, *d_cls = nullptr, *d_steps = nullptr, *d_rej = nullptr, *d_yf = nullptr, *d_dir = nullptr;
    GEON_TRY(ctx_buffer(ctx, "ad_y0", n * 64, &d_y0));
    if (cls) GEON_TRY(ctx_buffer(ctx, "ad_cls", n * 4, &d_cls));
    if (steps) GEON_TRY(ctx_buffer(ctx, "ad_steps", n * 4, &d_steps));
    if (rejected) GEON_TRY(ctx_buffer(ctx, "ad_rej", n * 4, &d_rej));
    if (y_final_soa) GEON_TRY(ctx_buffer(ctx, "ad_yf", n * 64, &d_yf));
    if (dir_soa) GEON_TRY(ctx_buffer(ctx, "ad_dir", n * 24, &d_dir));
    hipStream_t s = ctx->stream;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_y0, (void*)y0_soa, n * 64, s), "H2D");
    GEON_TRY(geon_trace_rays_adaptive_dev(ctx, metric, (const double*)d_y0, n, h0, rtol, atol,
                                          max_steps, (int32_t*)d_cls, (uint32_t*)d_steps,
                                          (uint32_t*)d_rej, (double*)d_yf, (double*)d_dir, s));
    if (cls) GEON_HIP(hipMemcpyDtoHAsync(cls, (hipDeviceptr_t)d_cls, n * 4, s), "D2H");
    if (steps) GEON_HIP(hipMemcpyDtoHAsync(steps, (hipDeviceptr_t)d_steps, n * 4, s), "D2H");
    if (rejected) GEON_HIP(hipMemcpyDtoHAsync(rejected, (hipDeviceptr_t)d_rej, n * 4, s), "D2H");
    if (y_final_soa) GEON_HIP(hipMemcpyDtoHAsync(y_final_soa, (hipDeviceptr_t)d_yf, n * 64, s), "D2H");
    if (dir_soa) GEON_HIP(hipMemcpyDtoHAsync(dir_soa, (hipDeviceptr_t)d_dir, n * 24, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return GEON_OK;
}

int geon_render_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const geon_camera_t* cam,
                    const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                    double* dir_per_pixel, geon_render_stats_t* stats, void* stream) {
    if (!ctx || !cfg) return fail(GEON_E_INVALID, "bad argument");
    hipSetDevice(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    geon_lane* lane;
    GEON_TRY(ctx_lane(ctx, s, &lane));
    lane_reap(lane, false);
    if (cfg->precision == GEON_PREC_MIXED) {
        // FP32-mixed renders run on the persistent scheduler. The batch's
        // buffers live until its work on `s` has finished: with statistics
        // requested the call waits for it anyway; otherwise the lane keeps
        // the batch and frees it once its end event has passed.
        geon_batch_t* b = nullptr;
        GEON_TRY(geon_batch_create(ctx, &b));
        uint32_t fid = 0;
        int st = geon_batch_add_frame(b, metric, cam, cfg, rgb, cls_per_pixel, dir_per_pixel, &fid);
        if (st == GEON_OK) st = geon_batch_run(b, nullptr, s);
        if (st == GEON_OK && stats) st = geon_batch_frame_stats(b, fid, stats);
        if (st == GEON_OK && !stats) lane->retired.push_back(b);
        else geon_batch_destroy(b);
        return st;
    }
    if (cfg->precision != GEON_PREC_FP64)
        return fail(GEON_E_INVALID, "unknown precision %d", cfg->precision);
    geon_dframe F;
    uint64_t n_pix, n_rays;
    GEON_TRY(fp64_frame(metric, cam, cfg, &F, &n_pix, &n_rays));
    fp64_planes P;
    void *d_cls, *d_steps, *d_res, *d_dir;
    GEON_TRY(lane_buffer(lane, "ray_cls", n_rays * 4, &d_cls));
    GEON_TRY(lane_buffer(lane, "ray_steps", n_rays * 4, &d_steps));
    GEON_TRY(lane_buffer(lane, "ray_res", n_rays * 8, &d_res));
    GEON_TRY(lane_buffer(lane, "ray_dir", n_rays * 24, &d_dir));
    P.cls = (int32_t*)d_cls;
    P.steps = (uint32_t*)d_steps;
    P.res = (double*)d_res;
    P.dir = (double*)d_dir;
    P.stats = lane->d_stats;
    return fp64_run(ctx, &F, n_pix, n_rays, P, rgb, cls_per_pixel, dir_per_pixel, stats, s, lane->t0,
                    lane->t1);
}

int geon_render(geon_ctx_t* ctx, const geon_metric_t* metric, const geon_camera_t* cam,
                const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                double* dir_per_pixel, geon_render_stats_t* stats) {
    if (!ctx || !cam) return fail(GEON_E_INVALID, "bad argument");
    hipSetDevice(ctx->device);
    const uint64_t n_pix = (uint64_t)cam->width * cam->height;
    void *d_rgb = nullptr, *d_cls = nullptr, *d_dir = nullptr;
    if (rgb) GEON_TRY(ctx_buffer(ctx, "pix_rgb", n_pix * 12, &d_rgb));
    if (cls_per_pixel) GEON_TRY(ctx_buffer(ctx, "pix_cls", n_pix * 4, &d_cls));
    if (dir_per_pixel) GEON_TRY(ctx_buffer(ctx, "pix_dir", n_pix * 24, &d_dir));
    geon_render_stats_t st;
    GEON_TRY(geon_render_dev(ctx, metric, cam, cfg, (float*)d_rgb, (int32_t*)d_cls, (double*)d_dir, &st,
                             nullptr));
    hipStream_t s = ctx->stream;
    if (rgb) GEON_HIP(hipMemcpyDtoHAsync(rgb, (hipDeviceptr_t)d_rgb, n_pix * 12, s), "D2H");
    if (cls_per_pixel) GEON_HIP(hipMemcpyDtoHAsync(cls_per_pixel, (hipDeviceptr_t)d_cls, n_pix * 4, s), "D2H");
    if (dir_per_pixel) GEON_HIP(hipMemcpyDtoHAsync(dir_per_pixel, (hipDeviceptr_t)d_dir, n_pix * 24, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (stats) *stats = st;
    return GEON_OK;
}

int geon_curvature_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa, size_t n,
                       int32_t what, double* out, void* stream) {
    if (!ctx || !x_soa || !out || what < GEON_CURV_RIEMANN || what > GEON_CURV_SCALAR)
        return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    geon_dframe F;
    GEON_TRY(fill_frame(metric, &F));
    hipStream_t s = pick_stream(ctx, stream);
    uint32_t slot;
    GEON_TRY(ctx_slot(ctx, &F, s, &slot));
    uint64_t nn = n;
    double factor = 1.0 / (8.0 * kPi);  // curvature.rs:118
    void* scratch = nullptr;
    if (what == GEON_CURV_RIEMANN && metric->kind != GEON_METRIC_ALCUBIERRE) {
        geon_lane* lane;
        GEON_TRY(ctx_lane(ctx, s, &lane));
        GEON_TRY(lane_buffer(lane, "curv_scratch", n * 256 * 8, &scratch));
    }
    void* args[] = {&slot, &nn, (void*)&x_soa, &what, &out, &scratch, &factor};
    if (metric->kind == GEON_METRIC_ALCUBIERRE && what == GEON_CURV_RIEMANN)
        // one 8-lane team per point, 8 points per block
        GEON_TRY(launch(ctx->fn_curv[metric->kind], ((nn + 7) / 8) * 64, 64, s, args));
    else if (metric->kind == GEON_METRIC_ALCUBIERRE)   // one quad per point, 16 points per block
        GEON_TRY(launch(ctx->fn_curv_alc_ricci, ((nn + 15) / 16) * 64, 64, s, args));
    else                                          // one lane per point
        GEON_TRY(launch(ctx->fn_curv[metric->kind], nn, 256, s, args));
    return ctx_slot_done(ctx, slot, s);
}

int geon_curvature(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa, size_t n,
                   int32_t what, double* out) {
    if (!ctx || !x_soa || !out) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    const size_t w = (what == GEON_CURV_RIEMANN) ? 256 : (what == GEON_CURV_SCALAR ? 1 : 16);
    void *d_x, *d_out;
    GEON_TRY(ctx_buffer(ctx, "cv_x", n * 32, &d_x));
    GEON_TRY(ctx_buffer(ctx, "cv_out", n * w * 8, &d_out));
    hipStream_t s = ctx->stream;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_x, (void*)x_soa, n * 32, s), "H2D");
    GEON_TRY(geon_curvature_dev(ctx, metric, (const double*)d_x, n, what, (double*)d_out, nullptr));
    GEON_HIP(hipMemcpyDtoHAsync(out, (hipDeviceptr_t)d_out, n * w * 8, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return GEON_OK;
}

int geon_nec_contraction_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa,
                             size_t n, const double k[4], double* out, void* stream) {
    if (!ctx || !metric || !x_soa || !k || !out) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    geon_lane* lane;
    GEON_TRY(ctx_lane(ctx, s, &lane));
    void* d_t;
    GEON_TRY(lane_buffer(lane, "nec_t", n * 16 * 8, &d_t));
    GEON_TRY(geon_curvature_dev(ctx, metric, x_soa, n, GEON_CURV_STRESS, (double*)d_t, s));
    uint64_t nn = n;
    double k0 = k[0], k1 = k[1], k2 = k[2], k3 = k[3];
    void* args[] = {&d_t, &nn, &k0, &k1, &k2, &k3, &out};
    return launch(ctx->fn_nec, nn, 256, s, args);
}

int geon_nec_contraction(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa, size_t n,
                         const double k[4], double* out) {
    if (!ctx || !metric || !x_soa || !k || !out) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    hipSetDevice(ctx->device);
    void *d_x, *d_out;
    GEON_TRY(ctx_buffer(ctx, "nec_x", n * 32, &d_x));
    GEON_TRY(ctx_buffer(ctx, "nec_out", n * 8, &d_out));
    hipStream_t s = ctx->stream;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_x, (void*)x_soa, n * 32, s), "H2D");
    GEON_TRY(geon_nec_contraction_dev(ctx, metric, (const double*)d_x, n, k, (double*)d_out, nullptr));
    GEON_HIP(hipMemcpyDtoHAsync(out, (hipDeviceptr_t)d_out, n * 8, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return GEON_OK;
}

// The energy-condition line scans (energy_conditions.rs:70-153): the
// contraction at each point, then, in point order, the first strict
// minimum and the sum of nec·dx over the negative values.
int geon_nec_scan(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa, size_t n,
                  const double k[4], double dx, double* nec_out, double* min_value,
                  uint64_t* min_index, double* neg_integral) {
    if (!ctx || !metric || !x_soa || !k || !min_value || !min_index || !neg_integral)
        return fail(GEON_E_INVALID, "bad argument");
    std::vector<double> v(n);
    GEON_TRY(geon_nec_contraction(ctx, metric, x_soa, n, k, v.data()));
    double mv = INFINITY, integral = 0.0;
    uint64_t mi = 0;
    for (size_t i = 0; i < n; ++i) {
        if (v[i] < mv) {
            mv = v[i];
            mi = i;
        }
        if (v[i] < 0.0) integral += v[i] * dx;
    }
    if (nec_out) std::copy(v.begin(), v.end(), nec_out);
    *min_value = mv;
    *min_index = mi;
    *neg_integral = integral;
    return GEON_OK;
}

int geon_shadow_bisect(geon_ctx_t* ctx, const double* mass, const double* r_obs, size_t n,
                       double step, uint32_t max_steps, uint32_t iters, double* alpha_c,
                       double* b_c) {
    if (!ctx || !mass || !r_obs || !alpha_c || !b_c) return fail(GEON_E_INVALID, "bad argument");
    if (n == 0) return GEON_OK;
    for (size_t k = 0; k < n; ++k)
        if (!(mass[k] > 0.0) || !(r_obs[k] > 2.0 * mass[k]))
            return fail(GEON_E_INVALID, "problem %zu: need mass > 0 and r_obs > 2M", k);
    hipSetDevice(ctx->device);
    void *d_m, *d_r, *d_a, *d_b;
    GEON_TRY(ctx_buffer(ctx, "bs_m", n * 8, &d_m));
    GEON_TRY(ctx_buffer(ctx, "bs_r", n * 8, &d_r));
    GEON_TRY(ctx_buffer(ctx, "bs_a", n * 8, &d_a));
    GEON_TRY(ctx_buffer(ctx, "bs_b", n * 8, &d_b));
    hipStream_t s = ctx->stream;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_m, (void*)mass, n * 8, s), "H2D");
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)d_r, (void*)r_obs, n * 8, s), "H2D");
    uint64_t nn = n;
    void* args[] = {&nn, &d_m, &d_r, &step, &max_steps, &iters, &d_a, &d_b};
    GEON_TRY(launch(ctx->fn_bisect, nn, 64, s, args));
    GEON_HIP(hipMemcpyDtoHAsync(alpha_c, (hipDeviceptr_t)d_a, n * 8, s), "D2H");
    GEON_HIP(hipMemcpyDtoHAsync(b_c, (hipDeviceptr_t)d_b, n * 8, s), "D2H");
    GEON_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return GEON_OK;
}

// write_png's tonemap (render.rs:175-179): gamma_inv is an f32.
int geon_tonemap_u8(const float* rgb, size_t n, uint8_t* out) {
    if ((!rgb || !out) && n) return fail(GEON_E_INVALID, "bad argument");
    const float gamma_inv = 1.0f / 2.2f;
    for (size_t i = 0; i < n; ++i) {
        float c = rgb[i];
        if (c < 0.0f) c = 0.0f;
        if (c > 1.0f) c = 1.0f;
        const float q = roundf(powf(c, gamma_inv) * 255.0f);
        out[i] = (q <= 0.0f || q != q) ? 0 : (q >= 255.0f ? 255 : (uint8_t)q);
    }
    return GEON_OK;
}

}  // extern "C"
#pragma GCC visibility pop
