//! Raw bindings to `include/geon_hip.h` (libgeon_hip.so): the structures
//! with their C layout and every `extern "C"` entry point. Status codes are
//! returned as `i32` (0 = `GEON_OK`); see `geon-hip` for the safe layer with
//! the reference crate's types.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub const GEON_OK: i32 = 0;
pub const GEON_E_INVALID: i32 = -1;
pub const GEON_E_NOT_TIMELIKE: i32 = -2;
pub const GEON_E_GRAM_SCHMIDT: i32 = -3;
pub const GEON_E_SINGULAR: i32 = -4;
pub const GEON_E_HIP: i32 = -5;
pub const GEON_E_NOMEM: i32 = -6;
pub const GEON_E_ADMISSION: i32 = -7;
pub const GEON_E_NO_MODULE: i32 = -8;
pub const GEON_E_UNSUPPORTED: i32 = -9;
pub const GEON_E_NOT_FOUND: i32 = -10;
pub const GEON_E_PENDING: i32 = -11;

pub const GEON_METRIC_MINKOWSKI: i32 = 0;
pub const GEON_METRIC_SCHWARZSCHILD: i32 = 1;
pub const GEON_METRIC_ALCUBIERRE: i32 = 2;
pub const GEON_METRIC_MORRIS_THORNE: i32 = 3;

pub const GEON_PREC_FP64: i32 = 0;
pub const GEON_PREC_MIXED: i32 = 1;

pub const GEON_DIVERGED: i32 = 0;
pub const GEON_CAPTURED: i32 = 1;
pub const GEON_ESCAPED: i32 = 2;
pub const GEON_MAXSTEPS: i32 = 3;

pub const GEON_CURV_RIEMANN: i32 = 0;
pub const GEON_CURV_RICCI: i32 = 1;
pub const GEON_CURV_EINSTEIN: i32 = 2;
pub const GEON_CURV_STRESS: i32 = 3;
pub const GEON_CURV_SCALAR: i32 = 4;

pub const GEON_REQ_RENDER: i32 = 0;
pub const GEON_REQ_BISECT: i32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_metric_t {
    pub kind: i32,
    pub reserved: i32,
    pub p: [f64; 6],
    pub horizon_pad: f64,
    pub escape_radius: f64,
    pub fd_step: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_camera_t {
    pub observer: [f64; 4],
    pub fov_y_rad: f64,
    pub width: u32,
    pub height: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_render_cfg_t {
    pub step: f64,
    pub max_steps: u32,
    pub samples_per_axis: u32,
    pub captured_rgb: [f32; 3],
    pub diverged_rgb: [f32; 3],
    pub show_stars: u32,
    pub precision: i32,
    pub fp64_band: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_render_stats_t {
    pub captured: u64,
    pub escaped: u64,
    pub diverged: u64,
    pub max_null_residual: f64,
    pub avg_steps: f64,
    pub max_steps_hit: u64,
    pub ray_steps: u64,
    pub fp64_rays: u64,
    pub kernel_ms: f64,
    pub gamma_steps: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_sched_cfg_t {
    pub block: u32,
    pub waves_per_simd: u32,
    pub rays_per_lane: u32,
    pub refill: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_serve_cfg_t {
    pub hbm_budget: u64,
    pub quantum_steps: u32,
    pub max_jobs: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct geon_serve_result_t {
    pub kind: i32,
    pub status: i32,
    pub latency_ms: f64,
    pub queue_ms: f64,
    pub stats: geon_render_stats_t,
    pub ray_steps: u64,
}

/// Opaque handles.
#[repr(C)]
pub struct geon_ctx_t {
    _private: [u8; 0],
}
#[repr(C)]
pub struct geon_batch_t {
    _private: [u8; 0],
}
#[repr(C)]
pub struct geon_server_t {
    _private: [u8; 0],
}

extern "C" {
    pub fn geon_strerror(status: c_int) -> *const c_char;
    pub fn geon_last_error() -> *const c_char;
    pub fn geon_version() -> c_int;

    pub fn geon_metric_default(kind: i32, out: *mut geon_metric_t) -> c_int;
    pub fn geon_metric_g(m: *const geon_metric_t, x: *const f64, g_out: *mut f64) -> c_int;
    pub fn geon_metric_dg(m: *const geon_metric_t, x: *const f64, dg_out: *mut f64) -> c_int;
    pub fn geon_metric_g_inv(m: *const geon_metric_t, x: *const f64, ginv_out: *mut f64) -> c_int;
    pub fn geon_metric_christoffel(m: *const geon_metric_t, x: *const f64, gamma_out: *mut f64)
        -> c_int;
    pub fn geon_metric_terminate(m: *const geon_metric_t, x: *const f64, cls_out: *mut i32) -> c_int;
    pub fn geon_metric_asymptotic_direction(m: *const geon_metric_t, x: *const f64, p: *const f64,
                                            dir_out: *mut f64) -> c_int;
    pub fn geon_sky_sample(dir: *const f64, show_stars: i32, rgb_out: *mut f32) -> c_int;
    pub fn geon_camera_new(observer: *const f64, fov_y_deg: f64, width: u32, height: u32,
                           out: *mut geon_camera_t) -> c_int;

    pub fn geon_ctx_create(device: c_int, code_object: *const c_char, out: *mut *mut geon_ctx_t)
        -> c_int;
    pub fn geon_ctx_destroy(ctx: *mut geon_ctx_t);

    pub fn geon_tetrad(m: *const geon_metric_t, observer: *const f64, tetrad_out: *mut f64) -> c_int;
    pub fn geon_ray_state(cam: *const geon_camera_t, tetrad: *const f64, i: u32, j: u32, du: f64,
                          dv: f64, y_out: *mut f64) -> c_int;

    pub fn geon_trace_rays(ctx: *mut geon_ctx_t, m: *const geon_metric_t, y0_soa: *const f64,
                           n: usize, step: f64, max_steps: u32, cls: *mut i32, steps: *mut u32,
                           y_final_soa: *mut f64, null_resid: *mut f64, dir_soa: *mut f64) -> c_int;
    pub fn geon_trace_rays_dev(ctx: *mut geon_ctx_t, m: *const geon_metric_t, y0_soa: *const f64,
                               n: usize, step: f64, max_steps: u32, cls: *mut i32, steps: *mut u32,
                               y_final_soa: *mut f64, null_resid: *mut f64, dir_soa: *mut f64,
                               stream: *mut c_void) -> c_int;
    pub fn geon_trace_rays_adaptive(ctx: *mut geon_ctx_t, m: *const geon_metric_t,
                                    y0_soa: *const f64, n: usize, h0: f64, rtol: f64, atol: f64,
                                    max_steps: u32, cls: *mut i32, steps: *mut u32,
                                    rejected: *mut u32, y_final_soa: *mut f64, dir_soa: *mut f64)
        -> c_int;
    pub fn geon_trace_rays_adaptive_dev(ctx: *mut geon_ctx_t, m: *const geon_metric_t,
                                        y0_soa: *const f64, n: usize, h0: f64, rtol: f64, atol: f64,
                                        max_steps: u32, cls: *mut i32, steps: *mut u32,
                                        rejected: *mut u32, y_final_soa: *mut f64,
                                        dir_soa: *mut f64, stream: *mut c_void) -> c_int;

    pub fn geon_render(ctx: *mut geon_ctx_t, m: *const geon_metric_t, cam: *const geon_camera_t,
                       cfg: *const geon_render_cfg_t, rgb: *mut f32, cls_per_pixel: *mut i32,
                       dir_per_pixel: *mut f64, stats: *mut geon_render_stats_t) -> c_int;
    pub fn geon_render_dev(ctx: *mut geon_ctx_t, m: *const geon_metric_t, cam: *const geon_camera_t,
                           cfg: *const geon_render_cfg_t, rgb: *mut f32, cls_per_pixel: *mut i32,
                           dir_per_pixel: *mut f64, stats: *mut geon_render_stats_t,
                           stream: *mut c_void) -> c_int;
    pub fn geon_render_dev_noalloc(ctx: *mut geon_ctx_t, m: *const geon_metric_t,
                                   cam: *const geon_camera_t, cfg: *const geon_render_cfg_t,
                                   rgb: *mut f32, cls_per_pixel: *mut i32, dir_per_pixel: *mut f64,
                                   stats: *mut geon_render_stats_t, stream: *mut c_void) -> c_int;
    pub fn geon_render_dev_noalloc_max_rays() -> u64;

    pub fn geon_curvature(ctx: *mut geon_ctx_t, m: *const geon_metric_t, x_soa: *const f64,
                          n: usize, what: i32, out: *mut f64) -> c_int;
    pub fn geon_curvature_dev(ctx: *mut geon_ctx_t, m: *const geon_metric_t, x_soa: *const f64,
                              n: usize, what: i32, out: *mut f64, stream: *mut c_void) -> c_int;
    pub fn geon_nec_contraction(ctx: *mut geon_ctx_t, m: *const geon_metric_t, x_soa: *const f64,
                                n: usize, k: *const f64, out: *mut f64) -> c_int;
    pub fn geon_nec_contraction_dev(ctx: *mut geon_ctx_t, m: *const geon_metric_t,
                                    x_soa: *const f64, n: usize, k: *const f64, out: *mut f64,
                                    stream: *mut c_void) -> c_int;
    pub fn geon_nec_scan(ctx: *mut geon_ctx_t, m: *const geon_metric_t, x_soa: *const f64, n: usize,
                         k: *const f64, dx: f64, nec_out: *mut f64, min_value: *mut f64,
                         min_index: *mut u64, neg_integral: *mut f64) -> c_int;
    pub fn geon_shadow_bisect(ctx: *mut geon_ctx_t, mass: *const f64, r_obs: *const f64, n: usize,
                              step: f64, max_steps: u32, iters: u32, alpha_c: *mut f64,
                              b_c: *mut f64) -> c_int;

    pub fn geon_batch_create(ctx: *mut geon_ctx_t, out: *mut *mut geon_batch_t) -> c_int;
    pub fn geon_batch_destroy(batch: *mut geon_batch_t);
    pub fn geon_batch_add_frame(batch: *mut geon_batch_t, m: *const geon_metric_t,
                                cam: *const geon_camera_t, cfg: *const geon_render_cfg_t,
                                rgb: *mut f32, cls_per_pixel: *mut i32, dir_per_pixel: *mut f64,
                                frame_id: *mut u32) -> c_int;
    pub fn geon_batch_add_rays(batch: *mut geon_batch_t, m: *const geon_metric_t,
                               y0_soa: *const f64, n: usize, step: f64, max_steps: u32,
                               precision: i32, early_exit: i32, cls: *mut i32, steps: *mut u32,
                               y_final_soa: *mut f64, null_resid: *mut f64, dir_soa: *mut f64,
                               job_id: *mut u32) -> c_int;
    pub fn geon_batch_run(batch: *mut geon_batch_t, tune: *const geon_sched_cfg_t,
                          stream: *mut c_void) -> c_int;
    pub fn geon_batch_frame_stats(batch: *mut geon_batch_t, frame_id: u32,
                                  out: *mut geon_render_stats_t) -> c_int;
    pub fn geon_batch_totals(batch: *mut geon_batch_t, ray_steps: *mut u64, fp64_rays: *mut u64,
                             device_ms: *mut f64) -> c_int;
    pub fn geon_autotune(ctx: *mut geon_ctx_t, metric_kind: i32, precision: i32,
                         best: *mut geon_sched_cfg_t) -> c_int;

    pub fn geon_serve_create(ctx: *mut geon_ctx_t, cfg: *const geon_serve_cfg_t,
                             out: *mut *mut geon_server_t) -> c_int;
    pub fn geon_serve_destroy(srv: *mut geon_server_t);
    pub fn geon_serve_submit_render(srv: *mut geon_server_t, m: *const geon_metric_t,
                                    cam: *const geon_camera_t, cfg: *const geon_render_cfg_t,
                                    keep_outputs: i32, ticket: *mut u64) -> c_int;
    pub fn geon_serve_submit_bisect(srv: *mut geon_server_t, mass: *const f64, r_obs: *const f64,
                                    n: usize, step: f64, max_steps: u32, iters: u32,
                                    full_trace: i32, ticket: *mut u64) -> c_int;
    pub fn geon_serve_poll(srv: *mut geon_server_t, ticket: u64, out: *mut geon_serve_result_t)
        -> c_int;
    pub fn geon_serve_fetch_bisect(srv: *mut geon_server_t, ticket: u64, alpha_c: *mut f64,
                                   b_c: *mut f64) -> c_int;
    pub fn geon_serve_fetch_render(srv: *mut geon_server_t, ticket: u64, rgb: *mut f32,
                                   cls_per_pixel: *mut i32, dir_per_pixel: *mut f64) -> c_int;
    pub fn geon_serve_release(srv: *mut geon_server_t, ticket: u64) -> c_int;
    pub fn geon_serve_drain(srv: *mut geon_server_t) -> c_int;
    pub fn geon_serve_counters(srv: *mut geon_server_t, quanta: *mut u64, done: *mut u64,
                               ray_steps: *mut u64, busy_ms: *mut f64) -> c_int;

    pub fn geon_tonemap_u8(rgb: *const f32, n: usize, out: *mut u8) -> c_int;
}
