/*
 * geon_hip.h — C ABI of the MI355X (gfx950) geodesic engine.
 *
 * Plain C: fixed-width integers, doubles and opaque pointers only. Every
 * entry point returns an int status (GEON_OK == 0, negative on error) and
 * never throws or aborts across the boundary.
 *
 * Reference interfaces each entry point replaces (paths in Arth-Singh/geon):
 *   geon_metric_default      Schwarzschild::new / Alcubierre::default /
 *                            MorrisThorne::default / Minkowski
 *                            (src/metric/schwarzschild.rs:23-29,
 *                             src/metric/alcubierre.rs:21-30,
 *                             src/metric/morris_thorne.rs:14-19,
 *                             src/metric/minkowski.rs:8)
 *   geon_metric_{g,dg,g_inv,christoffel,terminate,asymptotic_direction}
 *                            the Metric trait's methods on the host
 *                            (src/metric/mod.rs:19-109)
 *   geon_camera_new          Camera::new               (src/camera.rs:29-36)
 *   geon_tetrad              Camera::tetrad            (src/camera.rs:47-97)
 *   geon_ray_state           Camera::ray_state_sub     (src/camera.rs:113-160)
 *   geon_trace_rays[_dev]    geodesic::trace, batched  (src/geodesic.rs:62-117)
 *   geon_trace_rays_adaptive[_dev] no counterpart: the adaptive stepping the
 *                            reference names as its next step
 *                            (src/integrator.rs:5-7), same trace semantics
 *   geon_render[_dev]        render::render            (src/render.rs:55-165)
 *   geon_curvature[_dev]     curvature::{riemann,ricci,ricci_scalar,einstein,
 *                            stress_energy}            (src/curvature.rs:48-126)
 *   geon_nec_contraction[_dev] curvature::nec_contraction (src/curvature.rs:127-136)
 *   geon_nec_scan            the line scans of bin/energy_conditions.rs:70-153
 *                            (strict minimum, ∫ of the negative part)
 *   geon_shadow_bisect       bin/validate_shadow.rs:20-95 (is_captured +
 *                            bisection on the launch angle)
 *   geon_tonemap_u8          render::write_png tonemap (src/render.rs:169-180)
 *   geon_sky_sample          sky::sample               (src/sky.rs:107-120)
 *
 * Memory: "_dev" entry points take device pointers (HBM) and a hipStream_t
 * passed as void* (NULL: the context's own stream); they enqueue work and
 * return without synchronising the host, except where a host output
 * (statistics) is requested, which waits for that stream only. Scratch
 * memory the context keeps for them is held per stream, so calls on
 * different streams may overlap on the device. The other entry points take
 * caller-owned host buffers, copy in and out, and synchronise. The context
 * owns every device allocation it makes.
 *
 * Threading: a context is bound to one device and used by one host thread
 * at a time. Separate contexts may be used concurrently.
 */
#ifndef GEON_HIP_H
#define GEON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define GEON_OK                 0
#define GEON_E_INVALID         -1  /* bad argument / size                          */
#define GEON_E_NOT_TIMELIKE    -2  /* camera.rs:53  observer has g_tt >= 0          */
#define GEON_E_GRAM_SCHMIDT    -3  /* camera.rs:85  non-spacelike tetrad leg        */
#define GEON_E_SINGULAR        -4  /* types.rs:39   metric not invertible           */
#define GEON_E_HIP             -5  /* a HIP runtime call failed                     */
#define GEON_E_NOMEM           -6  /* device allocation failed                      */
#define GEON_E_ADMISSION       -7  /* request does not fit the HBM budget           */
#define GEON_E_NO_MODULE       -8  /* gfx950 code object missing / not loadable     */
#define GEON_E_UNSUPPORTED     -9  /* combination not provided by this build        */
#define GEON_E_NOT_FOUND      -10  /* unknown ticket                                */
#define GEON_E_PENDING        -11  /* request still running                         */

/* ---- metric ---------------------------------------------------------- */
#define GEON_METRIC_MINKOWSKI      0
#define GEON_METRIC_SCHWARZSCHILD  1
#define GEON_METRIC_ALCUBIERRE     2
#define GEON_METRIC_MORRIS_THORNE  3

/*
 * kind-specific parameters p[]:
 *   SCHWARZSCHILD  p[0] = mass
 *   ALCUBIERRE     p[0] = bubble_radius R, p[1] = bubble_thickness s,
 *                  p[2] = ship_velocity v, p[3] = ship_x
 *   MORRIS_THORNE  p[0] = throat_radius b0
 * Termination (metric::terminate of each reference metric):
 *   SCHWARZSCHILD  captured r < horizon_pad*2M, escaped r > escape_radius*M
 *   MINKOWSKI,
 *   ALCUBIERRE     escaped x²+y²+z² > escape_radius²
 *   MORRIS_THORNE  escaped |l| > escape_radius
 * Reference defaults: horizon_pad 1.001, escape_radius 1000, fd_step 1e-5.
 */
typedef struct geon_metric {
    int32_t kind;
    int32_t reserved;
    double  p[6];
    double  horizon_pad;
    double  escape_radius;
    double  fd_step;
} geon_metric_t;

/* Camera (camera.rs:21-26): the vertical field of view in radians, as the
 * reference stores it. Camera::new(observer, fov_y_deg, width, height)
 * (camera.rs:29-36) is geon_camera_new. */
typedef struct geon_camera {
    double   observer[4];
    double   fov_y_rad;
    uint32_t width;
    uint32_t height;
} geon_camera_t;

#define GEON_PREC_FP64   0   /* plain FP64 RK4                                   */
#define GEON_PREC_MIXED  1   /* FP32 RK4 + null re-projection, FP64 for flagged  */

/* RenderConfig (render.rs:17-41) plus the sky toggle (sky.rs:23) */
typedef struct geon_render_cfg {
    double   step;
    uint32_t max_steps;
    uint32_t samples_per_axis;
    float    captured_rgb[3];
    float    diverged_rgb[3];
    uint32_t show_stars;
    int32_t  precision;
    /* FP32-mixed: rays whose impact parameter is within this distance of the
     * photon orbit (3√3 M for Schwarzschild, b0 for Morris–Thorne) are
     * integrated in FP64; 0 selects 1e-3 of the mass (throat radius). */
    double   fp64_band;
} geon_render_cfg_t;

/* RenderStats (render.rs:45-51) plus counters the reference does not keep */
typedef struct geon_render_stats {
    uint64_t captured;          /* pixels, reference majority rule            */
    uint64_t escaped;
    uint64_t diverged;          /* includes max-steps pixels (reference rule) */
    double   max_null_residual;
    double   avg_steps;         /* mean over pixels of per-pixel integer avg  */
    uint64_t max_steps_hit;     /* rays that exhausted max_steps              */
    uint64_t ray_steps;         /* sum of steps_taken over every traced ray   */
    uint64_t fp64_rays;         /* rays integrated in FP64 (mixed precision)  */
    double   kernel_ms;         /* device time of the integration kernels     */
    uint64_t gamma_steps;       /* of ray_steps, the steps that evaluated Γ at
                                 * their stages (the rest are steps through
                                 * regions where Γ is exactly zero: Minkowski,
                                 * Alcubierre beyond its walls); plain FP64
                                 * renders only, 0 elsewhere                  */
} geon_render_stats_t;

/* Per-ray outcome codes. The reference's Termination has three variants;
 * max-steps is its Diverged with steps_taken == max_steps (geodesic.rs:107-116). */
#define GEON_DIVERGED  0
#define GEON_CAPTURED  1
#define GEON_ESCAPED   2
#define GEON_MAXSTEPS  3

/* curvature outputs (row-major, per point) */
#define GEON_CURV_RIEMANN   0   /* 256 doubles R^ρ_σμν as [ρ][σ][μ][ν]  */
#define GEON_CURV_RICCI     1   /* 16 doubles R_μν                       */
#define GEON_CURV_EINSTEIN  2   /* 16 doubles G_μν                       */
#define GEON_CURV_STRESS    3   /* 16 doubles T_μν = G_μν/(8π)           */
#define GEON_CURV_SCALAR    4   /* 1 double  R                           */

typedef struct geon_ctx geon_ctx_t;

const char* geon_strerror(int status);
/* Detail of the last failure on the calling thread (empty if none). */
const char* geon_last_error(void);
int  geon_version(void);

int  geon_metric_default(int32_t kind, geon_metric_t* out);

/* ---- the Metric trait on the host (src/metric/mod.rs:19-109) ------------
 * Host evaluations of one metric at one point, so that a binding can
 * implement the reference's trait for the metrics this library runs on the
 * GPU. Matrices are row-major; dg is [α][μ][ν], Γ is [μ][α][β]. Each follows
 * the reference's definition for that metric: analytic g⁻¹ and Γ for
 * Schwarzschild (schwarzschild.rs:47-107), zero ∂g and Γ for Minkowski
 * (minkowski.rs:20-30), and the trait defaults elsewhere: central
 * differences of g with fd_step (mod.rs:25-43), the cofactor inverse
 * (mod.rs:46-48, types.rs:34-46; GEON_E_SINGULAR when g is not invertible)
 * and Γ = ½ g^μσ(∂_α g_σβ + ∂_β g_σα − ∂_σ g_αβ) (mod.rs:51-69). */
int  geon_metric_g(const geon_metric_t* metric, const double x[4], double g_out[16]);
int  geon_metric_dg(const geon_metric_t* metric, const double x[4], double dg_out[64]);
int  geon_metric_g_inv(const geon_metric_t* metric, const double x[4], double ginv_out[16]);
int  geon_metric_christoffel(const geon_metric_t* metric, const double x[4], double gamma_out[64]);
/* Metric::terminate (mod.rs:78-81 and each metric's override): *cls_out is
 * GEON_CAPTURED, GEON_ESCAPED, GEON_DIVERGED, or -1 (None: keep integrating). */
int  geon_metric_terminate(const geon_metric_t* metric, const double x[4], int32_t* cls_out);
/* Metric::asymptotic_direction (mod.rs:91-108; alcubierre.rs:81-88;
 * morris_thorne.rs:54-77): unit Cartesian direction of an escaped ray. */
int  geon_metric_asymptotic_direction(const geon_metric_t* metric, const double x[4],
                                      const double p[4], double dir_out[3]);

/* sky::sample (sky.rs:107-120) on the host: linear RGB of the procedural sky
 * in the unit Cartesian direction dir, stars on or off (sky.rs:23). */
int  geon_sky_sample(const double dir[3], int32_t show_stars, float rgb_out[3]);

/* Camera::new (camera.rs:29-36): fov_y_rad = fov_y_deg·(π/180), as
 * f64::to_radians computes it. */
int  geon_camera_new(const double observer[4], double fov_y_deg, uint32_t width,
                     uint32_t height, geon_camera_t* out);

/* code_object: path of geon_kernels.hsaco, or NULL for the copy installed
 * beside libgeon_hip.so. */
int  geon_ctx_create(int device, const char* code_object, geon_ctx_t** out);
void geon_ctx_destroy(geon_ctx_t* ctx);

/* Host-side Gram–Schmidt tetrad, rows e_(a)^μ, 16 doubles row-major. */
int  geon_tetrad(const geon_metric_t* metric, const double observer[4],
                 double tetrad_out[16]);

/* Camera pixel -> initial state (camera.rs:113-160), host side. */
int  geon_ray_state(const geon_camera_t* cam, const double tetrad[16],
                    uint32_t i, uint32_t j, double du, double dv, double y_out[8]);

/*
 * Trace n rays from initial states y0 (SoA: 8 planes of n doubles).
 * Outputs (any may be NULL): cls[n] (GEON_* codes), steps[n],
 * y_final (SoA 8×n), null_resid[n], dir (SoA 3×n, asymptotic direction of
 * escaped rays, zeros otherwise).
 */
int  geon_trace_rays(geon_ctx_t* ctx, const geon_metric_t* metric,
                     const double* y0_soa, size_t n, double step, uint32_t max_steps,
                     int32_t* cls, uint32_t* steps, double* y_final_soa,
                     double* null_resid, double* dir_soa);
int  geon_trace_rays_dev(geon_ctx_t* ctx, const geon_metric_t* metric,
                         const double* y0_soa, size_t n, double step, uint32_t max_steps,
                         int32_t* cls, uint32_t* steps, double* y_final_soa,
                         double* null_resid, double* dir_soa, void* stream);

/*
 * Render a frame. rgb: 3·W·H floats (row-major, top-left origin, as
 * render.rs:146-152); cls_per_pixel: W·H GEON_* codes of the pixel's
 * representative class (render.rs:129-135; GEON_MAXSTEPS when every
 * diverged sub-ray ran out of steps); dir_per_pixel: SoA 3×(W·H), the
 * asymptotic direction of the pixel's first escaped sub-ray. Any output
 * pointer may be NULL.
 */
int  geon_render(geon_ctx_t* ctx, const geon_metric_t* metric, const geon_camera_t* cam,
                 const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                 double* dir_per_pixel, geon_render_stats_t* stats);
/* Device variant: outputs are device pointers; stats (host) is filled after
 * an internal synchronisation of `stream` when non-NULL. */
int  geon_render_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const geon_camera_t* cam,
                     const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                     double* dir_per_pixel, geon_render_stats_t* stats, void* stream);
/* geon_render_dev without any allocation (no counterpart in the reference:
 * for callers that cannot allocate — code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). FP64 precision only
 * (GEON_E_UNSUPPORTED otherwise); at most geon_render_dev_noalloc_max_rays()
 * rays = W·H·samples² per call (4,194,304, e.g. one 2048² frame; larger:
 * GEON_E_INVALID). The per-ray planes and statistics live in one fixed
 * device array per GPU (168 MiB, held until the process ends) shared by
 * every call on that GPU: calls that overlap on one GPU race, so order them
 * on one stream. Results are those of geon_render_dev; stats->kernel_ms is
 * 0 (no events). */
int  geon_render_dev_noalloc(geon_ctx_t* ctx, const geon_metric_t* metric,
                             const geon_camera_t* cam, const geon_render_cfg_t* cfg, float* rgb,
                             int32_t* cls_per_pixel, double* dir_per_pixel,
                             geon_render_stats_t* stats, void* stream);
uint64_t geon_render_dev_noalloc_max_rays(void);

/* geon_trace_rays with adaptive steps: Cash–Karp RK4(5), a step accepted
 * when max_i |err_i| / (atol + rtol·max(|y_i|, |y_new_i|)) ≤ 1, next step
 * h·0.9·e^(−1/4) in [h/10, 5h], |h| within [1e-8, 8]·|h0| (below: DIVERGED).
 * A trial step with a non-finite component is rejected with h/10.
 * max_steps counts accepted steps; rejected (optional) the rejected attempts.
 * h0 ≠ 0, rtol ≥ 0, atol > 0. Outputs as geon_trace_rays (no residual). */
int  geon_trace_rays_adaptive(geon_ctx_t* ctx, const geon_metric_t* metric, const double* y0_soa,
                              size_t n, double h0, double rtol, double atol, uint32_t max_steps,
                              int32_t* cls, uint32_t* steps, uint32_t* rejected,
                              double* y_final_soa, double* dir_soa);
int  geon_trace_rays_adaptive_dev(geon_ctx_t* ctx, const geon_metric_t* metric,
                                  const double* y0_soa, size_t n, double h0, double rtol,
                                  double atol, uint32_t max_steps, int32_t* cls, uint32_t* steps,
                                  uint32_t* rejected, double* y_final_soa, double* dir_soa,
                                  void* stream);

/* Curvature at n points x (SoA 4×n). out: n×{256|16|1} doubles. */
int  geon_curvature(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa,
                    size_t n, int32_t what, double* out);
int  geon_curvature_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa,
                        size_t n, int32_t what, double* out, void* stream);

/* T_μν k^μ k^ν at n points x (SoA 4×n) for one vector k, summed over μ then
 * ν as in curvature::nec_contraction (src/curvature.rs:127-136); T_μν is
 * GEON_CURV_STRESS. out: n doubles. */
int  geon_nec_contraction(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa,
                          size_t n, const double k[4], double* out);
int  geon_nec_contraction_dev(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa,
                              size_t n, const double k[4], double* out, void* stream);
/* The energy-condition line scans (src/bin/energy_conditions.rs:70-153): the
 * contraction at the n points (nec_out, optional), the first strict minimum
 * in point order (min_value, min_index; +inf and 0 for n = 0) and Σ nec·dx
 * over the points where nec < 0, accumulated in point order (the ∫ T_kk dx
 * proxy of :105-118). */
int  geon_nec_scan(geon_ctx_t* ctx, const geon_metric_t* metric, const double* x_soa, size_t n,
                   const double k[4], double dx, double* nec_out, double* min_value,
                   uint64_t* min_index, double* neg_integral);

/*
 * Shadow bisection (validate_shadow.rs): for each of n independent
 * problems k (Schwarzschild mass[k], observer radius r_obs[k]), bisect
 * `iters` times on the equatorial launch angle α ∈ [0, π/2 − 1e-3] using
 * trace(step, max_steps) capture as the predicate, and return
 * alpha_c[k] = ½(lo+hi) and b_c[k] = r_obs·sin α_c/√(1−2M/r_obs).
 */
int  geon_shadow_bisect(geon_ctx_t* ctx, const double* mass, const double* r_obs,
                        size_t n, double step, uint32_t max_steps, uint32_t iters,
                        double* alpha_c, double* b_c);

/* ---- persistent scheduler (no reference counterpart: the reference runs one
 * rayon task per pixel, render.rs:73-75; these batch many frames and ray sets
 * of ragged length into persistent gfx950 launches) ------------------------ */

/* Launch shape of a scheduler kernel; zero fields take the tuned defaults. */
typedef struct geon_sched_cfg {
    uint32_t block;           /* threads per workgroup (64, 128 or 256)        */
    uint32_t waves_per_simd;  /* resident waves per SIMD the grid is sized for */
    uint32_t rays_per_lane;   /* 1 or 2 rays integrated per lane               */
    uint32_t refill;          /* RK4 steps between refill points (1..65536)    */
} geon_sched_cfg_t;

typedef struct geon_batch geon_batch_t;

int  geon_batch_create(geon_ctx_t* ctx, geon_batch_t** out);
void geon_batch_destroy(geon_batch_t* batch);
/* Add a camera frame (render semantics). Outputs are device pointers, any may
 * be NULL; layout as geon_render. cfg->precision selects FP64 or FP32-mixed. */
int  geon_batch_add_frame(geon_batch_t* batch, const geon_metric_t* metric,
                          const geon_camera_t* cam, const geon_render_cfg_t* cfg,
                          float* rgb, int32_t* cls_per_pixel, double* dir_per_pixel,
                          uint32_t* frame_id);
/* Add n explicit rays (device SoA 8×n initial states, trace semantics).
 * Outputs are device pointers, any may be NULL. early_exit (Schwarzschild
 * only) ends a ray as soon as it is outgoing beyond r = 3M (escaped) or
 * ingoing inside it (captured), which decides capture without the tail. */
int  geon_batch_add_rays(geon_batch_t* batch, const geon_metric_t* metric,
                         const double* y0_soa, size_t n, double step, uint32_t max_steps,
                         int32_t precision, int32_t early_exit, int32_t* cls, uint32_t* steps,
                         double* y_final_soa, double* null_resid, double* dir_soa,
                         uint32_t* job_id);
/* Enqueue every added job on `stream` (NULL: the context stream); metric
 * kinds run concurrently on forked streams joined back into `stream`. */
int  geon_batch_run(geon_batch_t* batch, const geon_sched_cfg_t* tune, void* stream);
/* Statistics of a frame after the run (synchronises). */
int  geon_batch_frame_stats(geon_batch_t* batch, uint32_t frame_id, geon_render_stats_t* out);
/* Totals after the run (synchronises): executed RK4 ray-steps, rays re-traced
 * in FP64, device time between the run's first and last kernel. */
int  geon_batch_totals(geon_batch_t* batch, uint64_t* ray_steps, uint64_t* fp64_rays,
                       double* device_ms);
/* Time candidate launch shapes on a calibration frame of this metric kind
 * and precision; the winner becomes the context default and is returned. */
int  geon_autotune(geon_ctx_t* ctx, int32_t metric_kind, int32_t precision,
                   geon_sched_cfg_t* best);

/* ---- serving (no reference counterpart: the reference renders one frame per
 * process, bin/render.rs, and bisects serially, bin/validate_shadow.rs) ------
 * A server owns a host thread that admits requests as HBM frees up and runs
 * every admitted request's rays together in step-bounded quanta of the
 * persistent scheduler: requests that arrive while others are on the GPU join
 * the rays already in flight. */
typedef struct geon_serve_cfg {
    uint64_t hbm_budget;       /* bytes of request buffers in flight (0: 8 GiB)       */
    uint32_t quantum_steps;    /* RK4 steps per lane per quantum (0: 16384)           */
    uint32_t max_jobs;         /* job-table capacity over the server's life (0: 16384) */
} geon_serve_cfg_t;

#define GEON_REQ_RENDER  0
#define GEON_REQ_BISECT  1

typedef struct geon_serve_result {
    int32_t  kind;             /* GEON_REQ_*                                           */
    int32_t  status;           /* GEON_OK, or the error that ended the request         */
    double   latency_ms;       /* submit → result available                            */
    double   queue_ms;         /* submit → admitted                                    */
    geon_render_stats_t stats; /* render requests                                      */
    uint64_t ray_steps;        /* Σ steps_taken of the request's final rays           */
} geon_serve_result_t;

typedef struct geon_server geon_server_t;

int  geon_serve_create(geon_ctx_t* ctx, const geon_serve_cfg_t* cfg, geon_server_t** out);
/* Stops the host thread after the requests in flight finish, then frees. */
void geon_serve_destroy(geon_server_t* srv);
/* Render request; outputs are kept by the server when keep_outputs != 0 and
 * can be copied out with geon_serve_fetch_render before geon_serve_release. */
int  geon_serve_submit_render(geon_server_t* srv, const geon_metric_t* metric,
                              const geon_camera_t* cam, const geon_render_cfg_t* cfg,
                              int32_t keep_outputs, uint64_t* ticket);
/* Shadow bisection request: n problems (mass[k], r_obs[k]) of one mass,
 * `iters` halvings each with the capture predicate trace(step, max_steps) on
 * equatorial rays (validate_shadow.rs:20-79), run as parallel k-section
 * rounds that visit exactly the serial bisection's midpoints.
 * full_trace = 0 decides each probe at the photon sphere (outgoing beyond
 * 3M: escaped, ingoing inside: captured), which converges on 3√3 M; with
 * full_trace = 1 every probe is traced to capture or escape as the reference
 * does, reproducing its fixed-step result. */
int  geon_serve_submit_bisect(geon_server_t* srv, const double* mass, const double* r_obs,
                              size_t n, double step, uint32_t max_steps, uint32_t iters,
                              int32_t full_trace, uint64_t* ticket);
/* GEON_OK with the result when done, GEON_E_PENDING while running. */
int  geon_serve_poll(geon_server_t* srv, uint64_t ticket, geon_serve_result_t* out);
/* Bisection results (alpha_c, b_c: n doubles each, host). */
int  geon_serve_fetch_bisect(geon_server_t* srv, uint64_t ticket, double* alpha_c, double* b_c);
/* Render outputs (host buffers, any may be NULL; layout as geon_render). */
int  geon_serve_fetch_render(geon_server_t* srv, uint64_t ticket, float* rgb,
                             int32_t* cls_per_pixel, double* dir_per_pixel);
int  geon_serve_release(geon_server_t* srv, uint64_t ticket);
/* Blocks until every submitted request is done. */
int  geon_serve_drain(geon_server_t* srv);
/* Server-side counters: quanta launched, requests done, Σ final ray-steps,
 * device-busy milliseconds. */
int  geon_serve_counters(geon_server_t* srv, uint64_t* quanta, uint64_t* done,
                         uint64_t* ray_steps, double* busy_ms);

/* write_png's per-channel map: clamp(0,1)^(1/2.2)·255, rounded (render.rs:175-179). */
int  geon_tonemap_u8(const float* rgb, size_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif /* GEON_HIP_H */
