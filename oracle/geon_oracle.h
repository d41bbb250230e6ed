/*
 * geon_oracle.h — CPU restatement of the reference geodesic path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is
 * compared against (tests, smoke check, CPU baseline leg of the bench).
 * Nothing in the product links or calls it.
 *
 * Every function follows the Rust source of Arth-Singh/geon line by line
 * (operation order kept, no FMA contraction, glibc libm), with the file:line
 * it restates in the comment above it in geon_oracle.c.
 */
#ifndef GEON_ORACLE_H
#define GEON_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/geon_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_trace_result {
    int32_t  cls;          /* GEON_DIVERGED/CAPTURED/ESCAPED/MAXSTEPS */
    int32_t  reserved;
    uint64_t steps;
    double   y[8];
    double   null_residual;
} or_trace_result_t;

/* L0 types */
double or_dot4(const double g[16], const double a[4], const double b[4]);
int    or_inv4(const double m[16], double out[16]);

/* L1 metric */
/* Conditioning tests only: shift every tanh of the Alcubierre shape function
 * by one ulp towards ±∞ (k = ±1), or each evaluation up or down by one ulp
 * as a hash of its argument and k decides (|k| ≥ 2); 0 restores the
 * reference. Not thread-safe against traces in flight. */
void   or_set_tanh_ulp_shift(int k);
void   or_metric_default(int32_t kind, geon_metric_t* out);
void   or_g(const geon_metric_t* m, const double x[4], double g[16]);
void   or_dg(const geon_metric_t* m, const double x[4], double dg[64]);
int    or_g_inv(const geon_metric_t* m, const double x[4], double out[16]);
int    or_christoffel(const geon_metric_t* m, const double x[4], double gam[64]);
int    or_terminate(const geon_metric_t* m, const double x[4]);
void   or_asymptotic_direction(const geon_metric_t* m, const double x[4],
                               const double p[4], double out[3]);

/* L2 integration */
void   or_rhs(const geon_metric_t* m, const double y[8], double out[8]);
void   or_rk4_step(const geon_metric_t* m, const double y[8], double h, double out[8]);
void   or_trace(const geon_metric_t* m, const double y0[8], double step,
                uint64_t max_steps, or_trace_result_t* res);
/* Adaptive Cash–Karp RK4(5) trace of n rays (no reference counterpart) */
void   or_ck_step(const geon_metric_t* m, const double y[8], double h, double yn[8],
                  double err[8]);
void   or_trace_adaptive_rays(const geon_metric_t* m, const double* y0, size_t n, double h0,
                              double rtol, double atol, uint64_t max_steps, int32_t* cls,
                              uint64_t* steps, uint64_t* rejected, double* yf, double* dir);
/* Batch over n rays (AoS n×8 in, any output may be NULL), nthreads workers
 * with dynamic chunking. Returns elapsed wall seconds. */
double or_trace_rays(const geon_metric_t* m, const double* y0_aos, size_t n,
                     double step, uint64_t max_steps, int nthreads,
                     int32_t* cls, uint64_t* steps, double* yf_aos,
                     double* resid, double* dir_aos);

/* L3 camera */
void   or_camera_new(const double observer[4], double fov_y_deg, uint32_t width,
                     uint32_t height, geon_camera_t* out);
int    or_tetrad(const geon_metric_t* m, const double observer[4], double tet[16]);
void   or_ray_state_sub(const geon_camera_t* cam, const double tet[16], uint64_t i,
                        uint64_t j, double du, double dv, double y[8]);

/* L4 render + sky */
uint64_t or_sky_hash1(uint64_t x);
void   or_sky_sample(const double dir[3], int show_stars, float rgb[3]);
/* pixel outputs: rgb 3·W·H, cls W·H, dir 3·W·H (AoS, first escaped sub-ray),
 * steps_pix W·H (per-pixel integer average). Returns status. */
int    or_render(const geon_metric_t* m, const geon_camera_t* cam,
                 const geon_render_cfg_t* cfg, int nthreads, float* rgb,
                 int32_t* cls_pix, double* dir_pix, uint64_t* steps_pix,
                 geon_render_stats_t* stats);
void   or_tonemap_u8(const float* px, size_t n, uint8_t* out);

/* L2' curvature */
int    or_riemann(const geon_metric_t* m, const double x[4], double out[256]);
int    or_ricci(const geon_metric_t* m, const double x[4], double out[16]);
double or_ricci_scalar(const geon_metric_t* m, const double x[4]);
int    or_einstein(const geon_metric_t* m, const double x[4], double out[16]);
int    or_stress_energy(const geon_metric_t* m, const double x[4], double out[16]);
double or_nec_contraction(const geon_metric_t* m, const double x[4], const double k[4]);
/* batch: x AoS n×4, what = GEON_CURV_*, out n×{256,16,16,16,1}; threaded */
double or_curvature(const geon_metric_t* m, const double* x_aos, size_t n, int32_t what,
                    int nthreads, double* out);

/* harness */
void   or_equatorial_ray(double mass, double r_obs, double alpha, double y[8]);
int    or_is_captured(double mass, double r_obs, double alpha, double step, uint64_t max_steps);
void   or_validate_shadow(double mass, double r_obs, double step, uint64_t max_steps,
                          int iters, double* alpha_c, double* b_c);
/* Alcubierre ∫ T_kk⁻ dx along the axis (energy_conditions.rs:104-118) */
double or_alcubierre_nec_integral(const geon_metric_t* m);

#ifdef __cplusplus
}
#endif
#endif
