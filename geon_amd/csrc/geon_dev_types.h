// geon_dev_types.h — plain structures shared by the host runtime and the
// gfx950 kernels. Everything here is precomputed on the host (glibc libm,
// same association order as the reference) so that constants the reference
// recomputes per call, e.g. 2·tanh(R/s) or horizon_pad·2·M, carry the exact
// values the reference sees.
#pragma once
#include <stdint.h>

#define GEON_CONST_SLOTS 32   // ring of __constant__ frame slots

struct geon_dmetric {
    int32_t kind;
    int32_t s_pow2;      // Alcubierre: thickness is a power of two → x/s == x*inv_s exactly
    double  M;           // Schwarzschild mass
    double  cap_r;       // horizon_pad * 2.0 * M             (schwarzschild.rs:114)
    double  esc_r;       // escape_radius * M                 (schwarzschild.rs:116)
    double  esc2;        // escape_radius²  (minkowski.rs:34, alcubierre.rs:70: 1e6)
    double  esc_l;       // |l| escape (morris_thorne.rs:43)
    double  R, s, inv_s, v, ship_x;
    double  denom;       // 2.0 * tanh(R / s)                 (alcubierre.rs:42)
    double  flat_rs;     // r_s above which f(r_s ± h) == 0 exactly
    double  b0;          // throat radius
    double  fd_h;        // metric/mod.rs:73-75
    double  fd_two_h;    // 2.0 * h                           (metric/mod.rs:35)
    double  fd_inv_2h;   // 1.0 / (2.0 * h)                   (curvature.rs:34)
    double  fd_sin_h, fd_cos_h;   // sin h, cos h: θ ± h rotations (Morris–Thorne)
    double  inv_denom;   // 1.0 / denom (Alcubierre)
    int32_t mk;          // 1: quotients may take the shorter sequences (Alcubierre: by 2h and denom; Schwarzschild: M in range; see make_dmetric)
    int32_t shape_exp;   // Alcubierre: 1 if a curvature stencil's tanh arguments lie within 1e-3 of its centre's
};

struct geon_dframe {
    geon_dmetric m;
    double   tet[16];      // rows e_(a)^μ                   (camera.rs:47-97)
    double   obs[4];
    double   w_im, h_im;   // tan(fov/2)·aspect, tan(fov/2)   (camera.rs:119-121)
    double   step;
    double   inv_spa;      // 1.0 / spa                       (render.rs:71)
    uint32_t W, H, spa, max_steps;
    float    captured_rgb[3];
    float    diverged_rgb[3];
    uint32_t show_stars;
    int32_t  precision;
    // mixed precision: rays whose impact parameter lies within fp64_band of
    // crit_b are integrated in FP64
    double   crit_b;
    double   fp64_band;
    double   b_scale;      // b = b_scale * sin(alpha)
    uint32_t reserved[4];
};

struct geon_dstats {
    unsigned long long captured, escaped, diverged;
    unsigned long long max_resid_bits;   // max over pixels, as the u64 image of a double ≥ 0
    unsigned long long steps_pix_sum;    // Σ per-pixel integer-average steps
    unsigned long long max_steps_hit;
    unsigned long long ray_steps;
    unsigned long long fp64_rays;
    unsigned long long gamma_steps;      // steps that evaluated Γ (plain renders)
};
#define GEON_DSTATS_WORDS (sizeof(struct geon_dstats) / sizeof(unsigned long long))

// One unit of scheduled work: a camera frame (gen 0) or a set of explicit
// initial states (gen 1). Rays of a queue are numbered globally; this job
// owns ids [ray_base, ray_base + n_rays).
struct geon_djob {
    geon_dframe F;
    unsigned long long ray_base;
    unsigned long long n_rays;
    int32_t  gen;
    int32_t  early_exit;      // Schwarzschild only: decide capture/escape at the photon sphere
    const double* y0;         // gen 1: SoA 8 × n_rays
    int32_t*  cls;            // per-ray outputs (local index)
    uint32_t* steps;
    double*   resid;
    double*   dir;            // SoA 3 × n_rays
    double*   yfin;           // optional SoA 8 × n_rays
    unsigned long long* done; // finished-ray counter
    unsigned long long* fp64; // rays of this job handed to the FP64 path
    // per-pixel products of a camera job (render.rs:96-163), any may be NULL
    float*    rgb;            // 3 × W·H interleaved
    int32_t*  cls_pix;
    double*   dir_pix;        // SoA 3 × W·H
    struct geon_dstats* stats;
};

// Launch arguments of a scheduler kernel (passed by value).
struct geon_sched_args {
    const geon_djob* jobs;
    const unsigned long long* job_base;   // jobs[k].ray_base, ascending
    uint32_t n_jobs;
    uint32_t refill;                      // steps between refill points
    unsigned long long* cursor;           // next unclaimed position
    unsigned long long end;               // claimable positions are [cursor, end)
    const unsigned long long* end_ptr;    // if set, end = *end_ptr at kernel start
    const unsigned long long* list;       // list mode: position → ray id (else position = id)
    unsigned long long list_cap;          // list mode: ring capacity (position taken modulo)
    uint32_t quantum;                     // step budget per lane (0 = unbounded)
    uint32_t n_slots;                     // lanes × rays-per-lane
    void* slot_y;                         // SoA 8 × n_slots of T (quantum mode)
    unsigned long long* slot_g;           // ray id per slot, ~0 = empty
    uint32_t* slot_n;                     // steps taken per slot
    float* slot_aux;                      // SoA 8 × n_slots: FP32 conditioning trackers, flat run
    unsigned long long* fb_list;          // FP64 fallback ray ids (FP32 kernels), a ring
    unsigned long long* fb_count;         // entries ever pushed
    unsigned long long fb_cap;            // ring capacity
    unsigned long long* ray_steps;        // Σ executed RK4 steps (for throughput)
    unsigned long long* parked;           // quantum mode: rays left in flight (added to)
};
