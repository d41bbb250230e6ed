// geon_batch.cpp — host side of the persistent scheduler.
//
// A batch collects camera frames and explicit ray sets, groups them into
// queues by (metric kind, precision), and runs each queue with one
// persistent launch. Metric kinds run concurrently on streams forked from
// the caller's stream. FP32-mixed queues are followed by an FP64 launch in
// list mode over the rays the FP32 kernel handed back, then every camera job
// gets its per-pixel reduction (render.rs:96-163).
#include <math.h>
#include <string.h>

#include <vector>

#include "geon_runtime.hpp"

using namespace geon_host;

namespace {

const double kPi = 3.141592653589793;

// counters block layout (u64 words)
enum { C_CURSOR32, C_CURSOR64, C_FBCURSOR, C_FBCOUNT, C_STEPS32, C_STEPS64, C_STEPSFB, C_PER_KIND };

struct HostJob {
    geon_djob dj;
    int kind = 0;
    int queue = 0;          // 0: FP64 queue, 1: FP32 queue (mixed)
    bool camera = false;
    uint64_t n_rays = 0;
    size_t stats_word = 0;  // offset of this job's geon_dstats in the counters block
    size_t done_word = 0;
    size_t fp64_word = 0;
};

}  // namespace

struct geon_batch {
    geon_ctx* ctx = nullptr;
    std::vector<HostJob> jobs;
    std::vector<void*> owned;
    unsigned long long* d_counters = nullptr;
    size_t n_counter_words = 0;
    // per (kind, queue) device tables built at run time
    geon_djob* d_jobs[4][2] = {};
    unsigned long long* d_base[4][2] = {};
    uint64_t q_rays[4][2] = {};
    unsigned long long* d_fb_list[4] = {};
    uint64_t fb_cap[4] = {};
    // host images of the tables, kept alive while their uploads are in flight
    std::vector<geon_djob> h_tab[4][2];
    std::vector<unsigned long long> h_base[4][2];
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    bool enqueued = false;   // some work of the run was enqueued
    bool ran = false;        // the whole run was enqueued (ev_end marks its end)
};

namespace geon_host {

// FP32-mixed band defaults: 1e-3 of the mass (throat radius).
static void set_precision(const geon_metric_t* m, const geon_render_cfg_t* cfg, geon_dframe* F) {
    F->precision = cfg ? cfg->precision : GEON_PREC_FP64;
    if (F->precision != GEON_PREC_MIXED) return;
    double scale = m->p[0];
    if (m->kind == GEON_METRIC_SCHWARZSCHILD) F->crit_b = 3.0 * sqrt(3.0) * m->p[0];
    else if (m->kind == GEON_METRIC_MORRIS_THORNE) F->crit_b = m->p[0];
    else scale = 0.0;
    F->fp64_band = (cfg && cfg->fp64_band > 0.0) ? cfg->fp64_band : 1e-3 * scale;
}

int frame_from(const geon_metric_t* metric, const geon_camera_t* cam, const geon_render_cfg_t* cfg,
               geon_dframe* F) {
    memset(F, 0, sizeof *F);
    GEON_TRY(make_dmetric(metric, &F->m));
    if (cam) {
        if (cam->width == 0 || cam->height == 0) return fail(GEON_E_INVALID, "camera has zero size");
        GEON_TRY(tetrad(metric, cam->observer, F->tet));
        memcpy(F->obs, cam->observer, sizeof F->obs);
        const double aspect = (double)cam->width / (double)cam->height;
        F->h_im = tan(cam->fov_y_rad * 0.5);
        F->w_im = F->h_im * aspect;
        F->W = cam->width;
        F->H = cam->height;
    }
    if (cfg) {
        const uint32_t spa = cfg->samples_per_axis ? cfg->samples_per_axis : 1;
        F->spa = spa;
        F->inv_spa = 1.0 / (double)spa;
        F->step = cfg->step;
        F->max_steps = cfg->max_steps;
        memcpy(F->captured_rgb, cfg->captured_rgb, sizeof F->captured_rgb);
        memcpy(F->diverged_rgb, cfg->diverged_rgb, sizeof F->diverged_rgb);
        F->show_stars = cfg->show_stars;
    }
    set_precision(metric, cfg, F);
    return GEON_OK;
}

static int dalloc(geon_batch* b, size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e != hipSuccess) return fail(GEON_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    b->owned.push_back(*p);
    return GEON_OK;
}

static geon_sched_cfg_t resolve_cfg(geon_ctx* c, int kind, int prec, const geon_sched_cfg_t* t) {
    geon_sched_cfg_t r = c->tuned[kind][prec];
    if (t) {
        if (t->block) r.block = t->block;
        if (t->waves_per_simd) r.waves_per_simd = t->waves_per_simd;
        if (t->rays_per_lane) r.rays_per_lane = t->rays_per_lane;
        if (t->refill) r.refill = t->refill;
    }
    if (r.block != 64 && r.block != 128 && r.block != 256) r.block = 256;
    if (r.rays_per_lane != 2) r.rays_per_lane = 1;
    if (r.waves_per_simd < 1) r.waves_per_simd = 1;
    if (r.waves_per_simd > 8) r.waves_per_simd = 8;
    if (r.refill < 1) r.refill = 1;
    if (r.refill > 65536) r.refill = 65536;   // the kernels count a chunk's steps in 32 bits
    return r;
}

// One persistent scheduler launch over `rays` queue positions.
static int sched_launch(geon_ctx* c, int kind, int f32, const geon_sched_cfg_t& cfg,
                        geon_sched_args args, uint64_t rays, hipStream_t s) {
    const uint32_t rpl = cfg.rays_per_lane;
    uint64_t lanes = (uint64_t)c->n_cu * 4ull * cfg.waves_per_simd * 64ull;
    const uint64_t need = ((rays + rpl - 1) / rpl + 63) / 64 * 64;
    if (rays > 0 && need < lanes) lanes = need;   // rays == 0: unknown, full GPU
    if (lanes < 64) lanes = 64;
    lanes = (lanes + cfg.block - 1) / cfg.block * cfg.block;
    args.refill = cfg.refill;
    args.n_slots = (uint32_t)(lanes * rpl);
    hipFunction_t fn = c->fn_sched[kind][f32][rpl - 1];
    void* kargs[] = {&args};
    hipError_t e = hipModuleLaunchKernel(fn, (uint32_t)(lanes / cfg.block), 1, 1, cfg.block, 1, 1, 0,
                                         s, kargs, nullptr);
    if (e != hipSuccess) return hip_fail(e, "scheduler launch");
    return GEON_OK;
}

int batch_run_on(geon_batch* b, const geon_sched_cfg_t* tune, hipStream_t user) {
    geon_ctx* c = b->ctx;
    if (b->jobs.empty()) return GEON_OK;
    b->enqueued = true;
    // ---- build the per-queue job tables ----
    auto& tab = b->h_tab;
    auto& base = b->h_base;
    for (auto& j : b->jobs) {
        j.dj.ray_base = b->q_rays[j.kind][j.queue];
        b->q_rays[j.kind][j.queue] += j.n_rays;
        tab[j.kind][j.queue].push_back(j.dj);
        base[j.kind][j.queue].push_back(j.dj.ray_base);
    }
    for (int k = 0; k < 4; ++k)
        for (int q = 0; q < 2; ++q) {
            if (tab[k][q].empty()) continue;
            void* p;
            GEON_TRY(dalloc(b, tab[k][q].size() * sizeof(geon_djob), &p));
            b->d_jobs[k][q] = (geon_djob*)p;
            GEON_TRY(dalloc(b, base[k][q].size() * sizeof(unsigned long long), &p));
            b->d_base[k][q] = (unsigned long long*)p;
            GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)b->d_jobs[k][q], tab[k][q].data(),
                                        tab[k][q].size() * sizeof(geon_djob), user), "H2D jobs");
            GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)b->d_base[k][q], base[k][q].data(),
                                        base[k][q].size() * 8, user), "H2D bases");
            if (q == 1) {
                b->fb_cap[k] = b->q_rays[k][1];
                GEON_TRY(dalloc(b, b->fb_cap[k] * 8, &p));
                b->d_fb_list[k] = (unsigned long long*)p;
            }
        }
    // ---- zero every counter, then fork one stream per metric kind ----
    uint64_t nw = b->n_counter_words;
    void* zargs[] = {&b->d_counters, &nw};
    GEON_HIP(hipModuleLaunchKernel(c->fn_zero, 64, 1, 1, 256, 1, 1, 0, user, zargs, nullptr), "zero");
    GEON_HIP(hipEventRecord(b->ev_start, user), "event");
    GEON_HIP(hipEventRecord(c->fork_ev, user), "event");
    for (int k = 0; k < 4; ++k) {
        const bool any = b->q_rays[k][0] || b->q_rays[k][1];
        if (!any) continue;
        hipStream_t s = c->kind_stream[k];
        GEON_HIP(hipStreamWaitEvent(s, c->fork_ev, 0), "fork");
        unsigned long long* ck = b->d_counters + (size_t)k * C_PER_KIND;
        if (b->q_rays[k][1]) {
            geon_sched_cfg_t cfg = resolve_cfg(c, k, 1, tune);
            geon_sched_args a;
            memset(&a, 0, sizeof a);
            a.jobs = b->d_jobs[k][1];
            a.job_base = b->d_base[k][1];
            a.n_jobs = (uint32_t)tab[k][1].size();
            a.cursor = ck + C_CURSOR32;
            a.end = b->q_rays[k][1];
            a.fb_list = b->d_fb_list[k];
            a.fb_count = ck + C_FBCOUNT;
            a.fb_cap = b->fb_cap[k];
            a.ray_steps = ck + C_STEPS32;
            GEON_TRY(sched_launch(c, k, 1, cfg, a, b->q_rays[k][1], s));
            // FP64 pass over the rays the FP32 pass handed back. Its length is
            // known only on the device; it runs near a tenth of the FP32 queue.
            // A short list (4096² Schwarzschild: 1.37 M rays, ~5 per slot of
            // the one-ray grid) ends on its tail and runs faster at one ray per
            // lane (286 against 297 ms for the frame); long lists (the ragged
            // batch's 3–4 M per kind) keep the tuned shape, two rays per lane
            // in flight (1.47–1.50 against 1.54 s for the batch).
            geon_sched_cfg_t cf = resolve_cfg(c, k, 0, nullptr);
            if (b->q_rays[k][1] < 24000000ull) cf.rays_per_lane = 1;
            geon_sched_args f;
            memset(&f, 0, sizeof f);
            f.jobs = b->d_jobs[k][1];
            f.job_base = b->d_base[k][1];
            f.n_jobs = (uint32_t)tab[k][1].size();
            f.cursor = ck + C_FBCURSOR;
            f.end_ptr = ck + C_FBCOUNT;
            f.list = b->d_fb_list[k];
            f.list_cap = b->fb_cap[k];
            f.ray_steps = ck + C_STEPSFB;
            // the list length is only known on the device: size for the full GPU
            GEON_TRY(sched_launch(c, k, 0, cf, f, 0, s));
        }
        if (b->q_rays[k][0]) {
            geon_sched_cfg_t cfg = resolve_cfg(c, k, 0, tune);
            geon_sched_args a;
            memset(&a, 0, sizeof a);
            a.jobs = b->d_jobs[k][0];
            a.job_base = b->d_base[k][0];
            a.n_jobs = (uint32_t)tab[k][0].size();
            a.cursor = ck + C_CURSOR64;
            a.end = b->q_rays[k][0];
            a.ray_steps = ck + C_STEPS64;
            GEON_TRY(sched_launch(c, k, 0, cfg, a, b->q_rays[k][0], s));
        }
        // per-pixel reduction of this kind's camera jobs
        for (int q = 0; q < 2; ++q)
            for (size_t i = 0; i < tab[k][q].size(); ++i) {
                const geon_djob& dj = tab[k][q][i];
                if (dj.gen != 0) continue;
                const geon_djob* dptr = b->d_jobs[k][q] + i;
                const uint64_t n_pix = (uint64_t)dj.F.W * dj.F.H;
                void* fargs[] = {&dptr};
                GEON_HIP(hipModuleLaunchKernel(c->fn_job_finalize, (uint32_t)((n_pix + 255) / 256), 1, 1,
                                               256, 1, 1, 0, s, fargs, nullptr), "finalize");
            }
        GEON_HIP(hipEventRecord(c->join_ev[k], s), "join");
        GEON_HIP(hipStreamWaitEvent(user, c->join_ev[k], 0), "join wait");
    }
    GEON_HIP(hipEventRecord(b->ev_end, user), "event");
    b->ran = true;
    return GEON_OK;
}

}  // namespace geon_host

#pragma GCC visibility push(default)
extern "C" {

int geon_batch_create(geon_ctx_t* ctx, geon_batch_t** out) {
    if (!ctx || !out) return fail(GEON_E_INVALID, "bad argument");
    hipSetDevice(ctx->device);
    geon_batch* b = new geon_batch();
    b->ctx = ctx;
    hipEventCreate(&b->ev_start);
    hipEventCreate(&b->ev_end);
    *out = b;
    return GEON_OK;
}

// Frees a batch whose enqueued work has completed.
static void batch_free(geon_batch* b) {
    for (void* p : b->owned) hipFree(p);
    if (b->d_counters) hipFree(b->d_counters);
    if (b->ev_start) hipEventDestroy(b->ev_start);
    if (b->ev_end) hipEventDestroy(b->ev_end);
    delete b;
}

// Waits for the batch's own work only: a run that was enqueued whole ends at
// ev_end; one that failed part-way has no end marker, so the device is
// drained before its buffers go.
void geon_batch_destroy(geon_batch_t* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    if (b->ran) hipEventSynchronize(b->ev_end);
    else if (b->enqueued) hipDeviceSynchronize();
    batch_free(b);
}

static int add_job(geon_batch_t* b, HostJob& j) {
    if (b->ran) return fail(GEON_E_INVALID, "batch already ran");
    j.stats_word = b->n_counter_words;  // reserve counters after the per-kind block
    b->jobs.push_back(j);
    return GEON_OK;
}

int geon_batch_add_frame(geon_batch_t* b, const geon_metric_t* metric, const geon_camera_t* cam,
                         const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                         double* dir_per_pixel, uint32_t* frame_id) {
    if (!b || !metric || !cam || !cfg) return fail(GEON_E_INVALID, "bad argument");
    if (cfg->precision != GEON_PREC_FP64 && cfg->precision != GEON_PREC_MIXED)
        return fail(GEON_E_INVALID, "unknown precision %d", cfg->precision);
    hipSetDevice(b->ctx->device);
    HostJob j;
    memset(&j.dj, 0, sizeof j.dj);
    GEON_TRY(frame_from(metric, cam, cfg, &j.dj.F));
    j.kind = metric->kind;
    j.queue = cfg->precision == GEON_PREC_MIXED ? 1 : 0;
    j.camera = true;
    j.n_rays = (uint64_t)cam->width * cam->height * j.dj.F.spa * j.dj.F.spa;
    j.dj.n_rays = j.n_rays;
    j.dj.gen = 0;
    void* p;
    GEON_TRY(dalloc(b, j.n_rays * 4, &p)); j.dj.cls = (int32_t*)p;
    GEON_TRY(dalloc(b, j.n_rays * 4, &p)); j.dj.steps = (uint32_t*)p;
    GEON_TRY(dalloc(b, j.n_rays * 8, &p)); j.dj.resid = (double*)p;
    GEON_TRY(dalloc(b, j.n_rays * 24, &p)); j.dj.dir = (double*)p;
    j.dj.rgb = rgb;
    j.dj.cls_pix = cls_per_pixel;
    j.dj.dir_pix = dir_per_pixel;
    if (frame_id) *frame_id = (uint32_t)b->jobs.size();
    return add_job(b, j);
}

int geon_batch_add_rays(geon_batch_t* b, const geon_metric_t* metric, const double* y0_soa, size_t n,
                        double step, uint32_t max_steps, int32_t precision, int32_t early_exit,
                        int32_t* cls, uint32_t* steps, double* y_final_soa, double* null_resid,
                        double* dir_soa, uint32_t* job_id) {
    if (!b || !metric || (!y0_soa && n)) return fail(GEON_E_INVALID, "bad argument");
    if (precision != GEON_PREC_FP64 && precision != GEON_PREC_MIXED)
        return fail(GEON_E_INVALID, "unknown precision %d", precision);
    if (early_exit && metric->kind != GEON_METRIC_SCHWARZSCHILD)
        return fail(GEON_E_UNSUPPORTED, "early exit is defined for Schwarzschild only");
    HostJob j;
    memset(&j.dj, 0, sizeof j.dj);
    geon_render_cfg_t cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.step = step;
    cfg.max_steps = max_steps;
    cfg.samples_per_axis = 1;
    cfg.precision = precision;
    GEON_TRY(frame_from(metric, nullptr, &cfg, &j.dj.F));
    j.kind = metric->kind;
    j.queue = precision == GEON_PREC_MIXED ? 1 : 0;
    j.n_rays = n;
    j.dj.n_rays = n;
    j.dj.gen = 1;
    j.dj.early_exit = early_exit ? 1 : 0;
    j.dj.y0 = y0_soa;
    j.dj.cls = cls;
    j.dj.steps = steps;
    j.dj.yfin = y_final_soa;
    j.dj.resid = null_resid;
    j.dj.dir = dir_soa;
    if (job_id) *job_id = (uint32_t)b->jobs.size();
    return add_job(b, j);
}

int geon_batch_run(geon_batch_t* b, const geon_sched_cfg_t* tune, void* stream) {
    if (!b) return fail(GEON_E_INVALID, "bad argument");
    if (b->ran) return fail(GEON_E_INVALID, "batch already ran");
    hipSetDevice(b->ctx->device);
    // counters: per-kind block, then per job: stats, done, fp64
    const size_t per_kind = 4 * C_PER_KIND;
    size_t w = per_kind;
    for (auto& j : b->jobs) {
        j.stats_word = w;
        j.done_word = w + GEON_DSTATS_WORDS;
        j.fp64_word = w + GEON_DSTATS_WORDS + 1;
        w += GEON_DSTATS_WORDS + 2;
    }
    b->n_counter_words = w;
    hipError_t e = hipMalloc((void**)&b->d_counters, w * 8);
    if (e != hipSuccess) return fail(GEON_E_NOMEM, "counters: %s", hipGetErrorString(e));
    for (auto& j : b->jobs) {
        j.dj.stats = (geon_dstats*)(b->d_counters + j.stats_word);
        j.dj.done = b->d_counters + j.done_word;
        j.dj.fp64 = b->d_counters + j.fp64_word;
    }
    hipStream_t s = stream ? (hipStream_t)stream : b->ctx->stream;
    return batch_run_on(b, tune, s);
}

int geon_batch_frame_stats(geon_batch_t* b, uint32_t frame_id, geon_render_stats_t* out) {
    if (!b || !out || frame_id >= b->jobs.size()) return fail(GEON_E_INVALID, "bad argument");
    if (!b->ran) return fail(GEON_E_INVALID, "batch has not run");
    const HostJob& j = b->jobs[frame_id];
    GEON_HIP(hipEventSynchronize(b->ev_end), "sync");
    unsigned long long w[GEON_DSTATS_WORDS + 2];
    GEON_HIP(hipMemcpyDtoH(w, (hipDeviceptr_t)(b->d_counters + j.stats_word), sizeof w), "D2H");
    geon_dstats ds;
    memcpy(&ds, w, sizeof ds);
    memset(out, 0, sizeof *out);
    const uint64_t n_pix = j.camera ? (uint64_t)j.dj.F.W * j.dj.F.H : j.n_rays;
    out->captured = ds.captured;
    out->escaped = ds.escaped;
    out->diverged = ds.diverged;
    double mr;
    memcpy(&mr, &ds.max_resid_bits, sizeof mr);
    out->max_null_residual = mr;
    out->avg_steps = n_pix ? (double)ds.steps_pix_sum / (double)n_pix : 0.0;
    out->max_steps_hit = ds.max_steps_hit;
    out->ray_steps = ds.ray_steps;
    out->fp64_rays = j.dj.F.precision == GEON_PREC_MIXED ? w[GEON_DSTATS_WORDS + 1] : j.n_rays;
    float ms = 0.0f;
    hipEventElapsedTime(&ms, b->ev_start, b->ev_end);
    out->kernel_ms = ms;
    return GEON_OK;
}

int geon_batch_totals(geon_batch_t* b, uint64_t* ray_steps, uint64_t* fp64_rays, double* device_ms) {
    if (!b) return fail(GEON_E_INVALID, "bad argument");
    if (!b->ran) return fail(GEON_E_INVALID, "batch has not run");
    GEON_HIP(hipEventSynchronize(b->ev_end), "sync");
    std::vector<unsigned long long> w(b->n_counter_words);
    GEON_HIP(hipMemcpyDtoH(w.data(), (hipDeviceptr_t)b->d_counters, w.size() * 8), "D2H");
    uint64_t steps = 0, fb = 0;
    for (int k = 0; k < 4; ++k) {
        const unsigned long long* ck = w.data() + (size_t)k * C_PER_KIND;
        steps += ck[C_STEPS32] + ck[C_STEPS64] + ck[C_STEPSFB];
        fb += ck[C_FBCOUNT];
    }
    for (auto& j : b->jobs)
        if (j.dj.F.precision == GEON_PREC_FP64) fb += j.n_rays;
    if (ray_steps) *ray_steps = steps;
    if (fp64_rays) *fp64_rays = fb;
    if (device_ms) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, b->ev_start, b->ev_end);
        *device_ms = ms;
    }
    return GEON_OK;
}

}  // extern "C"
#pragma GCC visibility pop

namespace geon_host {

// Batches handed to a lane by a device-pointer render: freed once their
// end event has passed (all of them, waiting, when `wait`).
void lane_reap(geon_lane* l, bool wait) {
    std::vector<geon_batch*> keep;
    for (geon_batch* b : l->retired) {
        if (wait || !b->ran || hipEventQuery(b->ev_end) == hipSuccess) geon_batch_destroy(b);
        else keep.push_back(b);
    }
    l->retired.swap(keep);
}

}  // namespace geon_host

// ---- autotuner ------------------------------------------------------------------
namespace {

// Calibration frame of each metric kind: small enough to time in milliseconds,
// ragged enough (captured, escaped and long rays) to exercise refill, and
// large enough (4M rays, ~16 per slot of the persistent grid's 256 CUs × 4
// SIMDs × 4 waves × 64 lanes) that shapes are timed at their steady-state
// throughput. On frames that fill the grid only a few times the fill and the
// tail weigh enough to favour two rays per lane, which hold every ray from
// the start; renders of millions of rays run faster at one (4096²
// Schwarzschild, FP32-mixed: 296 against 398 ms). 1000 steps per ray keep a
// candidate's timing near 0.1 s in FP64.
void calibration(int kind, geon_metric_t* m, geon_camera_t* cam, geon_render_cfg_t* cfg) {
    geon_metric_default(kind, m);
    memset(cam, 0, sizeof *cam);
    memset(cfg, 0, sizeof *cfg);
    cam->fov_y_rad = 60.0 * (kPi / 180.0);
    cam->width = cam->height = 2048;
    cfg->step = 0.05;
    cfg->max_steps = 1000;
    cfg->samples_per_axis = 1;
    m->escape_radius = 60.0;
    switch (kind) {
    case GEON_METRIC_SCHWARZSCHILD:
        m->horizon_pad = 1.005;
        cam->observer[1] = 30.0; cam->observer[2] = 1.5707963267948966;
        break;
    case GEON_METRIC_MORRIS_THORNE:
        cam->observer[1] = 15.0; cam->observer[2] = 1.5707963267948966;
        break;
    case GEON_METRIC_ALCUBIERRE:
        m->p[1] = 0.125; m->p[2] = 0.5;
        cam->observer[1] = 10.0;
        break;
    default:
        cam->observer[1] = 20.0;
    }
}

// device ms per million final ray-steps (lower is better), best of `reps`
int time_cfg(geon_ctx* c, int kind, int prec, const geon_sched_cfg_t& cfg, int reps, double* cost) {
    geon_metric_t m;
    geon_camera_t cam;
    geon_render_cfg_t rc;
    calibration(kind, &m, &cam, &rc);
    rc.precision = prec ? GEON_PREC_MIXED : GEON_PREC_FP64;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        geon_batch_t* b = nullptr;
        GEON_TRY(geon_batch_create(c, &b));
        uint32_t fid;
        int st = geon_batch_add_frame(b, &m, &cam, &rc, nullptr, nullptr, nullptr, &fid);
        if (st == GEON_OK) st = geon_batch_run(b, &cfg, nullptr);
        geon_render_stats_t s;
        if (st == GEON_OK) st = geon_batch_frame_stats(b, fid, &s);
        geon_batch_destroy(b);
        GEON_TRY(st);
        const double v = s.kernel_ms / ((double)s.ray_steps * 1e-6 + 1e-30);
        if (v < best) best = v;
    }
    *cost = best;
    return GEON_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" int geon_autotune(geon_ctx_t* ctx, int32_t kind, int32_t precision, geon_sched_cfg_t* best) {
    if (!ctx || kind < 0 || kind > 3 || (precision != GEON_PREC_FP64 && precision != GEON_PREC_MIXED))
        return fail(GEON_E_INVALID, "bad argument");
    const int prec = precision == GEON_PREC_MIXED ? 1 : 0;
    geon_sched_cfg_t cur = ctx->tuned[kind][prec];
    double cur_cost;
    GEON_TRY(time_cfg(ctx, kind, prec, cur, 1, &cur_cost));   // warm-up
    GEON_TRY(time_cfg(ctx, kind, prec, cur, 2, &cur_cost));
    const uint32_t refills[] = {8, 32, 128};
    const uint32_t rpls[] = {1, 2};
    const uint32_t waves[] = {2, 4, 8};
    const uint32_t blocks[] = {64, 128, 256};
    for (int axis = 0; axis < 4; ++axis) {
        const uint32_t* vals = axis == 0 ? refills : axis == 1 ? rpls : axis == 2 ? waves : blocks;
        const int nv = axis == 1 ? 2 : 3;
        for (int i = 0; i < nv; ++i) {
            geon_sched_cfg_t t = cur;
            if (axis == 0) t.refill = vals[i];
            if (axis == 1) t.rays_per_lane = vals[i];
            if (axis == 2) t.waves_per_simd = vals[i];
            if (axis == 3) t.block = vals[i];
            if (!memcmp(&t, &cur, sizeof t)) continue;
            double cost;
            GEON_TRY(time_cfg(ctx, kind, prec, t, 2, &cost));
            if (cost < cur_cost * 0.98) {
                cur = t;
                cur_cost = cost;
            }
        }
    }
    ctx->tuned[kind][prec] = cur;
    if (best) *best = cur;
    return GEON_OK;
}
#pragma GCC visibility pop
