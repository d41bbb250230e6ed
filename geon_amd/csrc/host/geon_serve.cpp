// geon_serve.cpp — continuous-batching server over the persistent scheduler.
//
// One host thread owns the GPU work. Each metric kind has a pipeline on its
// own stream; the pipelines run independently of each other. The loop
//   1. admits queued requests in arrival order while their buffers fit in the
//      HBM arena (first-fit sub-allocation of one device allocation);
//   2. for every idle pipeline with work, enqueues one step-bounded quantum
//      of each of its active queues (FP32, its FP64 hand-back list, FP64),
//      then an asynchronous read-back of its counters and of its jobs'
//      completion words, and an event. Rays still running at the end of a
//      quantum are parked in the queue's slot pool and resume in the next
//      quantum, so rays of newly admitted requests start next to rays
//      already in flight;
//   3. polls the events; when a pipeline's quantum ends it harvests that
//      kind's finished jobs: renders get their per-pixel reduction (enqueued
//      on the pipeline with the read-back of their statistics; the request
//      completes when that work's event has passed), bisection requests walk
//      their k-section trees and start the next round of probes.
// The loop never waits for the device: job descriptors are staged in pinned
// memory and copied on the pipeline's stream ahead of its next quantum, and
// the probe rays of bisections keep their initial states and outcomes in
// pinned host memory that the kernels read and write directly. A blocking
// copy or stream synchronisation here would wait for the other pipelines'
// persistent kernels to free a slot for the copy, and every pipeline would
// stall with it.
// The FP64 hand-back list of a kind is drained while FP32 work no longer
// fills the GPU or once it holds a quarter of the FP64 slots, so FP64
// quanta run at useful occupancy under load.
// Per-ray arithmetic never depends on the slot, lane or neighbours of a ray,
// so a request's result is the same alone and in any mix.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "geon_runtime.hpp"

using namespace geon_host;

namespace {

const double kHalfPi = 1.5707963267948966;
const int kJobWords = 2 + (int)GEON_DSTATS_WORDS;  // per job: done, fp64, geon_dstats
const int kNoRoom = 1;     // internal: the arena cannot hold the request yet
const size_t kProbePool = 64ull << 20;   // pinned bytes for bisection probe rays

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation (offsets in bytes, 256-aligned).
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset → length
    void init(char* b, size_t n) {
        base = b;
        size = n;
        free_.clear();
        free_[0] = n;
    }
    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    bool alloc(size_t n, size_t* off) {
        n = round(n ? n : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o2 = it->first + n;
                free_.erase(it);
                if (rest) free_[o2] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = round(n ? n : 1);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Queue {
    geon_djob* d_jobs = nullptr;
    unsigned long long* d_base = nullptr;
    geon_djob* h_jobs = nullptr;           // pinned copies, written once per job
    unsigned long long* h_base = nullptr;
    uint32_t n_jobs = 0;
    uint32_t n_up = 0;                     // jobs already copied to the device
    uint64_t end = 0;
    geon_sched_cfg_t cfg;
    uint32_t lanes = 0, n_slots = 0;
    void* slot_y = nullptr;
    unsigned long long* slot_g = nullptr;
    uint32_t* slot_n = nullptr;
    float* slot_aux = nullptr;
    int outstanding = 0;       // jobs not yet complete
};

struct Request;

// A metric kind's quantum pipeline: its stream, the event that ends the
// last enqueued quantum and pinned read-back buffers.
struct Pipe {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    unsigned long long* h_ctr = nullptr;  // K_WORDS counters of the kind
    unsigned long long* h_job = nullptr;  // completion words of jobs [lo, hi]
    int lo = 0, hi = -1;
    bool ran32 = false, ranfb = false;    // queues run by the quantum in flight
    unsigned long long parked32 = 0, parkedfb = 0;  // rays parked by their last quanta
    // renders whose reduction and statistics read-back are enqueued, with
    // the event that ends that work
    std::deque<std::pair<hipEvent_t, std::vector<std::shared_ptr<Request>>>> fin;
    std::vector<hipEvent_t> spare;        // events for fin
};

struct Request {
    uint64_t ticket = 0;
    int kind = GEON_REQ_RENDER;
    int state = 0;             // 0 queued, 1 running, 2 done, 3 reduction enqueued
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t footprint = 0;
    std::vector<std::pair<size_t, size_t>> ranges;   // HBM arena allocations
    std::vector<std::pair<size_t, size_t>> hranges;  // pinned pool allocations
    // render
    geon_metric_t metric;
    geon_camera_t cam;
    geon_render_cfg_t cfg;
    bool keep = false;
    float* rgb = nullptr;
    int32_t* cls_pix = nullptr;
    double* dir_pix = nullptr;
    // bisection
    std::vector<double> mass, r_obs, lo, hi, alpha, bc, probes;
    std::vector<int32_t> round_cls;
    double step = 0.01;
    uint32_t max_steps = 0, iters = 0, done_iters = 0, levels = 0;
    bool full_trace = false;
    double* d_y0 = nullptr;    // pinned host memory, read and written by the kernels
    int32_t* d_cls = nullptr;
    uint32_t* d_steps = nullptr;
    std::vector<uint32_t> round_steps;
    // current job: global index, metric kind, queue, position in the queue
    int job = -1, mk = 0, mq = 0;
    uint32_t qpos = 0;
    uint64_t n_rays = 0;
    uint64_t ray_steps = 0;
    geon_serve_result_t result;
};

}  // namespace

struct geon_server {
    geon_ctx* ctx = nullptr;
    geon_serve_cfg_t cfg;
    Arena arena;
    void* arena_mem = nullptr;
    Arena hpool;                         // pinned, device-visible: bisection probes
    void* hpool_mem = nullptr;
    unsigned long long* h_fin = nullptr; // pinned: per job, its words after the reduction
    Queue q[4][3];                       // [kind][0 FP64 | 1 FP32 | 2 FP64 probes, one per lane]
    Queue fb[4];                         // FP64 pass over FP32 hand-backs
    unsigned long long* d_list[4] = {};  // fallback ray ids
    uint64_t fb_cap = 0;
    unsigned long long* d_ctr = nullptr; // per kind: K_WORDS counters
    unsigned long long* d_job = nullptr; // per job: kJobWords
    uint32_t n_jobs_total = 0;
    hipStream_t stream = nullptr;
    Pipe pipe[4];                        // one per metric kind
    // host state
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<std::shared_ptr<Request>> pending;
    std::vector<std::shared_ptr<Request>> running;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_ticket = 1;
    uint32_t last_qpos = 0;
    bool stop = false;
    int err = GEON_OK;
    std::thread th;
    // counters
    uint64_t quanta = 0, n_done = 0, steps_done = 0;
    double busy_ms = 0;
};

namespace {

// per-kind device counters; K_PARK* count the rays each queue left parked
// in its last quantum (zeroed before every quantum)
enum { K_CUR64, K_CUR32, K_FBCUR, K_FBCOUNT, K_STEPS, K_CURB, K_PARK32, K_PARKFB, K_PARK64, K_PARKB,
       K_WORDS };

int alloc_req(geon_server* s, Request& r, size_t bytes, void** p) {
    size_t off;
    if (!s->arena.alloc(bytes, &off)) return kNoRoom;
    r.ranges.push_back({off, bytes});
    *p = s->arena.base + off;
    return GEON_OK;
}

// the same from the pinned probe pool (device-visible host memory)
int alloc_host(geon_server* s, Request& r, size_t bytes, void** p) {
    size_t off;
    if (!s->hpool.alloc(bytes, &off)) return kNoRoom;
    r.hranges.push_back({off, bytes});
    *p = s->hpool.base + off;
    return GEON_OK;
}

void free_req(geon_server* s, Request& r) {
    for (auto& rg : r.ranges) s->arena.release(rg.first, rg.second);
    r.ranges.clear();
    for (auto& rg : r.hranges) s->hpool.release(rg.first, rg.second);
    r.hranges.clear();
}

int init_queue(geon_server* s, Queue& Q, int kind, int prec, uint32_t cap_jobs, bool one_per_lane) {
    geon_ctx* c = s->ctx;
    Q.cfg = c->tuned[kind][prec];
    if (Q.cfg.rays_per_lane != 2 || one_per_lane) Q.cfg.rays_per_lane = 1;
    Q.lanes = (uint32_t)(c->n_cu * 4u * Q.cfg.waves_per_simd * 64u);
    Q.lanes = (Q.lanes + Q.cfg.block - 1) / Q.cfg.block * Q.cfg.block;
    Q.n_slots = Q.lanes * Q.cfg.rays_per_lane;
    const size_t tsz = prec ? 4 : 8;
    GEON_HIP(hipMalloc((void**)&Q.d_jobs, (size_t)cap_jobs * sizeof(geon_djob)), "jobs");
    GEON_HIP(hipMalloc((void**)&Q.d_base, (size_t)cap_jobs * 8), "bases");
    GEON_HIP(hipHostMalloc((void**)&Q.h_jobs, (size_t)cap_jobs * sizeof(geon_djob),
                           hipHostMallocDefault), "jobs (pinned)");
    GEON_HIP(hipHostMalloc((void**)&Q.h_base, (size_t)cap_jobs * 8, hipHostMallocDefault),
             "bases (pinned)");
    GEON_HIP(hipMalloc(&Q.slot_y, (size_t)Q.n_slots * 8 * tsz), "slots");
    GEON_HIP(hipMalloc((void**)&Q.slot_g, (size_t)Q.n_slots * 8), "slots");
    GEON_HIP(hipMalloc((void**)&Q.slot_n, (size_t)Q.n_slots * 4), "slots");
    GEON_HIP(hipMalloc((void**)&Q.slot_aux, (size_t)Q.n_slots * 8 * 4), "slots");
    uint64_t n = Q.n_slots;
    unsigned long long empty = ~0ull;
    void* a[] = {&Q.slot_g, &n, &empty};
    GEON_HIP(hipModuleLaunchKernel(c->fn_fill, 256, 1, 1, 256, 1, 1, 0, s->stream, a, nullptr), "fill");
    return GEON_OK;
}

void free_queue(Queue& Q) {
    if (Q.d_jobs) hipFree(Q.d_jobs);
    if (Q.d_base) hipFree(Q.d_base);
    if (Q.h_jobs) hipHostFree(Q.h_jobs);
    if (Q.h_base) hipHostFree(Q.h_base);
    if (Q.slot_y) hipFree(Q.slot_y);
    if (Q.slot_g) hipFree(Q.slot_g);
    if (Q.slot_n) hipFree(Q.slot_n);
    if (Q.slot_aux) hipFree(Q.slot_aux);
}

// Append a job to queue (kind, prec); returns the global job index.
int add_job(geon_server* s, int kind, int prec, geon_djob& dj, int* job) {
    if (s->n_jobs_total >= s->cfg.max_jobs) return fail(GEON_E_ADMISSION, "job table full");
    Queue& Q = s->q[kind][prec];
    const int j = (int)s->n_jobs_total++;
    unsigned long long* w = s->d_job + (size_t)j * kJobWords;
    dj.done = w;
    dj.fp64 = w + 1;
    dj.stats = (geon_dstats*)(w + 2);
    dj.ray_base = Q.end;
    // fresh counters (zeroed at creation, never reused); the descriptor is
    // staged here and copied on the kind's stream before its next quantum
    Q.h_jobs[Q.n_jobs] = dj;
    Q.h_base[Q.n_jobs] = dj.ray_base;
    s->last_qpos = Q.n_jobs;
    Q.n_jobs++;
    Q.end += dj.n_rays;
    Q.outstanding++;
    *job = j;
    return GEON_OK;
}

// equatorial launch (validate_shadow.rs:20-35), host libm as the reference
void equatorial(double m, double r, double a, double* y) {
    const double f = 1.0 - 2.0 * m / r;
    y[0] = 0.0; y[1] = r; y[2] = kHalfPi; y[3] = 0.0;
    y[4] = 1.0 / sqrt(f);
    y[5] = -cos(a) * sqrt(f);
    y[6] = 0.0;
    y[7] = sin(a) / r;
}

// Probe angles of the next k-section round: the 2^L − 1 midpoints the next L
// serial halvings could visit, in heap order, computed with the serial rule
// mid = 0.5 * (lo + hi).
void build_round(Request& r) {
    const uint32_t left = r.iters - r.done_iters;
    r.levels = left < 6 ? left : 6;
    const uint32_t nodes = (1u << r.levels) - 1u;
    const size_t n = r.mass.size();
    r.probes.assign(n * nodes, 0.0);
    std::vector<double> blo(nodes + 1), bhi(nodes + 1);
    for (size_t k = 0; k < n; ++k) {
        blo[1] = r.lo[k];
        bhi[1] = r.hi[k];
        for (uint32_t i = 1; i <= nodes; ++i) {
            const double mid = 0.5 * (blo[i] + bhi[i]);
            r.probes[k * nodes + (i - 1)] = mid;
            if (2 * i <= nodes) {  // not captured → hi = mid (left); captured → lo = mid (right)
                blo[2 * i] = blo[i]; bhi[2 * i] = mid;
                blo[2 * i + 1] = mid; bhi[2 * i + 1] = bhi[i];
            }
        }
    }
}

int start_round(geon_server* s, Request& r) {
    build_round(r);
    const uint32_t nodes = (1u << r.levels) - 1u;
    const size_t n = r.mass.size(), nr = n * nodes;
    std::vector<double> y0(8 * nr);
    for (size_t k = 0; k < n; ++k)
        for (uint32_t i = 0; i < nodes; ++i) {
            double y[8];
            equatorial(r.mass[k], r.r_obs[k], r.probes[k * nodes + i], y);
            const size_t idx = k * nodes + i;
            for (int c = 0; c < 8; ++c) y0[(size_t)c * nr + idx] = y[c];
        }
    memcpy(r.d_y0, y0.data(), y0.size() * 8);   // pinned: no kernel reads it until the next quantum
    // problems may differ in mass: one job per distinct mass keeps the
    // metric constants per job; here every problem of a request shares the
    // job when masses are equal, else one job per problem.
    geon_render_cfg_t cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.step = r.step;
    cfg.max_steps = r.max_steps;
    cfg.samples_per_axis = 1;
    geon_metric_t m;
    geon_metric_default(GEON_METRIC_SCHWARZSCHILD, &m);
    m.p[0] = r.mass[0];
    geon_djob dj;
    memset(&dj, 0, sizeof dj);
    GEON_TRY(frame_from(&m, nullptr, &cfg, &dj.F));
    dj.n_rays = nr;
    dj.gen = 1;
    dj.early_exit = r.full_trace ? 0 : 1;
    dj.y0 = r.d_y0;
    dj.cls = r.d_cls;
    dj.steps = r.d_steps;
    r.n_rays = nr;
    r.mk = GEON_METRIC_SCHWARZSCHILD;
    r.mq = 2;    // latency-bound chains of rounds: one ray per lane
    GEON_TRY(add_job(s, r.mk, r.mq, dj, &r.job));
    r.qpos = s->last_qpos;
    return GEON_OK;
}

int admit(geon_server* s, Request& r) {
    if (r.kind == GEON_REQ_RENDER) {
        geon_djob dj;
        memset(&dj, 0, sizeof dj);
        GEON_TRY(frame_from(&r.metric, &r.cam, &r.cfg, &dj.F));
        const uint64_t n_pix = (uint64_t)r.cam.width * r.cam.height;
        const uint64_t nr = n_pix * dj.F.spa * dj.F.spa;
        void* p;
        GEON_TRY(alloc_req(s, r, nr * 40, &p));
        dj.cls = (int32_t*)p;
        dj.steps = (uint32_t*)((char*)p + nr * 4);
        dj.resid = (double*)((char*)p + nr * 8);
        dj.dir = (double*)((char*)p + nr * 16);
        if (r.keep) {
            GEON_TRY(alloc_req(s, r, n_pix * 40, &p));
            r.rgb = (float*)p;
            r.cls_pix = (int32_t*)((char*)p + n_pix * 12);
            r.dir_pix = (double*)((char*)p + n_pix * 16);
        }
        dj.rgb = r.rgb;
        dj.cls_pix = r.cls_pix;
        dj.dir_pix = r.dir_pix;
        dj.n_rays = nr;
        dj.gen = 0;
        r.n_rays = nr;
        r.mk = r.metric.kind;
        r.mq = r.cfg.precision == GEON_PREC_MIXED ? 1 : 0;
        GEON_TRY(add_job(s, r.mk, r.mq, dj, &r.job));
        r.qpos = s->last_qpos;
        return GEON_OK;
    }
    const size_t n = r.mass.size();
    void* p;
    GEON_TRY(alloc_host(s, r, n * 63 * 8 * 8, &p));
    r.d_y0 = (double*)p;
    GEON_TRY(alloc_host(s, r, n * 63 * 4, &p));
    r.d_cls = (int32_t*)p;
    GEON_TRY(alloc_host(s, r, n * 63 * 4, &p));
    r.d_steps = (uint32_t*)p;
    r.lo.assign(n, 0.0);
    r.hi.assign(n, kHalfPi - 1e-3);
    r.done_iters = 0;
    return start_round(s, r);
}

// Bytes a request holds while it runs: of the HBM arena for a render, of the
// pinned probe pool for a bisection.
size_t footprint(const Request& r) {
    if (r.kind == GEON_REQ_RENDER) {
        const uint64_t n_pix = (uint64_t)r.cam.width * r.cam.height;
        const uint64_t spa = r.cfg.samples_per_axis ? r.cfg.samples_per_axis : 1;
        return Arena::round(n_pix * spa * spa * 40) + (r.keep ? Arena::round(n_pix * 40) : 0);
    }
    return Arena::round(r.mass.size() * 63 * 64) + 2 * Arena::round(r.mass.size() * 63 * 4);
}

int launch_quantum(geon_server* s, Queue& Q, int kind, int f32, unsigned long long* cursor,
                   unsigned long long end, const unsigned long long* end_ptr,
                   const unsigned long long* list, hipStream_t st, const geon_djob* jobs,
                   const unsigned long long* base, uint32_t n_jobs, unsigned long long* parked) {
    geon_ctx* c = s->ctx;
    geon_sched_args a;
    memset(&a, 0, sizeof a);
    a.jobs = jobs;
    a.job_base = base;
    a.n_jobs = n_jobs;
    a.refill = Q.cfg.refill;
    a.cursor = cursor;
    a.end = end;
    a.end_ptr = end_ptr;
    a.list = list;
    a.list_cap = s->fb_cap;
    a.quantum = s->cfg.quantum_steps;
    a.n_slots = Q.n_slots;
    a.slot_y = Q.slot_y;
    a.slot_g = Q.slot_g;
    a.slot_n = Q.slot_n;
    a.slot_aux = Q.slot_aux;
    a.ray_steps = s->d_ctr + (size_t)kind * K_WORDS + K_STEPS;
    a.parked = parked;
    if (f32) {
        a.fb_list = s->d_list[kind];
        a.fb_count = s->d_ctr + (size_t)kind * K_WORDS + K_FBCOUNT;
        a.fb_cap = s->fb_cap;
    }
    hipFunction_t fn = c->fn_sched[kind][f32][Q.cfg.rays_per_lane - 1];
    void* args[] = {&a};
    GEON_HIP(hipModuleLaunchKernel(fn, Q.lanes / Q.cfg.block, 1, 1, Q.cfg.block, 1, 1, 0, st, args,
                                   nullptr), "quantum");
    unsigned long long b = end;
    void* cargs[] = {&cursor, &b, (void*)&end_ptr};
    GEON_HIP(hipModuleLaunchKernel(c->fn_clamp, 1, 1, 1, 64, 1, 1, 0, st, cargs, nullptr), "clamp");
    return GEON_OK;
}

// A render whose rays are all done: enqueue its per-pixel reduction and the
// read-back of its job words (statistics) on the pipeline's stream.
int start_finish_render(geon_server* s, Request& r, hipStream_t st) {
    geon_ctx* c = s->ctx;
    Queue& Q = s->q[r.mk][r.mq];
    // job index within its queue: jobs are appended in order
    const geon_djob* dptr = Q.d_jobs + r.qpos;
    void* fargs[] = {&dptr};
    const uint64_t n_pix = (uint64_t)r.cam.width * r.cam.height;
    GEON_HIP(hipModuleLaunchKernel(c->fn_job_finalize, (uint32_t)((n_pix + 255) / 256), 1, 1, 256, 1,
                                   1, 0, st, fargs, nullptr), "finalize");
    GEON_HIP(hipMemcpyDtoHAsync(s->h_fin + (size_t)r.job * kJobWords,
                                (hipDeviceptr_t)(s->d_job + (size_t)r.job * kJobWords),
                                kJobWords * 8, st), "stats");
    return GEON_OK;
}

// The same render once the event after that work has passed.
void finish_render(geon_server* s, Request& r) {
    const uint64_t n_pix = (uint64_t)r.cam.width * r.cam.height;
    const unsigned long long* w = s->h_fin + (size_t)r.job * kJobWords;
    geon_dstats ds;
    memcpy(&ds, w + 2, sizeof ds);
    geon_render_stats_t& S = r.result.stats;
    memset(&S, 0, sizeof S);
    S.captured = ds.captured;
    S.escaped = ds.escaped;
    S.diverged = ds.diverged;
    double mr;
    memcpy(&mr, &ds.max_resid_bits, sizeof mr);
    S.max_null_residual = mr;
    S.avg_steps = (double)ds.steps_pix_sum / (double)n_pix;
    S.max_steps_hit = ds.max_steps_hit;
    S.ray_steps = ds.ray_steps;
    S.fp64_rays = r.mq ? w[1] : r.n_rays;
    r.ray_steps = ds.ray_steps;
}

// After a k-section round: walk each problem's tree with the round's classes.
int finish_round(geon_server* s, Request& r, bool* more) {
    const uint32_t nodes = (1u << r.levels) - 1u;
    const size_t n = r.mass.size();
    // pinned memory written by the quantum whose event has passed
    r.round_cls.assign(r.d_cls, r.d_cls + n * nodes);
    r.round_steps.assign(r.d_steps, r.d_steps + n * nodes);
    for (uint32_t v : r.round_steps) r.ray_steps += v;
    for (size_t k = 0; k < n; ++k) {
        uint32_t i = 1;
        for (uint32_t lv = 0; lv < r.levels; ++lv) {
            const double mid = r.probes[k * nodes + (i - 1)];
            if (r.round_cls[k * nodes + (i - 1)] == GEON_CAPTURED) {
                r.lo[k] = mid;
                i = 2 * i + 1;
            } else {
                r.hi[k] = mid;
                i = 2 * i;
            }
        }
    }
    r.done_iters += r.levels;
    if (r.done_iters < r.iters) {
        *more = true;
        return start_round(s, r);
    }
    *more = false;
    r.alpha.resize(n);
    r.bc.resize(n);
    for (size_t k = 0; k < n; ++k) {
        const double a = 0.5 * (r.lo[k] + r.hi[k]);
        const double f = 1.0 - 2.0 * r.mass[k] / r.r_obs[k];
        r.alpha[k] = a;
        r.bc[k] = r.r_obs[k] * sin(a) / sqrt(f);
    }
    return GEON_OK;
}

// Copy the descriptors of jobs added since the last upload (pinned source,
// written once) on stream st.
int upload_jobs(Queue& Q, hipStream_t st) {
    if (Q.n_up >= Q.n_jobs) return GEON_OK;
    const size_t n = Q.n_jobs - Q.n_up;
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)(Q.d_jobs + Q.n_up), Q.h_jobs + Q.n_up,
                                n * sizeof(geon_djob), st), "job upload");
    GEON_HIP(hipMemcpyHtoDAsync((hipDeviceptr_t)(Q.d_base + Q.n_up), Q.h_base + Q.n_up, n * 8, st),
             "base upload");
    Q.n_up = Q.n_jobs;
    return GEON_OK;
}

// Enqueue one quantum of every active queue of kind k, then the read-back
// of the kind's counters and of its running jobs' completion words.
int launch_kind(geon_server* s, int k) {
    geon_ctx* c = s->ctx;
    Pipe& P = s->pipe[k];
    Queue& Q64 = s->q[k][0];
    Queue& Q32 = s->q[k][1];
    Queue& FB = s->fb[k];
    unsigned long long* ck = s->d_ctr + (size_t)k * K_WORDS;
    const hipStream_t st = P.stream;
    // job descriptors added since the last quantum, ahead of the kernels
    GEON_TRY(upload_jobs(Q32, st));
    GEON_TRY(upload_jobs(Q64, st));
    GEON_TRY(upload_jobs(s->q[k][2], st));
    unsigned long long* park = ck + K_PARK32;
    uint64_t n_park = 4;
    void* za[] = {&park, &n_park};
    GEON_HIP(hipModuleLaunchKernel(c->fn_zero, 1, 1, 1, 64, 1, 1, 0, st, za, nullptr), "zero");
    if (Q32.outstanding) {
        GEON_TRY(launch_quantum(s, Q32, k, 1, ck + K_CUR32, Q32.end, nullptr, nullptr, st, Q32.d_jobs,
                                Q32.d_base, Q32.n_jobs, ck + K_PARK32));
        // hand-backs known at the last read-back (this quantum may add more)
        const unsigned long long* h = P.h_ctr;
        const unsigned long long fb_wait = (h[K_FBCOUNT] - h[K_FBCUR]) + P.parkedfb;
        const unsigned long long f32_left = (Q32.end > h[K_CUR32] ? Q32.end - h[K_CUR32] : 0) + P.parked32;
        P.ranfb = fb_wait > 0 && (f32_left < Q32.n_slots || 4 * fb_wait >= FB.n_slots);
        if (P.ranfb)
            GEON_TRY(launch_quantum(s, FB, k, 0, ck + K_FBCUR, 0, ck + K_FBCOUNT, s->d_list[k], st,
                                    Q32.d_jobs, Q32.d_base, Q32.n_jobs, ck + K_PARKFB));
    } else {
        P.ranfb = false;
    }
    P.ran32 = Q32.outstanding > 0;
    if (Q64.outstanding)
        GEON_TRY(launch_quantum(s, Q64, k, 0, ck + K_CUR64, Q64.end, nullptr, nullptr, st, Q64.d_jobs,
                                Q64.d_base, Q64.n_jobs, ck + K_PARK64));
    Queue& QB = s->q[k][2];
    if (QB.outstanding)
        GEON_TRY(launch_quantum(s, QB, k, 0, ck + K_CURB, QB.end, nullptr, nullptr, st, QB.d_jobs,
                                QB.d_base, QB.n_jobs, ck + K_PARKB));
    GEON_HIP(hipMemcpyDtoHAsync(P.h_ctr, (hipDeviceptr_t)ck, K_WORDS * 8, st), "counters");
    int lo = 1 << 30, hi = -1;
    for (auto& r : s->running) {
        if (r->mk != k || r->state == 3) continue;
        lo = r->job < lo ? r->job : lo;
        hi = r->job > hi ? r->job : hi;
    }
    P.lo = lo;
    P.hi = hi;
    if (hi >= lo)
        GEON_HIP(hipMemcpyDtoHAsync(P.h_job, (hipDeviceptr_t)(s->d_job + (size_t)lo * kJobWords),
                                    (size_t)(hi - lo + 1) * kJobWords * 8, st), "job words");
    GEON_HIP(hipEventRecord(P.done, st), "event");
    P.busy = true;
    s->quanta++;
    return GEON_OK;
}

// Mark a request done and release what it no longer holds.
void complete(geon_server* s, Request& r, int st) {
    std::lock_guard<std::mutex> lk(s->mu);
    r.result.status = st;
    r.result.ray_steps = r.ray_steps;
    r.t_done = now_ms();
    r.result.latency_ms = r.t_done - r.t_submit;
    r.result.queue_ms = r.t_admit - r.t_submit;
    r.state = 2;
    if (!r.keep || r.kind == GEON_REQ_BISECT) free_req(s, r);
    s->n_done++;
    s->steps_done += r.ray_steps;
    s->done_cv.notify_all();
}

// After kind k's quantum: finish the jobs whose rays are all done. Renders
// get their reduction enqueued and complete in reap_kind; bisections walk
// their trees now (their probe outcomes are in pinned memory).
void harvest_kind(geon_server* s, int k) {
    Pipe& P = s->pipe[k];
    std::vector<std::shared_ptr<Request>> still, fin;
    for (auto& r : s->running) {
        if (r->state == 3 || r->mk != k || r->job < P.lo || r->job > P.hi ||
            P.h_job[(size_t)(r->job - P.lo) * kJobWords] < r->n_rays) {
            still.push_back(r);
            continue;
        }
        s->q[r->mk][r->mq].outstanding--;
        if (r->kind == GEON_REQ_RENDER) {
            const int st = start_finish_render(s, *r, P.stream);
            if (st != GEON_OK) {
                complete(s, *r, st);
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                r->state = 3;
            }
            fin.push_back(r);
            still.push_back(r);
            continue;
        }
        bool more = false;
        const int st = finish_round(s, *r, &more);
        if (more && st == GEON_OK) {
            still.push_back(r);
            continue;
        }
        complete(s, *r, st);
    }
    s->running.swap(still);
    if (fin.empty()) return;
    hipEvent_t ev = nullptr;
    if (!P.spare.empty()) {
        ev = P.spare.back();
        P.spare.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
    }
    if (!ev || hipEventRecord(ev, P.stream) != hipSuccess) {
        if (ev) P.spare.push_back(ev);
        for (auto& r : fin) complete(s, *r, fail(GEON_E_HIP, "event"));
        s->running.erase(std::remove_if(s->running.begin(), s->running.end(),
                                        [](const std::shared_ptr<Request>& r) { return r->state == 2; }),
                         s->running.end());
        return;
    }
    P.fin.emplace_back(ev, std::move(fin));
}

// Complete the renders of kind k whose reduction has finished.
bool reap_kind(geon_server* s, int k) {
    Pipe& P = s->pipe[k];
    bool any = false;
    while (!P.fin.empty() && hipEventQuery(P.fin.front().first) == hipSuccess) {
        for (auto& r : P.fin.front().second) {
            finish_render(s, *r);
            complete(s, *r, GEON_OK);
        }
        P.spare.push_back(P.fin.front().first);
        P.fin.pop_front();
        any = true;
    }
    if (any)
        s->running.erase(std::remove_if(s->running.begin(), s->running.end(),
                                        [](const std::shared_ptr<Request>& r) { return r->state == 2; }),
                         s->running.end());
    return any;
}

void serve_loop(geon_server* s) {
    geon_ctx* c = s->ctx;
    hipSetDevice(c->device);
    double busy_since = -1.0;
    for (;;) {
        bool any_busy = false;
        for (int k = 0; k < 4; ++k) any_busy = any_busy || s->pipe[k].busy || !s->pipe[k].fin.empty();
        // ---- admission ----
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->running.empty() && s->pending.empty() && !any_busy) {
                if (s->stop) break;
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
            while (!s->pending.empty()) {
                auto r = s->pending.front();
                if (r->footprint > (r->kind == GEON_REQ_RENDER ? s->arena.size : s->hpool.size)) {
                    r->result.status = GEON_E_ADMISSION;
                    r->state = 2;
                    r->t_done = now_ms();
                    s->pending.pop_front();
                    s->done_cv.notify_all();
                    continue;
                }
                int st = admit(s, *r);
                if (st == kNoRoom) {
                    free_req(s, *r);    // does not fit yet: wait for memory to free up
                    break;
                }
                s->pending.pop_front();
                r->t_admit = now_ms();
                if (st != GEON_OK) {
                    free_req(s, *r);
                    r->result.status = st;
                    r->state = 2;
                    r->t_done = now_ms();
                    s->done_cv.notify_all();
                    continue;
                }
                r->state = 1;
                s->running.push_back(r);
            }
        }
        // ---- start idle pipelines that have work ----
        for (int k = 0; k < 4; ++k) {
            Pipe& P = s->pipe[k];
            if (P.busy || (!s->q[k][0].outstanding && !s->q[k][1].outstanding &&
                           !s->q[k][2].outstanding))
                continue;
            if (launch_kind(s, k) != GEON_OK) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->err = GEON_E_HIP;
            }
        }
        any_busy = false;
        for (int k = 0; k < 4; ++k) any_busy = any_busy || s->pipe[k].busy || !s->pipe[k].fin.empty();
        if (any_busy && busy_since < 0.0) busy_since = now_ms();
        // ---- harvest pipelines whose quantum ended ----
        bool progressed = false;
        for (int k = 0; k < 4; ++k) {
            Pipe& P = s->pipe[k];
            if (!P.busy || hipEventQuery(P.done) != hipSuccess) continue;
            P.busy = false;
            progressed = true;
            if (P.ran32) P.parked32 = P.h_ctr[K_PARK32];
            if (P.ranfb) P.parkedfb = P.h_ctr[K_PARKFB];
            harvest_kind(s, k);
        }
        for (int k = 0; k < 4; ++k) progressed = reap_kind(s, k) || progressed;
        any_busy = false;
        for (int k = 0; k < 4; ++k) any_busy = any_busy || s->pipe[k].busy || !s->pipe[k].fin.empty();
        if (!any_busy && busy_since >= 0.0) {
            s->busy_ms += now_ms() - busy_since;
            busy_since = -1.0;
        }
        if (any_busy && !progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int geon_serve_create(geon_ctx_t* ctx, const geon_serve_cfg_t* cfg, geon_server_t** out) {
    if (!ctx || !out) return fail(GEON_E_INVALID, "bad argument");
    hipSetDevice(ctx->device);
    auto* s = new geon_server();
    s->ctx = ctx;
    memset(&s->cfg, 0, sizeof s->cfg);
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.hbm_budget) s->cfg.hbm_budget = 8ull << 30;
    if (!s->cfg.quantum_steps) s->cfg.quantum_steps = 16384;
    if (!s->cfg.max_jobs) s->cfg.max_jobs = 16384;
    int st = GEON_OK;
    auto bail = [&](int code) {
        geon_serve_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(GEON_E_HIP, "stream"));
    for (int k = 0; k < 4; ++k) {   // default-priority streams, one per metric kind
        Pipe& P = s->pipe[k];
        if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&P.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&P.h_ctr, K_WORDS * 8, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&P.h_job, (size_t)s->cfg.max_jobs * kJobWords * 8,
                          hipHostMallocDefault) != hipSuccess)
            return bail(fail(GEON_E_HIP, "pipeline resources"));
        memset(P.h_ctr, 0, K_WORDS * 8);
    }
    if (hipMalloc(&s->arena_mem, s->cfg.hbm_budget) != hipSuccess)
        return bail(fail(GEON_E_NOMEM, "arena of %llu bytes", (unsigned long long)s->cfg.hbm_budget));
    s->arena.init((char*)s->arena_mem, s->cfg.hbm_budget);
    // probe pool: coherent pinned memory the scheduler reads and writes
    // directly (4.6 KB per bisection problem: ~230 requests of 64 problems)
    if (hipHostMalloc(&s->hpool_mem, kProbePool, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void**)&s->h_fin, (size_t)s->cfg.max_jobs * kJobWords * 8,
                      hipHostMallocDefault) != hipSuccess)
        return bail(fail(GEON_E_NOMEM, "pinned pools"));
    s->hpool.init((char*)s->hpool_mem, kProbePool);
    for (int k = 0; k < 4 && st == GEON_OK; ++k) {
        for (int q = 0; q < 2 && st == GEON_OK; ++q)
            st = init_queue(s, s->q[k][q], k, q, s->cfg.max_jobs, false);
        if (st == GEON_OK && k == GEON_METRIC_SCHWARZSCHILD)
            st = init_queue(s, s->q[k][2], k, 0, s->cfg.max_jobs, true);
        // hand-back quanta often run at low occupancy: one ray per lane
        if (st == GEON_OK) st = init_queue(s, s->fb[k], k, 0, 1, true);
    }
    if (st != GEON_OK) return bail(st);
    // FP64 hand-back rings: unclaimed entries never exceed the rays admitted,
    // and every admitted ray holds at least 40 bytes of the arena
    s->fb_cap = s->cfg.hbm_budget / 40 + 1;
    for (int k = 0; k < 4; ++k)
        if (hipMalloc((void**)&s->d_list[k], s->fb_cap * 8) != hipSuccess)
            return bail(fail(GEON_E_NOMEM, "fallback list"));
    if (hipMalloc((void**)&s->d_ctr, 4 * K_WORDS * 8) != hipSuccess ||
        hipMalloc((void**)&s->d_job, (size_t)s->cfg.max_jobs * kJobWords * 8) != hipSuccess)
        return bail(fail(GEON_E_NOMEM, "counters"));
    uint64_t n1 = 4 * K_WORDS, n2 = (uint64_t)s->cfg.max_jobs * kJobWords;
    void* a1[] = {&s->d_ctr, &n1};
    void* a2[] = {&s->d_job, &n2};
    hipModuleLaunchKernel(ctx->fn_zero, 64, 1, 1, 256, 1, 1, 0, s->stream, a1, nullptr);
    hipModuleLaunchKernel(ctx->fn_zero, 256, 1, 1, 256, 1, 1, 0, s->stream, a2, nullptr);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(GEON_E_HIP, "init"));
    s->th = std::thread(serve_loop, s);
    *out = s;
    return GEON_OK;
}

void geon_serve_destroy(geon_server_t* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    hipSetDevice(s->ctx->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    hipDeviceSynchronize();
    for (int k = 0; k < 4; ++k) {
        for (int q = 0; q < 3; ++q) free_queue(s->q[k][q]);
        free_queue(s->fb[k]);
        if (s->d_list[k]) hipFree(s->d_list[k]);
    }
    if (s->d_ctr) hipFree(s->d_ctr);
    if (s->d_job) hipFree(s->d_job);
    if (s->arena_mem) hipFree(s->arena_mem);
    if (s->hpool_mem) hipHostFree(s->hpool_mem);
    if (s->h_fin) hipHostFree(s->h_fin);
    for (int k = 0; k < 4; ++k) {
        Pipe& P = s->pipe[k];
        if (P.stream) hipStreamDestroy(P.stream);
        if (P.done) hipEventDestroy(P.done);
        for (auto& f : P.fin) P.spare.push_back(f.first);
        for (hipEvent_t e : P.spare) hipEventDestroy(e);
        if (P.h_ctr) hipHostFree(P.h_ctr);
        if (P.h_job) hipHostFree(P.h_job);
    }
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

static int submit(geon_server_t* s, std::shared_ptr<Request> r, uint64_t* ticket) {
    r->footprint = footprint(*r);
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->t_submit = now_ms();
    memset(&r->result, 0, sizeof r->result);
    r->result.kind = r->kind;
    s->pending.push_back(r);
    s->all[r->ticket] = r;
    if (ticket) *ticket = r->ticket;
    s->cv.notify_all();
    return GEON_OK;
}

int geon_serve_submit_render(geon_server_t* s, const geon_metric_t* metric, const geon_camera_t* cam,
                             const geon_render_cfg_t* cfg, int32_t keep_outputs, uint64_t* ticket) {
    if (!s || !metric || !cam || !cfg || !metric_valid(metric) || !cam->width || !cam->height)
        return fail(GEON_E_INVALID, "bad argument");
    if (cfg->precision != GEON_PREC_FP64 && cfg->precision != GEON_PREC_MIXED)
        return fail(GEON_E_INVALID, "unknown precision");
    double tet[16];
    GEON_TRY(tetrad(metric, cam->observer, tet));   // report chart errors at submission
    auto r = std::make_shared<Request>();
    r->kind = GEON_REQ_RENDER;
    r->metric = *metric;
    r->cam = *cam;
    r->cfg = *cfg;
    r->keep = keep_outputs != 0;
    return submit(s, r, ticket);
}

int geon_serve_submit_bisect(geon_server_t* s, const double* mass, const double* r_obs, size_t n,
                             double step, uint32_t max_steps, uint32_t iters, int32_t full_trace,
                             uint64_t* ticket) {
    if (!s || !mass || !r_obs || n == 0 || iters == 0) return fail(GEON_E_INVALID, "bad argument");
    for (size_t k = 0; k < n; ++k) {
        if (!(mass[k] > 0.0) || !(r_obs[k] > 2.0 * mass[k]))
            return fail(GEON_E_INVALID, "problem %zu: need mass > 0 and r_obs > 2M", k);
        if (mass[k] != mass[0])
            return fail(GEON_E_UNSUPPORTED, "one mass per bisection request (vary r_obs)");
    }
    auto r = std::make_shared<Request>();
    r->kind = GEON_REQ_BISECT;
    r->mass.assign(mass, mass + n);
    r->r_obs.assign(r_obs, r_obs + n);
    r->step = step;
    r->max_steps = max_steps;
    r->iters = iters;
    r->full_trace = full_trace != 0;
    return submit(s, r, ticket);
}

int geon_serve_poll(geon_server_t* s, uint64_t ticket, geon_serve_result_t* out) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(GEON_E_NOT_FOUND, "ticket %llu", (unsigned long long)ticket);
    if (it->second->state != 2) return GEON_E_PENDING;
    if (out) *out = it->second->result;
    return GEON_OK;
}

int geon_serve_fetch_bisect(geon_server_t* s, uint64_t ticket, double* alpha_c, double* b_c) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(GEON_E_NOT_FOUND, "ticket");
    Request& r = *it->second;
    if (r.state != 2) return GEON_E_PENDING;
    if (r.kind != GEON_REQ_BISECT || r.result.status != GEON_OK) return fail(GEON_E_INVALID, "not a finished bisection");
    if (alpha_c) memcpy(alpha_c, r.alpha.data(), r.alpha.size() * 8);
    if (b_c) memcpy(b_c, r.bc.data(), r.bc.size() * 8);
    return GEON_OK;
}

int geon_serve_fetch_render(geon_server_t* s, uint64_t ticket, float* rgb, int32_t* cls, double* dir) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) return fail(GEON_E_NOT_FOUND, "ticket");
        r = it->second;
    }
    if (r->state != 2) return GEON_E_PENDING;
    if (r->kind != GEON_REQ_RENDER || !r->keep || r->ranges.empty())
        return fail(GEON_E_INVALID, "render outputs were not kept");
    const uint64_t n = (uint64_t)r->cam.width * r->cam.height;
    if (rgb) GEON_HIP(hipMemcpyDtoH(rgb, (hipDeviceptr_t)r->rgb, n * 12), "D2H");
    if (cls) GEON_HIP(hipMemcpyDtoH(cls, (hipDeviceptr_t)r->cls_pix, n * 4), "D2H");
    if (dir) GEON_HIP(hipMemcpyDtoH(dir, (hipDeviceptr_t)r->dir_pix, n * 24), "D2H");
    return GEON_OK;
}

int geon_serve_release(geon_server_t* s, uint64_t ticket) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(GEON_E_NOT_FOUND, "ticket");
    if (it->second->state != 2) return GEON_E_PENDING;
    free_req(s, *it->second);
    s->all.erase(it);
    return GEON_OK;
}

int geon_serve_drain(geon_server_t* s) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] {
        for (auto& kv : s->all)
            if (kv.second->state != 2) return false;
        return true;
    });
    return GEON_OK;
}

int geon_serve_counters(geon_server_t* s, uint64_t* quanta, uint64_t* done, uint64_t* ray_steps,
                        double* busy_ms) {
    if (!s) return fail(GEON_E_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (quanta) *quanta = s->quanta;
    if (done) *done = s->n_done;
    if (ray_steps) *ray_steps = s->steps_done;
    if (busy_ms) *busy_ms = s->busy_ms;
    return GEON_OK;
}

}  // extern "C"
#pragma GCC visibility pop
