// geon_runtime.cpp — host runtime behind include/geon_hip.h.
//
// Owns a HIP context per device: the gfx950 code object (hipModuleLoad),
// the __constant__ frame-slot ring (hipModuleGetGlobal), one stream and a
// grow-only cache of device buffers. Host-side pieces of the reference path
// that run once per frame (metric constants, the Gram–Schmidt tetrad of
// camera.rs:47-97) are computed here with glibc libm, exactly as the
// reference computes them, and handed to the kernels through the slots.
#include "geon_runtime.hpp"

#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>

namespace geon_host {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(GEON_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---- metric on the host -----------------------------------------------------

bool metric_valid(const geon_metric_t* m) {
    if (!m || m->kind < GEON_METRIC_MINKOWSKI || m->kind > GEON_METRIC_MORRIS_THORNE) return false;
    if (!(m->escape_radius > 0.0)) return false;
    if ((m->kind == GEON_METRIC_ALCUBIERRE || m->kind == GEON_METRIC_MORRIS_THORNE) &&
        !(m->fd_step > 0.0))
        return false;
    if (m->kind == GEON_METRIC_ALCUBIERRE && !(m->p[1] != 0.0)) return false;
    return true;
}

// g_μν(x) of each reference metric (schwarzschild.rs:33-45, minkowski.rs:20-28,
// alcubierre.rs:33-63, morris_thorne.rs:22-36), row-major.
void host_g(const geon_metric_t* m, const double x[4], double g[16]) {
    memset(g, 0, 16 * sizeof(double));
    switch (m->kind) {
    case GEON_METRIC_MINKOWSKI:
        g[0] = -1.0; g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
        break;
    case GEON_METRIC_SCHWARZSCHILD: {
        const double r = x[1], mass = m->p[0];
        const double f = 1.0 - 2.0 * mass / r;
        const double st = sin(x[2]);
        g[0] = -f; g[5] = 1.0 / f; g[10] = r * r; g[15] = r * r * st * st;
        break;
    }
    case GEON_METRIC_ALCUBIERRE: {
        const double dx = x[1] - m->p[3], dy = x[2], dz = x[3];
        const double r_s = sqrt(dx * dx + dy * dy + dz * dz);
        const double s = m->p[1], R = m->p[0], v = m->p[2];
        const double f = (tanh((r_s + R) / s) - tanh((r_s - R) / s)) / (2.0 * tanh(R / s));
        g[0] = -1.0 + v * v * f * f; g[1] = -v * f; g[4] = -v * f;
        g[5] = 1.0; g[10] = 1.0; g[15] = 1.0;
        break;
    }
    case GEON_METRIC_MORRIS_THORNE: {
        const double l = x[1], b0 = m->p[0];
        const double r2 = b0 * b0 + l * l;
        const double st = sin(x[2]);
        g[0] = -1.0; g[5] = 1.0; g[10] = r2; g[15] = r2 * st * st;
        break;
    }
    }
}

static double dot4(const double g[16], const double a[4], const double b[4]) {
    double s = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) s += g[i * 4 + j] * a[i] * b[j];
    return s;
}

// camera.rs:47-97
int tetrad(const geon_metric_t* m, const double obs[4], double tet[16]) {
    double g[16];
    host_g(m, obs, g);
    memset(tet, 0, 16 * sizeof(double));
    const double gtt = g[0];
    if (!(gtt < 0.0))
        return fail(GEON_E_NOT_TIMELIKE, "observer is not in a timelike region (g_tt = %g >= 0)", gtt);
    tet[0] = 1.0 / sqrt(-gtt);
    for (int i = 0; i < 3; ++i) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        v[i + 1] = 1.0;
        for (int k = 0; k <= i; ++k) {
            const double* ek = tet + 4 * k;
            const double coeff = dot4(g, ek, v) / dot4(g, ek, ek);
            for (int mu = 0; mu < 4; ++mu) v[mu] -= coeff * ek[mu];
        }
        const double n2 = dot4(g, v, v);
        if (!(n2 > 0.0))
            return fail(GEON_E_GRAM_SCHMIDT, "Gram-Schmidt produced a non-spacelike vector (n2 = %g)", n2);
        const double n = sqrt(n2);
        for (int mu = 0; mu < 4; ++mu) v[mu] /= n;
        memcpy(tet + 4 * (i + 1), v, sizeof v);
    }
    return GEON_OK;
}

static bool is_pow2(double s) {
    int e;
    const double fr = frexp(fabs(s), &e);
    return fr == 0.5;
}

// Constants the kernels read from the slot, computed in the reference's order.
int make_dmetric(const geon_metric_t* m, geon_dmetric* d) {
    if (!metric_valid(m)) return fail(GEON_E_INVALID, "invalid metric description");
    memset(d, 0, sizeof *d);
    d->kind = m->kind;
    d->M = m->p[0];
    d->cap_r = m->horizon_pad * 2.0 * m->p[0];        // schwarzschild.rs:114
    d->esc_r = m->escape_radius * m->p[0];            // schwarzschild.rs:116
    d->esc2 = m->escape_radius * m->escape_radius;    // r² > 1e6 at the default
    d->esc_l = m->escape_radius;                      // |l| > 1000 at the default
    d->fd_h = m->fd_step;
    d->fd_two_h = 2.0 * m->fd_step;                   // metric/mod.rs:35
    d->fd_inv_2h = 1.0 / (2.0 * m->fd_step);          // curvature.rs:34
    if (m->kind == GEON_METRIC_ALCUBIERRE) {
        d->R = m->p[0];
        d->s = m->p[1];
        d->v = m->p[2];
        d->ship_x = m->p[3];
        d->s_pow2 = is_pow2(d->s) ? 1 : 0;
        d->inv_s = 1.0 / d->s;
        d->denom = 2.0 * tanh(d->R / d->s);           // alcubierre.rs:42
        // tanh(u) rounds to exactly 1.0 for u ≥ 19.07 (1 − tanh u < 2^-54):
        // glibc's tanh and the kernels' return 1.0 from 19.1 on. With one
        // more FD step of margin (h/s ≫ the rounding of u) every stencil
        // point of a ray beyond this radius sees f == 0.
        d->flat_rs = d->R + 19.1 * fabs(d->s) + 2.0 * m->fd_step;
        d->inv_denom = 1.0 / d->denom;
        // Quotients by 2h and by denom are formed from the correctly rounded
        // reciprocals (one multiply and two FMAs) where every operand and
        // quotient is a normal number far from overflow: f ≤ 2/denom, the
        // metric entries are bounded by 1 + v²f², and their differences are
        // 0 or at least an ulp of 1.
        d->mk = (d->s > 0.0 && d->denom >= 1e-3 && d->denom <= 2.0 && fabs(d->v) <= 1e6 &&
                 d->fd_two_h >= 1e-150 && d->fd_two_h <= 1e150) ? 1 : 0;
        // A curvature stencil point is at most two FD steps from the point
        // (|Δr_s| ≤ 2h), so its tanh arguments differ from the centre's by
        // at most 2h/|s|: the kernels expand tanh about the centre where that
        // is ≤ 1e-3 (fifth-order Taylor polynomial, remainder below 1e-18).
        d->shape_exp = (d->mk && 2.0 * fabs(m->fd_step) <= 1e-3 * fabs(d->s)) ? 1 : 0;
    }
    // Schwarzschild's quotients take the shorter sequences where M, r and
    // f = 1 − 2M/r lie within 2^±100 (the kernels test r and f)
    if (m->kind == GEON_METRIC_SCHWARZSCHILD)
        d->mk = (d->M >= 0x1p-100 && d->M <= 0x1p100) ? 1 : 0;
    if (m->kind == GEON_METRIC_MORRIS_THORNE) {
        d->b0 = m->p[0];
        d->fd_sin_h = sin(m->fd_step);
        d->fd_cos_h = cos(m->fd_step);
    }
    return GEON_OK;
}

// ---- context -----------------------------------------------------------------

static std::string default_code_object() {
    Dl_info info;
    if (dladdr((void*)&default_code_object, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t k = p.find_last_of('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/geon_kernels.hsaco";
    }
    return "geon_kernels.hsaco";
}

static const char* kMetricNames[4] = {"mink", "schw", "alc", "mt"};

int ctx_create(int device, const char* code_object, geon_ctx** out) {
    if (!out) return fail(GEON_E_INVALID, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return fail(GEON_E_INVALID, "device %d of %d", device, ndev);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    geon_ctx* c = new geon_ctx();
    c->device = device;
    const std::string path = code_object ? std::string(code_object) : default_code_object();
    if ((e = hipModuleLoad(&c->mod, path.c_str())) != hipSuccess) {
        delete c;
        return fail(GEON_E_NO_MODULE, "hipModuleLoad(%s): %s", path.c_str(), hipGetErrorString(e));
    }
    auto get = [&](hipFunction_t* f, const std::string& name) -> bool {
        hipError_t r = hipModuleGetFunction(f, c->mod, name.c_str());
        if (r != hipSuccess) {
            fail(GEON_E_NO_MODULE, "kernel %s missing: %s", name.c_str(), hipGetErrorString(r));
            return false;
        }
        return true;
    };
    bool ok = true;
    for (int k = 0; k < 4 && ok; ++k) {
        ok = ok && get(&c->fn_render[k], std::string("geon_render_rays_") + kMetricNames[k]);
        ok = ok && get(&c->fn_trace[k], std::string("geon_trace_states_") + kMetricNames[k]);
        ok = ok && get(&c->fn_trace_adaptive[k], std::string("geon_trace_adaptive_") + kMetricNames[k]);
        ok = ok && get(&c->fn_curv[k], std::string("geon_curvature_") + kMetricNames[k]);
    }
    ok = ok && get(&c->fn_finalize, "geon_render_finalize");
    ok = ok && get(&c->fn_stats_clear, "geon_stats_clear");
    ok = ok && get(&c->fn_bisect, "geon_bisect_serial");
    ok = ok && get(&c->fn_nec, "geon_nec_kk");
    ok = ok && get(&c->fn_curv_alc_ricci, "geon_curvature_alc_ricci");
    ok = ok && get(&c->fn_job_finalize, "geon_job_finalize");
    ok = ok && get(&c->fn_zero, "geon_zero_u64");
    ok = ok && get(&c->fn_fill, "geon_fill_u64");
    ok = ok && get(&c->fn_clamp, "geon_clamp_cursor");
    for (int k = 0; k < 4 && ok; ++k)
        for (int f = 0; f < 2 && ok; ++f)
            for (int r = 0; r < 2 && ok; ++r)
                ok = ok && get(&c->fn_sched[k][f][r], std::string("geon_sched_") + kMetricNames[k] +
                                                          (f ? "_f32_r" : "_f64_r") + std::to_string(r + 1));
    if (!ok) {
        hipModuleUnload(c->mod);
        delete c;
        return GEON_E_NO_MODULE;
    }
    size_t bytes = 0;
    if ((e = hipModuleGetGlobal(&c->c_frames, &bytes, c->mod, "c_frames")) != hipSuccess ||
        bytes != sizeof(geon_dframe) * GEON_CONST_SLOTS) {
        hipModuleUnload(c->mod);
        delete c;
        return fail(GEON_E_NO_MODULE, "constant slots: %s (%zu bytes)", hipGetErrorString(e), bytes);
    }
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        hipModuleUnload(c->mod);
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    for (int s = 0; s < GEON_CONST_SLOTS; ++s) {
        hipEventCreateWithFlags(&c->slot_ev[s], hipEventDisableTiming);
        c->slot_used[s] = false;
    }
    // one stream per metric kind for concurrent scheduler launches (default priority)
    for (int k = 0; k < 4; ++k) {
        hipStreamCreateWithFlags(&c->kind_stream[k], hipStreamNonBlocking);
        hipEventCreateWithFlags(&c->join_ev[k], hipEventDisableTiming);
    }
    hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        c->n_cu = ncu;
    // launch-shape defaults until geon_autotune replaces them
    for (int k = 0; k < 4; ++k) {
        c->tuned[k][0] = {256, 4, 2, 64};
        c->tuned[k][1] = {256, 4, 2, 64};
    }
    geon_lane* l0;
    const int st = ctx_lane(c, c->stream, &l0);
    if (st != GEON_OK) {
        ctx_destroy(c);
        return st;
    }
    *out = c;
    return GEON_OK;
}

// The lane's retired batches are freed first by the caller (lane_reap).
static void lane_free(geon_lane* l) {
    if (l->stream) hipStreamSynchronize(l->stream);
    for (auto& kv : l->bufs)
        if (kv.second.p) hipFree(kv.second.p);
    if (l->d_stats) hipFree(l->d_stats);
    if (l->t0) hipEventDestroy(l->t0);
    if (l->t1) hipEventDestroy(l->t1);
    delete l;
}

void ctx_destroy(geon_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& kv : c->lanes) lane_free(kv.second);
    c->lanes.clear();
    for (int s = 0; s < GEON_CONST_SLOTS; ++s)
        if (c->slot_ev[s]) hipEventDestroy(c->slot_ev[s]);
    for (int k = 0; k < 4; ++k) {
        if (c->kind_stream[k]) hipStreamDestroy(c->kind_stream[k]);
        if (c->join_ev[k]) hipEventDestroy(c->join_ev[k]);
    }
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->mod) hipModuleUnload(c->mod);
    delete c;
}

int ctx_lane(geon_ctx* c, hipStream_t stream, geon_lane** out) {
    if (!stream) stream = c->stream;
    auto it = c->lanes.find(stream);
    if (it != c->lanes.end()) {
        *out = it->second;
        return GEON_OK;
    }
    geon_lane* l = new geon_lane();
    l->stream = stream;
    hipError_t e = hipMalloc((void**)&l->d_stats, sizeof(geon_dstats));
    if (e != hipSuccess) {
        delete l;
        return fail(GEON_E_NOMEM, "stats block: %s", hipGetErrorString(e));
    }
    if ((e = hipEventCreate(&l->t0)) != hipSuccess || (e = hipEventCreate(&l->t1)) != hipSuccess) {
        l->stream = nullptr;   // nothing was enqueued on it yet
        lane_free(l);
        return hip_fail(e, "hipEventCreate");
    }
    c->lanes[stream] = l;
    *out = l;
    return GEON_OK;
}

int lane_buffer(geon_lane* l, const char* name, size_t bytes, void** p) {
    geon_lane::Buf& b = l->bufs[name];
    if (b.bytes < bytes) {
        if (b.p) {
            // the lane's earlier work may still read or write the old buffer
            hipStreamSynchronize(l->stream);
            hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
        }
        hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
        if (e != hipSuccess) return fail(GEON_E_NOMEM, "hipMalloc(%zu) for %s: %s", bytes, name,
                                         hipGetErrorString(e));
        b.bytes = bytes;
    }
    *p = b.p;
    return GEON_OK;
}

int ctx_buffer(geon_ctx* c, const char* name, size_t bytes, void** p) {
    geon_lane* l;
    GEON_TRY(ctx_lane(c, c->stream, &l));
    return lane_buffer(l, name, bytes, p);
}

// Claim the next constant slot and enqueue its upload on `stream`. A slot is
// reused only after the event recorded behind its last launch completes.
int ctx_slot(geon_ctx* c, const geon_dframe* f, hipStream_t stream, uint32_t* slot) {
    const uint32_t s = c->next_slot;
    c->next_slot = (c->next_slot + 1) % GEON_CONST_SLOTS;
    if (c->slot_used[s]) {
        hipError_t e = hipEventSynchronize(c->slot_ev[s]);
        if (e != hipSuccess) return hip_fail(e, "slot event");
    }
    hipError_t e = hipMemcpyHtoDAsync((hipDeviceptr_t)((char*)c->c_frames + s * sizeof(geon_dframe)),
                                      (void*)f, sizeof(geon_dframe), stream);
    if (e != hipSuccess) return hip_fail(e, "constant slot upload");
    *slot = s;
    return GEON_OK;
}

int ctx_slot_done(geon_ctx* c, uint32_t slot, hipStream_t stream) {
    hipError_t e = hipEventRecord(c->slot_ev[slot], stream);
    if (e != hipSuccess) return hip_fail(e, "slot event record");
    c->slot_used[slot] = true;
    return GEON_OK;
}

int launch(hipFunction_t f, uint64_t n_threads, uint32_t block, hipStream_t stream, void** args) {
    if (n_threads == 0) return GEON_OK;
    const uint64_t grid = (n_threads + block - 1) / block;
    if (grid > 0x7fffffffull) return fail(GEON_E_INVALID, "grid too large (%llu)", (unsigned long long)grid);
    hipError_t e = hipModuleLaunchKernel(f, (uint32_t)grid, 1, 1, block, 1, 1, 0, stream, args, nullptr);
    if (e != hipSuccess) return hip_fail(e, "hipModuleLaunchKernel");
    return GEON_OK;
}

}  // namespace geon_host
