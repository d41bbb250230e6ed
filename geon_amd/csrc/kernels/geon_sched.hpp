// geon_sched.hpp — persistent ray scheduler for gfx950.
//
// A fixed grid of lanes drains a queue of rays. Each lane holds RPL rays in
// registers, integrates them step by step, retires a ray the moment it
// terminates (writing its outputs) and refills the freed slot from the
// queue. Refills happen at wave-uniform points every `refill` steps: the
// wave counts its empty slots with a ballot and one lane claims that many
// queue positions with a single atomicAdd, so every position is handed out
// exactly once. Ragged ray
// lengths across many frames therefore share one launch without idle lanes
// until the queue runs dry.
//
// Quantum mode (serving): a step budget bounds each launch; rays still in
// flight are written to a slot pool and resumed by the next launch, so
// requests admitted between launches join the rays already running.
// The FP64 hand-back list is a ring: entries pushed but not yet claimed
// never exceed the rays of the jobs admitted, which bounds its capacity.
#pragma once
#include "geon_fp32.hpp"

namespace geon {

GEON_D uint32_t find_job(const unsigned long long* base, uint32_t n, unsigned long long g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// initial state of a job's ray (camera.rs:113-160 or caller-supplied)
GEON_D void job_ray(const geon_djob& J, unsigned long long local, double* y) {
    if (J.gen == 1) {
#pragma unroll
        for (int c = 0; c < 8; ++c) y[c] = J.y0[c * J.n_rays + local];
        return;
    }
    const geon_dframe& F = J.F;
    const uint32_t spa = F.spa, spa2 = spa * spa;
    const unsigned long long pix = local / spa2;
    const uint32_t sub = (uint32_t)(local - pix * spa2);
    const double du = ((double)(sub % spa) + 0.5) * F.inv_spa;
    const double dv = ((double)(sub / spa) + 0.5) * F.inv_spa;
    camera_ray(F, (uint32_t)(pix % F.W), (uint32_t)(pix / F.W), du, dv, y);
}

// Impact parameter b = L/E of a ray in a static spherically symmetric chart
// (E = −g_tt p^t, L² = (g_θθ p^θ)² + (g_φφ p^φ)²/sin²θ); rays with
// |b − crit_b| < fp64_band orbit close to the photon sphere (throat) and
// are integrated in FP64. So are the rays whose FP32 trace would be handed
// to FP64 anyway: an ingoing ray below the critical impact parameter has no
// turning point, so a Schwarzschild ray from outside the photon sphere
// falls in (every capture, and every dip inside 3.2M, is re-traced in
// FP64) and a Morris–Thorne ray passes the throat (every dip inside 0.1 b0
// is); tracing them in FP32 first would only be discarded.
template <class MP>
GEON_D bool near_photon_orbit(const geon_dframe& F, const double* y) {
    if (!(F.fp64_band > 0.0)) return false;
    if constexpr (std::is_same<MP, Schwarzschild>::value || std::is_same<MP, MorrisThorne>::value) {
        double g[16];
        MP::g(F.m, y, g);
        const double s = sin(y[2]);
        const double E = -g[0] * y[4];
        const double Lt = g[10] * y[6];
        const double Lp = (s != 0.0) ? g[15] * y[7] / s : 0.0;
        const double b = sqrt(Lt * Lt + Lp * Lp) / E;
        if (fabs(b - F.crit_b) < F.fp64_band) return true;
        if (!(b < F.crit_b - F.fp64_band)) return false;
        if constexpr (std::is_same<MP, Schwarzschild>::value)
            return y[1] > 3.0 * F.m.M && y[5] < 0.0;
        else
            return y[1] * y[5] < 0.0;
    } else {
        return false;
    }
}

// Exact early decision for capture predicates (bisection): in Schwarzschild
// a null ray moving outward beyond r = 3M never turns back, one moving
// inward inside r = 3M never turns around.
GEON_D int photon_sphere_decision(const geon_dmetric& P, double r, double pr) {
    const double r3 = 3.0 * P.M;
    if (r > r3 && pr > 0.0) return ESCAPED;
    if (r < r3 && pr < 0.0) return CAPTURED;
    return RUNNING;
}

template <class MP>
GEON_D void write_result(const geon_djob& J, unsigned long long local, int cls, uint32_t n,
                         const double* y) {
    if (J.cls) J.cls[local] = cls;
    if (J.steps) J.steps[local] = n;
    if (J.resid) J.resid[local] = null_residual<MP>(J.F.m, y);
    if (J.dir) {
        double d[3] = {0.0, 0.0, 0.0};
        if (cls == ESCAPED) direction<MP>(y, d);
        J.dir[local] = d[0];
        J.dir[J.n_rays + local] = d[1];
        J.dir[2 * J.n_rays + local] = d[2];
    }
    if (J.yfin) {
#pragma unroll
        for (int c = 0; c < 8; ++c) J.yfin[c * J.n_rays + local] = y[c];
    }
}

// Adds each lane's v to the counter its p points to, one atomic per distinct
// counter among the lanes (all lanes of the wave must be here).
GEON_D void wave_add(unsigned long long* p, uint32_t v) {
    const unsigned long long mine = (unsigned long long)p;
    unsigned long long todo = __ballot(v != 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const unsigned long long lp = __shfl(mine, leader);
        const bool in = v != 0 && mine == lp;
        uint32_t sum = in ? v : 0u;
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (__lane_id() == (uint32_t)leader) atomicAdd((unsigned long long*)lp, (unsigned long long)sum);
        todo &= ~__ballot(in);
    }
}

// initial state of ray gid in precision T; returns false when the ray is
// routed to the FP64 list instead
template <class MP, class T>
GEON_D bool init_ray(const geon_djob& J, unsigned long long gid,
                                                   const geon_sched_args& A, T* y) {
    double yd[8];
    job_ray(J, gid - J.ray_base, yd);
    if constexpr (std::is_same<T, float>::value) {
        if (J.F.precision == 1 && near_photon_orbit<MP>(J.F, yd)) {
            const unsigned long long q = atomicAdd(A.fb_count, 1ull);
            A.fb_list[q % A.fb_cap] = gid;
            if (J.fp64) atomicAdd(J.fp64, 1ull);
            return false;
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) y[c] = (T)yd[c];
    return true;
}

template <class T>
GEON_D void push_fallback(const geon_sched_args& A, const geon_djob& J, unsigned long long gid) {
    const unsigned long long q = atomicAdd(A.fb_count, 1ull);
    A.fb_list[q % A.fb_cap] = gid;
    if (J.fp64) atomicAdd(J.fp64, 1ull);
}

// per-precision operations
template <class MP, class T>
struct Prec;

template <class MP>
struct Prec<MP, double> {
    using Params = geon_dmetric;
    using Run = FlatRun;
    GEON_D static Params params(const geon_dmetric& m) { return m; }
    GEON_D static int terminate(const Params& P, const double* y) { return MP::terminate(P, y); }
    GEON_D static float step(const Params& P, const double* y, double h, double* out, float&, float&,
                             FlatRun& fr) {
        rk4_step<MP>(P, y, h, out, fr);
        return 1.0f;
    }
    GEON_D static uint32_t span(const Params& P, double* y, double h, FlatRun& fr, uint32_t cap) {
        if constexpr (MP::flat_test != 0) return flat_span<MP>(P, y, h, fr, cap);
        else return 0;
    }
    GEON_D static void trig(const double*, float&, float&) {}
    // a parked ray keeps its flat run (plane 2 of the slot's auxiliary words)
    GEON_D static void park_run(const Run& fr, float* aux, size_t n, uint32_t s) {
        if constexpr (MP::flat_test == 2) aux[2 * n + s] = __uint_as_float(fr.left);
    }
    GEON_D static void resume_run(Run& fr, const float* aux, size_t n, uint32_t s) {
        if constexpr (MP::flat_test == 2) fr.left = __float_as_uint(aux[2 * n + s]);
    }
    GEON_D static float radius(const double*) { return 1e30f; }
    GEON_D static bool finite(const double* y) { return finite8(y); }
    GEON_D static bool suspicious(const Params&, int, const double*, float, float) { return false; }
    GEON_D static void track_init(float&, float&) {}
    GEON_D static void track(float&, float&, float, const double*) {}
};

// FP32 metrics that count their wall steps in the second tracker
template <class MF, class = void>
struct CountsWall : std::false_type {};
template <class MF>
struct CountsWall<MF, std::void_t<decltype(MF::counts_wall)>>
    : std::integral_constant<bool, MF::counts_wall> {};

template <class MP>
struct Prec<MP, float> {
    using MF = typename FP32Of<MP>::type;
    using Params = FParams;
    using Run = FlatRunF;
    GEON_D static Params params(const geon_dmetric& m) { return fparams(m); }
    GEON_D static int terminate(const Params& P, const float* y) { return MF::terminate(P, y); }
    GEON_D static float step(const Params& P, const float* y, float h, float* out, float& s, float& c,
                             FlatRunF& fr) {
        return rk4_step_f<MF>(P, y, h, out, s, c, fr);
    }
    // flat spans leave min_s and min_r alone: a flat step returns 1 and these
    // metrics' radius() is 1e30, so the trackers could not move
    GEON_D static uint32_t span(const Params& P, float* y, float h, FlatRunF& fr, uint32_t cap) {
        if constexpr (MF::flat_test != 0) {
            static_assert(!MF::uses_theta, "flat spans assume no θ trackers");
            return flat_span_f<MF>(P, y, h, fr, cap);
        } else {
            return 0;
        }
    }
    GEON_D static void trig(const float* y, float& s, float& c) {
        if constexpr (MF::uses_theta) sincosf_cw(y[2], &s, &c);
    }
    // a parked ray keeps its flat run, anchor included (planes 2–7), so the
    // positions it takes after resuming do not depend on where it was parked
    GEON_D static void park_run(const Run& fr, float* aux, size_t n, uint32_t s) {
        if constexpr (MF::flat_test != 0) {
            aux[2 * n + s] = __uint_as_float(fr.left);
            aux[3 * n + s] = __uint_as_float(fr.j);
#pragma unroll
            for (int i = 0; i < 4; ++i) aux[(4 + i) * n + s] = fr.a[i];
        }
    }
    GEON_D static void resume_run(Run& fr, const float* aux, size_t n, uint32_t s) {
        if constexpr (MF::flat_test != 0) {
            fr.left = __float_as_uint(aux[2 * n + s]);
            fr.j = __float_as_uint(aux[3 * n + s]);
#pragma unroll
            for (int i = 0; i < 4; ++i) fr.a[i] = aux[(4 + i) * n + s];
        }
    }
    GEON_D static float radius(const float* y) { return MF::radius(y); }
    GEON_D static bool finite(const float* y) { return finite8f(y); }
    GEON_D static bool suspicious(const Params& P, int cls, const float* y, float min_s, float min_r) {
        return MF::suspicious(P, cls, y, min_s, min_r);
    }
    // trackers: min |sin θ| and min radius along the path, or (metrics that
    // count wall steps) the number of steps that evaluated Γ
    GEON_D static void track_init(float& min_s, float& min_r) {
        min_s = 1.0f;
        min_r = CountsWall<MF>::value ? 0.0f : 1e30f;
    }
    GEON_D static void track(float& min_s, float& min_r, float asn, const float* yn) {
        if constexpr (CountsWall<MF>::value) {
            if (asn > 1.0f) min_r += 1.0f;   // exact below 2^24 steps
        } else {
            min_s = fminf(min_s, asn);
            min_r = fminf(min_r, MF::radius(yn));
        }
    }
};

template <class MP, class T, int RPL>
__device__ void sched_run(const geon_sched_args& A) {
    using PR = Prec<MP, T>;
    using Params = typename PR::Params;
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wl = __lane_id();
    const unsigned long long end = A.end_ptr ? *A.end_ptr : A.end;

    T y[RPL][8];
    bool act[RPL];
    unsigned long long g[RPL];
    uint32_t n[RPL], jb[RPL], maxs[RPL];
    T h[RPL];
    int early[RPL];
    float min_s[RPL], min_r[RPL];   // FP32 conditioning trackers
    float sn[RPL], cs[RPL];         // sin/cos of the current θ (FP32 charts with θ)
    typename PR::Run fr[RPL];       // flat-step runs
    Params P[RPL];
    unsigned long long executed = 0;
    uint32_t ex32 = 0;   // RK4 steps since the last refill point (at most refill × RPL)

    auto activate = [&](int r, unsigned long long gid, uint32_t steps_done, bool resume) {
        const uint32_t k = find_job(A.job_base, A.n_jobs, gid);
        const geon_djob& J = A.jobs[k];
        g[r] = gid;
        jb[r] = k;
        n[r] = steps_done;
        maxs[r] = J.F.max_steps;
        h[r] = (T)J.F.step;
        early[r] = J.early_exit;
        P[r] = PR::params(J.F.m);
        PR::track_init(min_s[r], min_r[r]);
        fr[r] = typename PR::Run{};
        act[r] = true;
        if (!resume) {
            act[r] = init_ray<MP, T>(J, gid, A, y[r]);
            if (act[r]) PR::trig(y[r], sn[r], cs[r]);
        }
    };

    // Finished rays are counted per job (the host polls a request's count
    // after the launch): each lane keeps a pending count for one job and the
    // wave adds the counts at its refill points and at exit, one atomic per
    // job there instead of one per ray on the same word.
    unsigned long long* done_p = nullptr;
    uint32_t done_n = 0;
    auto finish = [&](int r, int cls) {
        act[r] = false;
        const geon_djob& J = A.jobs[jb[r]];
        if (PR::suspicious(P[r], cls, y[r], min_s[r], min_r[r])) {
            push_fallback<T>(A, J, g[r]);
            return;
        }
        double yd[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) yd[c] = (double)y[r][c];
        write_result<MP>(J, g[r] - J.ray_base, cls, n[r], yd);
        if (J.done) {
            if (J.done != done_p) {
                if (done_n) atomicAdd(done_p, (unsigned long long)done_n);
                done_p = J.done;
                done_n = 0;
            }
            ++done_n;
        }
    };

    // resume rays left in flight by the previous quantum
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        act[r] = false;
        if (A.slot_g) {
            const uint32_t s = lane * RPL + r;
            const unsigned long long gid = A.slot_g[s];
            if (gid != ~0ull) {
                activate(r, gid, A.slot_n[s], true);
                const T* sy = (const T*)A.slot_y;
#pragma unroll
                for (int c = 0; c < 8; ++c) y[r][c] = sy[(size_t)c * A.n_slots + s];
                PR::trig(y[r], sn[r], cs[r]);
                if (A.slot_aux) {
                    min_s[r] = A.slot_aux[s];
                    min_r[r] = A.slot_aux[A.n_slots + s];
                    PR::resume_run(fr[r], A.slot_aux, A.n_slots, s);
                }
            }
        }
    }

    bool exhausted = false;
    uint32_t budget = A.quantum;
    for (;;) {
        // ---- refill point (all lanes of the wave converged here) ----
        wave_add(done_p, done_n);
        done_n = 0;
        executed += ex32;
        ex32 = 0;
        // A wave claims when a quarter of its slots are empty (or all are):
        // one atomicAdd per claim. A claim may run past `end`; positions at or
        // beyond it are simply not used, and between serving quanta the host
        // clamps the cursor back to `end` before admitting more work.
        if (!exhausted) {
            unsigned long long masks[RPL], orm = 0;
            uint32_t total = 0;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                masks[r] = __ballot(!act[r]);
                orm |= masks[r];
                total += __popcll(masks[r]);
            }
            const uint32_t slots = 64u * RPL;
            if (total > 0 && (4u * total >= slots || total == slots)) {
                const int leader = __ffsll((long long)orm) - 1;
                unsigned long long base = 0;
                if (wl == (uint32_t)leader) base = atomicAdd(A.cursor, (unsigned long long)total);
                base = __shfl(base, leader);
                const unsigned long long take =
                    base >= end ? 0ull : (end - base < total ? end - base : (unsigned long long)total);
                if (take < total) exhausted = true;
                uint32_t before = 0;
#pragma unroll
                for (int r = 0; r < RPL; ++r) {
                    const unsigned long long below = masks[r] & ((1ull << wl) - 1ull);
                    const uint32_t off = before + (uint32_t)__popcll(below);
                    if (!act[r] && off < take) {
                        const unsigned long long pos = base + off;
                        activate(r, A.list ? A.list[pos % A.list_cap] : pos, 0, false);
                    }
                    before += __popcll(masks[r]);
                }
            }
        }
        bool any = false;
#pragma unroll
        for (int r = 0; r < RPL; ++r) any = any || act[r];
        if (__ballot(any) == 0ull) {
            if (exhausted) break;
            continue;  // every claimed ray went to the FP64 list; claim again
        }
        if (A.quantum && budget == 0) break;
        uint32_t chunk = A.refill;
        if (A.quantum && chunk > budget) chunk = budget;

        // ---- integrate ----
        for (uint32_t it = 0; it < chunk; ++it) {
            bool live = false;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                if (!act[r]) continue;
                live = true;
                // the ray's outcome this iteration: RUNNING, or the class it
                // ends with (max steps first, then the termination test, then
                // the step's NaN guard, which keeps the pre-step state)
                int t = (n[r] >= maxs[r]) ? MAXSTEPS : PR::terminate(P[r], y[r]);
                if (t == RUNNING && early[r]) {
                    if constexpr (std::is_same<MP, Schwarzschild>::value) {
                        const geon_dmetric& Pm = A.jobs[jb[r]].F.m;
                        t = photon_sphere_decision(Pm, (double)y[r][1], (double)y[r][5]);
                    }
                }
                if (t == RUNNING) {
                    uint32_t m = 0;
                    if constexpr (MP::flat_test != 0) {
                        m = PR::span(P[r], y[r], h[r], fr[r], maxs[r] - n[r]);
                        n[r] += m;
                        executed += m;
                    }
                    if (m == 0) {
                        T yn[8];
                        float s2 = sn[r], c2 = cs[r];
                        const float asn = PR::step(P[r], y[r], h[r], yn, s2, c2, fr[r]);
                        if (PR::finite(yn)) {
#pragma unroll
                            for (int c = 0; c < 8; ++c) y[r][c] = yn[c];
                            sn[r] = s2;
                            cs[r] = c2;
                            PR::track(min_s[r], min_r[r], asn, yn);
                            ++n[r];
                            ++ex32;
                        } else {
                            t = DIVERGED;
                        }
                    }
                }
                if (t != RUNNING) finish(r, t);
            }
            if (!live) break;
        }
        if (A.quantum) budget -= chunk;
    }
    wave_add(done_p, done_n);
    executed += ex32;

    // park rays still in flight
    uint32_t in_flight = 0;
    if (A.slot_g) {
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const uint32_t s = lane * RPL + r;
            if (act[r]) {
                ++in_flight;
                A.slot_g[s] = g[r];
                A.slot_n[s] = n[r];
                T* sy = (T*)A.slot_y;
#pragma unroll
                for (int c = 0; c < 8; ++c) sy[(size_t)c * A.n_slots + s] = y[r][c];
                if (A.slot_aux) {
                    A.slot_aux[s] = min_s[r];
                    A.slot_aux[A.n_slots + s] = min_r[r];
                    PR::park_run(fr[r], A.slot_aux, A.n_slots, s);
                }
            } else {
                A.slot_g[s] = ~0ull;
            }
        }
    }
    // one atomic per wave for the executed-step and in-flight counters
    for (int off = 32; off > 0; off >>= 1) {
        executed += __shfl_xor(executed, off);
        in_flight += __shfl_xor(in_flight, off);
    }
    if (wl == 0 && A.ray_steps) atomicAdd(A.ray_steps, executed);
    if (wl == 0 && A.parked && in_flight) atomicAdd(A.parked, (unsigned long long)in_flight);
}

}  // namespace geon
