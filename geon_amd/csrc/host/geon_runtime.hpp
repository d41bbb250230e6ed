// geon_runtime.hpp — internal declarations of the host runtime.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../../include/geon_hip.h"
#include "../geon_dev_types.h"

struct geon_batch;

// Scratch the context keeps for the work enqueued on one stream: device
// buffers by name (grow-only), a statistics block and two timing events.
// Work on different streams never shares a lane, so "_dev" calls on
// different streams may overlap on the device.
struct geon_lane {
    hipStream_t stream = nullptr;
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Buf> bufs;
    geon_dstats* d_stats = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<geon_batch*> retired;   // batches whose work may still run on `stream`
};

struct geon_ctx {
    int device = 0;
    hipModule_t mod = nullptr;
    hipFunction_t fn_render[4] = {};
    hipFunction_t fn_trace[4] = {};
    hipFunction_t fn_trace_adaptive[4] = {};
    hipFunction_t fn_curv[4] = {};
    hipFunction_t fn_finalize = nullptr;
    hipFunction_t fn_stats_clear = nullptr;
    hipFunction_t fn_bisect = nullptr;
    hipFunction_t fn_nec = nullptr;
    hipFunction_t fn_curv_alc_ricci = nullptr;   // Alcubierre Ricci/R/G/T by quads
    hipFunction_t fn_sched[4][2][2] = {};   // [kind][0 f64 | 1 f32][rays_per_lane - 1]
    hipFunction_t fn_job_finalize = nullptr;
    hipFunction_t fn_zero = nullptr;
    hipFunction_t fn_fill = nullptr;
    hipFunction_t fn_clamp = nullptr;
    int n_cu = 256;
    geon_sched_cfg_t tuned[4][2] = {};       // [kind][0 FP64 | 1 FP32-mixed]
    hipStream_t kind_stream[4] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[4] = {};
    hipDeviceptr_t c_frames = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t slot_ev[GEON_CONST_SLOTS] = {};
    bool slot_used[GEON_CONST_SLOTS] = {};
    uint32_t next_slot = 0;
    std::map<hipStream_t, geon_lane*> lanes;   // per stream handle; the context stream's too
};

namespace geon_host {

extern thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
bool metric_valid(const geon_metric_t* m);
void host_g(const geon_metric_t* m, const double x[4], double g[16]);
int tetrad(const geon_metric_t* m, const double obs[4], double tet[16]);
int make_dmetric(const geon_metric_t* m, geon_dmetric* d);

int ctx_create(int device, const char* code_object, geon_ctx** out);
void ctx_destroy(geon_ctx* c);
// The lane of `stream` (NULL: the context stream), created on first use.
int ctx_lane(geon_ctx* c, hipStream_t stream, geon_lane** out);
// A device buffer of at least `bytes` owned by the lane; growing it waits
// for the lane's stream before the old buffer is freed.
int lane_buffer(geon_lane* l, const char* name, size_t bytes, void** p);
// The context stream's lane buffer (host-buffer entry points).
int ctx_buffer(geon_ctx* c, const char* name, size_t bytes, void** p);
// Frees the lane's retired batches whose work has completed (all when wait).
void lane_reap(geon_lane* l, bool wait);
int ctx_slot(geon_ctx* c, const geon_dframe* f, hipStream_t stream, uint32_t* slot);
int ctx_slot_done(geon_ctx* c, uint32_t slot, hipStream_t stream);
int launch(hipFunction_t f, uint64_t n_threads, uint32_t block, hipStream_t stream, void** args);
int frame_from(const geon_metric_t* metric, const geon_camera_t* cam, const geon_render_cfg_t* cfg,
               geon_dframe* F);

// The per-ray planes of one FP64 render (n_rays entries each) and its
// statistics block: lane buffers for geon_render_dev, the fixed device array
// for geon_render_dev_noalloc.
struct fp64_planes {
    int32_t* cls = nullptr;
    uint32_t* steps = nullptr;
    double* res = nullptr;
    double* dir = nullptr;
    geon_dstats* stats = nullptr;
};
// The constant frame of an FP64 render and its pixel and ray counts.
int fp64_frame(const geon_metric_t* metric, const geon_camera_t* cam, const geon_render_cfg_t* cfg,
               geon_dframe* F, uint64_t* n_pix, uint64_t* n_rays);
// Enqueues the FP64 render on `s` (statistics clear, ray kernel, pixel
// finalize); with `stats` it waits for the stream and fills them. t0/t1,
// when given, bracket the ray kernel for stats->kernel_ms.
int fp64_run(geon_ctx* ctx, const geon_dframe* F, uint64_t n_pix, uint64_t n_rays,
             const fp64_planes& P, float* rgb, int32_t* cls_per_pixel, double* dir_per_pixel,
             geon_render_stats_t* stats, hipStream_t s, hipEvent_t t0, hipEvent_t t1);

}  // namespace geon_host

#define GEON_TRY(expr)                 \
    do {                               \
        const int _st = (expr);        \
        if (_st != GEON_OK) return _st; \
    } while (0)

#define GEON_HIP(expr, what)                                        \
    do {                                                            \
        const hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return geon_host::hip_fail(_e, what); \
    } while (0)
