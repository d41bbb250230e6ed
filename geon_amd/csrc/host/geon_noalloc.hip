// geon_noalloc.hip — geon_render_dev_noalloc: the FP64 render of
// geon_render_dev with no allocation on any call, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// geon_render_dev keeps per-ray buffers (class, steps, residual, direction:
// 40 B per ray) and a statistics block in the context, grown with hipMalloc.
// Here they live in `scratch`, a file-scope device array of a fixed size:
// at most GEON_NOALLOC_MAX_RAYS rays (pixels × samples²) per call. Each GPU
// the library runs on holds its own copy until the process ends; its device
// address is taken once per device with hipGetSymbolAddress (the array's
// name is a host-side stand-in and is never passed to a kernel). Every call
// on a GPU shares that GPU's array: the statistics block is cleared by a
// kernel at the start of each call and every ray's slots are written before
// they are read, so nothing of an earlier call leaks into a later one, but
// calls that overlap on one GPU (different streams, no ordering) race.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "geon_runtime.hpp"

using namespace geon_host;

#define GEON_NOALLOC_MAX_RAYS (1ull << 22)
#define GEON_NOALLOC_STATS_BYTES 256ull
#define GEON_NOALLOC_BYTES (GEON_NOALLOC_STATS_BYTES + 40ull * GEON_NOALLOC_MAX_RAYS)

static_assert(sizeof(geon_dstats) <= GEON_NOALLOC_STATS_BYTES, "statistics block too large");

static __device__ __align__(256) unsigned char scratch[GEON_NOALLOC_BYTES];

namespace {

const int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (taken once, with the device
// current, and kept for the process).
int scratch_address(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(GEON_E_INVALID, "device %d", device);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(GEON_E_HIP, "hipGetSymbolAddress(scratch) gave no address");
        g_ptr[device] = p;
    }
    *out = (unsigned char*)g_ptr[device];
    return GEON_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int geon_render_dev_noalloc(geon_ctx_t* ctx, const geon_metric_t* metric, const geon_camera_t* cam,
                            const geon_render_cfg_t* cfg, float* rgb, int32_t* cls_per_pixel,
                            double* dir_per_pixel, geon_render_stats_t* stats, void* stream) {
    if (!ctx || !cfg || !cam) return fail(GEON_E_INVALID, "bad argument");
    if (cfg->precision != GEON_PREC_FP64)
        return fail(GEON_E_UNSUPPORTED, "the allocation-free render is FP64 only");
    const uint64_t spa = cfg->samples_per_axis ? cfg->samples_per_axis : 1;
    const uint64_t pix = (uint64_t)cam->width * cam->height;
    if (spa > 2048 || pix > GEON_NOALLOC_MAX_RAYS || pix * spa * spa > GEON_NOALLOC_MAX_RAYS)
        return fail(GEON_E_INVALID, "%llu pixels at %llu samples per axis exceed the "
                    "allocation-free limit of %llu rays", (unsigned long long)pix,
                    (unsigned long long)spa, (unsigned long long)GEON_NOALLOC_MAX_RAYS);
    unsigned char* base;
    GEON_TRY(scratch_address(ctx->device, &base));
    GEON_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    geon_dframe F;
    uint64_t n_pix, n_rays;
    GEON_TRY(fp64_frame(metric, cam, cfg, &F, &n_pix, &n_rays));
    // statistics block, then the class (4 B), steps (4 B), residual (8 B)
    // and direction (3 × 8 B) planes, each sized for the maximum
    const uint64_t M = GEON_NOALLOC_MAX_RAYS;
    unsigned char* rays = base + GEON_NOALLOC_STATS_BYTES;
    fp64_planes P;
    P.stats = (geon_dstats*)base;
    P.cls = (int32_t*)rays;
    P.steps = (uint32_t*)(rays + 4 * M);
    P.res = (double*)(rays + 8 * M);
    P.dir = (double*)(rays + 16 * M);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return fp64_run(ctx, &F, n_pix, n_rays, P, rgb, cls_per_pixel, dir_per_pixel, stats, s,
                    nullptr, nullptr);
}

uint64_t geon_render_dev_noalloc_max_rays(void) { return GEON_NOALLOC_MAX_RAYS; }

}  // extern "C"
#pragma GCC visibility pop
