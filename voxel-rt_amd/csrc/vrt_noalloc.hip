// vrt_noalloc.hip — the multi-view draw without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// vrt_batch_create allocates a descriptor array, a queue and a graph per
// plan.  vrt_draw_views_noalloc keeps both the descriptors and the queue
// counters in `scratch`, one fixed array per GPU in this file, and moves the
// host-built descriptors into it through kernel arguments (four per launch),
// so the call is stream-ordered, never reads pageable host memory after it
// returns, and records only kernel nodes when a caller captures it.  Every
// call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two contexts) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vrt_internal.h"

namespace vrt {

namespace {

constexpr int32_t kNoAllocMaxViews = 256;
constexpr size_t kViewBytes = sizeof(ViewDesc) * (size_t)kNoAllocMaxViews;
constexpr size_t kQueueOffset = (kViewBytes + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kQueueOffset + kQueueBytes;
constexpr int32_t kPutPerLaunch = 4;  // 4 x 912 B of kernel arguments
constexpr int32_t kMaxDevices = 64;

// The descriptors and queue counters of vrt_draw_views_noalloc, per GPU.
// Host code never uses this name as an address: it takes the device address
// from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ViewPack {
    ViewDesc v[kPutPerLaunch];
};

__global__ void k_put_views(ViewDesc* dst, int32_t count, ViewPack pack) {
    // one wave copies the descriptors word by word
    const uint32_t* src = reinterpret_cast<const uint32_t*>(pack.v);
    uint32_t* out = reinterpret_cast<uint32_t*>(dst);
    const int32_t words = count * (int32_t)(sizeof(ViewDesc) / 4);
    for (int32_t i = threadIdx.x; i < words; i += blockDim.x) out[i] = src[i];
}

std::once_flag g_once[kMaxDevices];
unsigned char* g_ptr[kMaxDevices];
hipError_t g_err[kMaxDevices];

// The device address of `scratch` on `device` (current), looked up once per device.
int scratch_ptr(int32_t device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(VRT_ERR_INVALID, "device index out of range");
    std::call_once(g_once[device], [device]() {
        void* p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
    });
    if (g_err[device] != hipSuccess) return hip_fail(g_err[device], "hipGetSymbolAddress(scratch)");
    if (!g_ptr[device]) return fail(VRT_ERR_HIP, "no device address for the scratch array");
    *out = g_ptr[device];
    return VRT_OK;
}

}  // namespace

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_draw_views_noalloc(vrt_ctx* ctx, const vrt_view* views, int32_t n, const vrt_batch_config* cfg_in) {
    if (!ctx || !ctx->grid.vox) return fail(VRT_ERR_INVALID, "invalid context");
    if (!views || n <= 0) return fail(VRT_ERR_INVALID, "no views");
    if (n > kNoAllocMaxViews) return fail(VRT_ERR_INVALID, "vrt_draw_views_noalloc takes at most 256 views per call");
    const vrt_batch_config cfg = cfg_in ? *cfg_in : ctx->tuned_cfg;
    if (int rc = check_batch_config(cfg)) return rc;
    VRT_HIP(hipSetDevice(ctx->device));
    unsigned char* base = nullptr;
    if (int rc = scratch_ptr(ctx->device, &base)) return rc;
    ViewDesc* views_dev = reinterpret_cast<ViewDesc*>(base);
    unsigned int* queue = reinterpret_cast<unsigned int*>(base + kQueueOffset);

    for (int32_t i = 0; i < n; i++) {  // every view checked before anything is launched
        if (int rc = check_frame(views[i].frame)) return rc;
        if (int rc = check_target(views[i].target)) return rc;
    }
    // Views largest first, as vrt_batch_create orders them (a stable
    // insertion sort on the stack: this path uses no heap either).
    int32_t order[kNoAllocMaxViews];
    for (int32_t i = 0; i < n; i++) {
        const int64_t px = (int64_t)views[i].target.width * views[i].target.height;
        int32_t j = i;
        while (j > 0 && (int64_t)views[order[j - 1]].target.width * views[order[j - 1]].target.height < px) {
            order[j] = order[j - 1];
            j--;
        }
        order[j] = i;
    }
    // descriptors built four at a time and passed by value to a copy kernel
    int64_t total_tiles = 0;
    for (int32_t i = 0; i < n; i += kPutPerLaunch) {
        const int32_t k = n - i < kPutPerLaunch ? n - i : kPutPerLaunch;
        ViewPack pack{};
        for (int32_t j = 0; j < k; j++) {
            if (int rc = make_view_desc(views[order[i + j]], cfg, total_tiles, pack.v[j])) return rc;
            total_tiles += pack.v[j].tiles;
        }
        if (total_tiles > kMaxBatchTiles) return fail(VRT_ERR_INVALID, "too many tiles in one call");
        hipLaunchKernelGGL(k_put_views, dim3(1), dim3(256), 0, ctx->stream, views_dev + i, k, pack);
        VRT_HIP(hipGetLastError());
    }
    int32_t wgs = 0;
    if (int rc = persistent_workgroups(ctx, cfg, total_tiles, &wgs)) return rc;
    if (cfg.chunk_tiles)
        if (int rc = launch_queue_reset(queue, ctx->stream)) return rc;  // a kernel, not a memset
    if (int rc = launch_render_batch(render_grid(ctx), views_dev, n, total_tiles, cfg, wgs, queue, ctx->counters,
                                     nullptr, ctx->stream))
        return rc;
    ctx->draws += (uint64_t)n;
    return VRT_OK;
}

}  // extern "C"
