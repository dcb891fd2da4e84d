// vrt_batch.hip — batched form: many camera views of uneven resolution in one
// launch.  A plan uploads the views' descriptors once (largest view first, so
// the long-running tiles start early and small views fill the tail), then
// captures the launch into a hipGraph that every vrt_batch_launch replays:
// either "the hardware-scheduled kernel, then the fold of its per-pass
// counts into the context's counters" (chunk_tiles 0), or "reset the tile
// queue; launch the persistent kernel" (the reset is a one-thread kernel, not
// a memset node).  The autotuner times every tile shape and chunk size and
// keeps the fastest for the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "vrt_internal.h"

struct vrt_batch {
    vrt_ctx* ctx = nullptr;
    vrt_batch_config cfg{};
    std::vector<vrt::ViewDesc> views;
    vrt::ViewDesc* views_dev = nullptr;
    unsigned int* queue = nullptr;
    unsigned long long* parts = nullptr;  // per-pass counts (hardware-scheduled plans)
    int64_t items = -1;                   // entries of `parts`, -1: counts by atomics
    int64_t total_tiles = 0;
    int32_t workgroups = 0;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace vrt {

int64_t batch_items(const std::vector<ViewDesc>& views, int64_t total_tiles, const vrt_batch_config& cfg) {
    if (cfg.chunk_tiles != 0) return -1;
    for (const ViewDesc& v : views)
        if (!(v.frame.render_lim < kTileSum32Lim)) return -1;
    return total_tiles * (cfg.tile_w * cfg.tile_h / 64);
}

int check_batch_config(const vrt_batch_config& c) {
    const bool shape_ok = (c.tile_w == 8 && c.tile_h == 8) || (c.tile_w == 16 && c.tile_h == 4) ||
                          (c.tile_w == 32 && c.tile_h == 2) || (c.tile_w == 16 && c.tile_h == 16) ||
                          (c.tile_w == 64 && c.tile_h == 1);
    if (!shape_ok) return fail(VRT_ERR_INVALID, "tile shape must be 8x8, 16x4, 32x2, 16x16 or 64x1");
    if (c.chunk_tiles < 0 || c.chunk_tiles > 4096) return fail(VRT_ERR_INVALID, "chunk_tiles must be in [0, 4096]");
    if (c.waves_per_cu < 0 || c.waves_per_cu > 64) return fail(VRT_ERR_INVALID, "waves_per_cu must be in [0, 64]");
    return VRT_OK;
}

namespace {

void destroy_plan(vrt_batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    if (b->capture) (void)hipStreamSynchronize(b->capture);
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->views_dev) (void)hipFree(b->views_dev);
    if (b->queue) (void)hipFree(b->queue);
    if (b->parts) (void)hipFree(b->parts);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    delete b;
}

}  // namespace

// One view's descriptor, its tiles starting at first_tile of the launch.
int make_view_desc(const vrt_view& v, const vrt_batch_config& cfg, int64_t first_tile, ViewDesc& d) {
    if (int rc = check_frame(v.frame)) return rc;
    if (int rc = check_target(v.target)) return rc;
    std::memset(&d, 0, sizeof(d));
    d.frame = make_frame_desc(v.frame, v.target);
    d.target = v.target;
    d.tiles_x = (v.target.width + cfg.tile_w - 1) / cfg.tile_w;
    const int32_t tiles_y = (v.target.height + cfg.tile_h - 1) / cfg.tile_h;
    d.tiles = d.tiles_x * tiles_y;
    d.first_tile = first_tile;
    return VRT_OK;
}

// Fills descriptors (largest view first) and the tile total.
int build_view_descs(const vrt_view* views, int32_t n, const vrt_batch_config& cfg, std::vector<ViewDesc>& out,
                     int64_t& total_tiles) {
    if (n <= 0 || !views) return fail(VRT_ERR_INVALID, "no views");
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return (int64_t)views[a].target.width * views[a].target.height >
               (int64_t)views[b].target.width * views[b].target.height;
    });
    out.clear();
    total_tiles = 0;
    for (int32_t i : order) {
        ViewDesc d;
        if (int rc = make_view_desc(views[i], cfg, total_tiles, d)) return rc;
        total_tiles += d.tiles;
        out.push_back(d);
    }
    if (total_tiles > kMaxBatchTiles) return fail(VRT_ERR_INVALID, "too many tiles in one batch");
    return VRT_OK;
}

int persistent_workgroups(vrt_ctx* c, const vrt_batch_config& cfg, int64_t total_tiles, int32_t* out) {
    if (cfg.chunk_tiles == 0) {  // hardware-scheduled: one wave per 64 pixels, four per workgroup, no queue
        const int64_t items = total_tiles * (cfg.tile_w * cfg.tile_h / 64);
        *out = (int32_t)std::min<int64_t>((items + 3) / 4, 0x7fffffff);
        return VRT_OK;
    }
    int32_t waves = cfg.waves_per_cu;
    if (waves == 0) {
        if (int rc = batch_occupancy(cfg, &waves)) return rc;
        if (waves <= 0) waves = 8;
    }
    int64_t wg = (int64_t)c->num_cus * std::max(1, waves / 4);
    const int64_t pops = (total_tiles + cfg.chunk_tiles - 1) / cfg.chunk_tiles;
    const int64_t need = (pops + 3) / 4;  // 4 waves per workgroup, one pop each at least
    wg = std::max<int64_t>(1, std::min(wg, need));
    *out = (int32_t)wg;
    return VRT_OK;
}

}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_batch_create(vrt_ctx* ctx, const vrt_view* views, int32_t n, const vrt_batch_config* cfg_in,
                     vrt_batch** out) {
    if (!ctx || !ctx->grid.vox || !out) return fail(VRT_ERR_INVALID, "invalid arguments");
    *out = nullptr;
    VRT_HIP(hipSetDevice(ctx->device));
    vrt_batch_config cfg = cfg_in ? *cfg_in : ctx->tuned_cfg;
    if (int rc = check_batch_config(cfg)) return rc;
    vrt_batch* b = new vrt_batch();
    b->ctx = ctx;
    b->cfg = cfg;
    auto bail = [&](int rc) {
        destroy_plan(b);
        return rc;
    };
    if (int rc = build_view_descs(views, n, cfg, b->views, b->total_tiles)) return bail(rc);
    if (int rc = persistent_workgroups(ctx, cfg, b->total_tiles, &b->workgroups)) return bail(rc);
    hipError_t e;
    if ((e = hipMalloc(&b->views_dev, b->views.size() * sizeof(ViewDesc))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc views"));
    if ((e = hipMalloc(&b->queue, kQueueBytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc queue"));
    // the hardware-scheduled kernel's per-pass counts, up to 2^26 entries (512 MiB)
    b->items = batch_items(b->views, b->total_tiles, cfg);
    if (b->items > ((int64_t)1 << 26)) b->items = -1;
    if (b->items > 0 && (e = hipMalloc(&b->parts, (size_t)b->items * sizeof(unsigned long long))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc partial counts"));
    if ((e = hipMemcpy(b->views_dev, b->views.data(), b->views.size() * sizeof(ViewDesc), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return bail(hip_fail(e, "hipMemcpy views"));
    if ((e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamBeginCapture"));
    int rc = cfg.chunk_tiles ? launch_queue_reset(b->queue, b->capture) : VRT_OK;
    if (!rc)
        rc = launch_render_batch(render_grid(ctx), b->views_dev, (int32_t)b->views.size(), b->total_tiles, cfg,
                                 b->workgroups, b->queue, ctx->counters, b->parts, b->capture);
    if (!rc && b->parts) rc = launch_fold_parts(b->parts, b->items, ctx->counters, b->capture);
    e = hipStreamEndCapture(b->capture, &b->graph);
    if (rc) return bail(rc);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamEndCapture"));
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
        return bail(hip_fail(e, "hipGraphInstantiate"));
    *out = b;
    return VRT_OK;
}

int vrt_batch_launch(vrt_batch* b) {
    if (!b || !b->exec) return fail(VRT_ERR_INVALID, "invalid batch");
    VRT_HIP(hipSetDevice(b->ctx->device));
    VRT_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    b->ctx->draws += b->views.size();
    return VRT_OK;
}

int vrt_batch_destroy(vrt_batch* b) {
    if (!b) return fail(VRT_ERR_INVALID, "invalid batch");
    destroy_plan(b);
    return VRT_OK;
}

int vrt_draw_views_plain(vrt_ctx* ctx, const vrt_view* views, int32_t n) {
    if (!views || n < 0) return fail(VRT_ERR_INVALID, "invalid views");
    for (int32_t i = 0; i < n; i++)
        if (int rc = vrt_draw_frame(ctx, &views[i].frame, &views[i].target)) return rc;
    return VRT_OK;
}

int vrt_autotune_batch(vrt_ctx* ctx, const vrt_view* views, int32_t n, int32_t reps, vrt_batch_config* best,
                       vrt_tune_entry* results, int32_t max_results, int32_t* n_results) {
    if (!ctx || !ctx->grid.vox) return fail(VRT_ERR_INVALID, "invalid context");
    if (reps < 1) reps = 1;
    VRT_HIP(hipSetDevice(ctx->device));
    const int32_t shapes[4][2] = {{8, 8}, {16, 4}, {32, 2}, {16, 16}};
    const int32_t chunks[6] = {0, 1, 2, 4, 8, 16};  // 0: hardware-scheduled, no queue
    std::vector<vrt_tune_entry> all;
    hipEvent_t e0, e1;
    VRT_HIP(hipEventCreate(&e0));
    VRT_HIP(hipEventCreate(&e1));
    int rc = VRT_OK;
    for (auto& sh : shapes) {
        for (int32_t ch : chunks) {
            vrt_batch_config cfg{sh[0], sh[1], ch, 0};
            vrt_batch* b = nullptr;
            if ((rc = vrt_batch_create(ctx, views, n, &cfg, &b))) break;
            std::vector<double> ms;
            rc = vrt_batch_launch(b);  // warm-up
            for (int32_t r = 0; r < reps && !rc; r++) {
                if (hipEventRecord(e0, ctx->stream) != hipSuccess) {
                    rc = fail(VRT_ERR_HIP, "event record");
                    break;
                }
                rc = vrt_batch_launch(b);
                float t = 0.0f;
                if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                    rc = fail(VRT_ERR_HIP, "autotune timing");
                ms.push_back(t);
            }
            vrt_batch_destroy(b);
            if (rc) break;
            std::sort(ms.begin(), ms.end());
            all.push_back({cfg, ms[ms.size() / 2]});
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    auto it = std::min_element(all.begin(), all.end(),
                               [](const vrt_tune_entry& a, const vrt_tune_entry& b) { return a.ms < b.ms; });
    ctx->tuned_cfg = it->config;
    ctx->tuned = true;
    if (best) *best = it->config;
    if (n_results) *n_results = (int32_t)std::min<size_t>(all.size(), (size_t)std::max(0, max_results));
    for (int32_t i = 0; results && i < max_results && i < (int32_t)all.size(); i++) results[i] = all[i];
    return VRT_OK;
}

}  // extern "C"
