// vrt_server.hip — served form: the world stays resident on the device,
// removeSphere edits are applied there between frames, and render requests
// are admitted by framebuffer memory into continuous batches.
//
// Ordering.  Requests and edits form one FIFO.  A request sees every edit
// queued before it and none after.  Batches read one of two padded render
// copies of the grid: the context's own and a second one the server keeps,
// used in turn, version v on copy v mod 2.  Edit v runs on the edit stream:
// the carve and depth-field repair of the grid start at once (render kernels
// never read the grid), then the boxes of edits v-1 and v are copied from the
// grid into copy v mod 2, which waits (hipStreamWaitEvent) only for the
// batches still reading that copy, those of version v-2 or older; every later
// batch waits for that copy.  The batches of version v-1 keep running on the
// other copy meanwhile, so an edit does not drain the device.  The host never
// blocks on an edit.
//
// Continuous batching.  Each pump admits queued requests, oldest first, while
// their frame buffers fit in the budget and a batch slot (stream) is free, and
// launches them as one multi-view launch on that slot's stream.  Requests
// arriving while earlier batches run go into the next slot, so their work
// overlaps the tail of the running batches on the device.  A completed
// request keeps its buffers (and its share of the budget) until released.
//
// Memory.  The budget is reserved as one device arena at creation; request
// buffers are carved from it first-fit and returned on release, so serving
// never calls hipMalloc or hipFree (hipFree waits for the whole device,
// which would stall the batches running on other slots).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "vrt_internal.h"

namespace vrt {
int build_view_descs(const vrt_view* views, int32_t n, const vrt_batch_config& cfg, std::vector<ViewDesc>& out,
                     int64_t& total_tiles);
}  // namespace vrt

namespace {

struct Request {
    uint64_t id = 0;
    vrt_frame frame{};
    int32_t width = 0, height = 0;
    uint64_t bytes = 0;
    double t_submit = 0, t_start = 0, t_exec = 0, t_gpu_done = 0, t_done = 0;
    uint64_t world_version = 0;
    vrt_target target{};
    bool has_mem = false;
    uint64_t offset = 0;  // into the arena
    uint64_t span = 0;    // bytes carved from the arena
    bool done = false;
};

// Per-pass count entries a slot keeps: a batch whose hardware-scheduled
// launch has more passes adds its counts by atomics instead.
constexpr int64_t kSlotParts = (int64_t)1 << 20;

struct Item {  // FIFO entry: a request or an edit
    bool is_edit = false;
    uint64_t id = 0;  // request id
    int32_t cx = 0, cy = 0, cz = 0, r = 0;
};

struct EditRec {
    uint64_t version;
    int32_t cx, cy, cz, r;
};

struct Slot {
    int32_t copy = 0;  // render copy the batch reads
    hipStream_t stream = nullptr;
    hipEvent_t exec = nullptr;  // the batch reached the device (timing)
    hipEvent_t done = nullptr;  // the batch ended (timing)
    vrt::ViewDesc* views_dev = nullptr;
    vrt::ViewDesc* views_host = nullptr;  // pinned staging
    unsigned int* queue = nullptr;
    unsigned long long* parts = nullptr;  // per-pass counts of the slot's batch (kSlotParts entries)
    bool busy = false;
    std::vector<uint64_t> ids;
};

// First-fit sub-allocation of one device range; free blocks coalesce.
struct Arena {
    uint8_t* base = nullptr;
    uint64_t size = 0;
    std::map<uint64_t, uint64_t> free_blocks;  // offset -> bytes

    bool take(uint64_t bytes, uint64_t* off) {
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_blocks.erase(it);
            if (rest) free_blocks.emplace(at, rest);
            return true;
        }
        return false;
    }
    void give(uint64_t off, uint64_t bytes) {
        auto it = free_blocks.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_blocks.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_blocks.erase(next);
        }
        if (it != free_blocks.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
};

}  // namespace

struct vrt_server {
    vrt_ctx* ctx = nullptr;
    vrt_server_config cfg{};
    double t0 = 0;
    uint64_t next_id = 1;
    uint64_t version_queued = 0;   // edits submitted
    uint64_t version_applied = 0;  // edits launched on the device
    hipStream_t edit_stream = nullptr;
    hipEvent_t edit_done = nullptr;
    hipEvent_t ref = nullptr;  // device timestamp of host time t_ref
    double t_ref = 0;
    bool have_edit = false;
    std::deque<Item> fifo;
    std::map<uint64_t, Request> reqs;
    std::vector<Slot> slots;
    std::deque<uint64_t> completed;
    uint64_t bytes_in_use = 0;
    int32_t queued = 0;
    Arena arena;
    // copy[0] is the context's render copy, copy[1] the server's; each holds
    // the grid as of copy_version[k]; `recent` the edits newer than the older copy
    vrt::GridDesc copy[2];
    uint64_t copy_version[2] = {0, 0};
    int32_t* own_copy = nullptr;
    std::deque<EditRec> recent;
};

namespace vrt {
namespace {

uint64_t align256(uint64_t v) { return (v + 255) & ~(uint64_t)255; }

uint64_t frame_bytes(const vrt_server_config& c, int32_t w, int32_t h) {
    const uint64_t px = (uint64_t)w * (uint64_t)h;
    uint64_t per = 3;                    // rgb
    if (c.want_depth) per += 4;          // depth
    if (c.want_aux) per += 8 + 1 + 4 + 4 + 4;  // hit, face, eval, occ, steps
    return px * per + 6 * 256;           // each buffer 256-byte aligned
}

// Carves the request's buffers out of the arena; false when no free block is
// large enough (the caller then waits for releases, as for the budget).
bool alloc_target(vrt_server* s, Request& r) {
    const uint64_t px = (uint64_t)r.width * (uint64_t)r.height;
    uint64_t span = align256(px * 3);
    if (s->cfg.want_depth) span += align256(px * 4);
    if (s->cfg.want_aux) span += align256(px * 8) + align256(px) + 2 * align256(px * 4) + align256(px * 4);
    uint64_t off = 0;
    if (!s->arena.take(span, &off)) return false;
    uint8_t* p = s->arena.base + off;
    std::memset(&r.target, 0, sizeof(r.target));
    r.target.width = r.width;
    r.target.height = r.height;
    r.target.rgb = p;
    p += align256(px * 3);
    if (s->cfg.want_depth) {
        r.target.depth = (float*)p;
        p += align256(px * 4);
    }
    if (s->cfg.want_aux) {
        r.target.hit = (int64_t*)p;
        p += align256(px * 8);
        r.target.face = p;
        p += align256(px);
        r.target.shadow_eval = (uint32_t*)p;
        p += align256(px * 4);
        r.target.shadow_occ = (uint32_t*)p;
        p += align256(px * 4);
        r.target.steps = (uint32_t*)p;
    }
    r.has_mem = true;
    r.offset = off;
    r.span = span;
    return true;
}

double clock_of(vrt_server* s) { return now_seconds() - s->t0; }

// Device time of a completed event on the server's clock.
double device_time(vrt_server* s, hipEvent_t e) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s->ref, e) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    return s->t_ref + 1e-3 * (double)ms;
}

int retire(vrt_server* s) {
    for (auto& sl : s->slots) {
        if (!sl.busy) continue;
        hipError_t e = hipEventQuery(sl.done);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) return hip_fail(e, "batch completion");
        const double t = clock_of(s);
        const double t_exec = device_time(s, sl.exec), t_gpu = device_time(s, sl.done);
        for (uint64_t id : sl.ids) {
            auto it = s->reqs.find(id);
            if (it == s->reqs.end()) continue;
            it->second.t_exec = t_exec;
            it->second.t_gpu_done = t_gpu;
            it->second.t_done = t;
            it->second.done = true;
            s->completed.push_back(id);
        }
        sl.ids.clear();
        sl.busy = false;
    }
    return VRT_OK;
}

// Brings render copy k up to version_applied: the boxes of the edits it has
// not seen, copied from the grid on the edit stream after the batches that
// read copy k.
int catch_up(vrt_server* s, int32_t k) {
    for (auto& sl : s->slots)
        if (sl.busy && sl.copy == k) VRT_HIP(hipStreamWaitEvent(s->edit_stream, sl.done, 0));
    for (const EditRec& e : s->recent)
        if (e.version > s->copy_version[k])
            if (int rc = remove_sphere_refresh_into(s->ctx, s->copy[k], e.cx, e.cy, e.cz, e.r, s->edit_stream))
                return rc;
    s->copy_version[k] = s->version_applied;
    while (!s->recent.empty() && s->recent.front().version <= std::min(s->copy_version[0], s->copy_version[1]))
        s->recent.pop_front();
    return VRT_OK;
}

int apply_edit(vrt_server* s, const Item& it) {
    vrt_ctx* c = s->ctx;
    // carve + repair of the grid: no render reads it, so no wait
    if (int rc = remove_sphere_grid(c, it.cx, it.cy, it.cz, it.r, s->edit_stream)) return rc;
    s->version_applied++;
    s->recent.push_back({s->version_applied, it.cx, it.cy, it.cz, it.r});
    if (int rc = catch_up(s, (int32_t)(s->version_applied & 1))) return rc;
    VRT_HIP(hipEventRecord(s->edit_done, s->edit_stream));
    s->have_edit = true;
    return VRT_OK;
}

int launch_slot(vrt_server* s, Slot& sl, std::vector<uint64_t>& ids) {
    vrt_ctx* c = s->ctx;
    std::vector<vrt_view> views(ids.size());
    for (size_t i = 0; i < ids.size(); i++) {
        Request& r = s->reqs[ids[i]];
        views[i].frame = r.frame;
        views[i].target = r.target;
    }
    vrt_batch_config cfg = s->cfg.batch.tile_w ? s->cfg.batch : c->tuned_cfg;
    std::vector<ViewDesc> descs;
    int64_t total = 0;
    if (int rc = build_view_descs(views.data(), (int32_t)views.size(), cfg, descs, total)) return rc;
    int32_t wgs = 0;
    if (int rc = persistent_workgroups(c, cfg, total, &wgs)) return rc;
    // The staging buffer is reused only after the slot's previous batch
    // finished (slot not busy), so the async copy never races a rewrite.
    std::memcpy(sl.views_host, descs.data(), descs.size() * sizeof(ViewDesc));
    if (s->have_edit) VRT_HIP(hipStreamWaitEvent(sl.stream, s->edit_done, 0));
    VRT_HIP(hipEventRecord(sl.exec, sl.stream));
    VRT_HIP(hipMemcpyAsync(sl.views_dev, sl.views_host, descs.size() * sizeof(ViewDesc), hipMemcpyHostToDevice,
                           sl.stream));
    if (cfg.chunk_tiles)
        if (int rc = launch_queue_reset(sl.queue, sl.stream)) return rc;
    sl.copy = (int32_t)(s->version_applied & 1);
    GridDesc g = render_grid(c);
    g.vox = s->copy[sl.copy].vox;
    const int64_t items = batch_items(descs, total, cfg);
    unsigned long long* parts = items > 0 && items <= kSlotParts ? sl.parts : nullptr;
    if (int rc = launch_render_batch(g, sl.views_dev, (int32_t)descs.size(), total, cfg, wgs, sl.queue, c->counters,
                                     parts, sl.stream))
        return rc;
    if (parts)
        if (int rc = launch_fold_parts(parts, items, c->counters, sl.stream)) return rc;
    VRT_HIP(hipEventRecord(sl.done, sl.stream));
    const double t = clock_of(s);
    for (uint64_t id : ids) {
        Request& r = s->reqs[id];
        r.t_start = t;
        r.world_version = s->version_applied;
    }
    sl.ids = ids;
    sl.busy = true;
    c->draws += ids.size();
    return VRT_OK;
}

// Admit and launch; returns the number of batches launched.
int admit(vrt_server* s, int* launched) {
    *launched = 0;
    for (;;) {
        while (!s->fifo.empty() && s->fifo.front().is_edit) {
            if (int rc = apply_edit(s, s->fifo.front())) return rc;
            s->fifo.pop_front();
        }
        if (s->fifo.empty()) return VRT_OK;
        Slot* free_slot = nullptr;
        for (auto& sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (!free_slot) return VRT_OK;
        std::vector<uint64_t> ids;
        int64_t pixels = 0;
        while (!s->fifo.empty() && !s->fifo.front().is_edit && (int32_t)ids.size() < s->cfg.max_batch_views) {
            Request& r = s->reqs[s->fifo.front().id];
            const int64_t px = (int64_t)r.width * r.height;
            if (!ids.empty() && s->cfg.max_batch_pixels > 0 && pixels + px > s->cfg.max_batch_pixels) break;
            if (s->bytes_in_use + r.bytes > s->cfg.framebuffer_budget) break;
            if (!alloc_target(s, r)) break;  // no block large enough: wait for releases
            s->bytes_in_use += r.bytes;
            pixels += px;
            ids.push_back(r.id);
            s->fifo.pop_front();
            s->queued--;
        }
        if (ids.empty()) return VRT_OK;  // budget full: wait for releases
        if (int rc = launch_slot(s, *free_slot, ids)) return rc;
        (*launched)++;
    }
}

void destroy_server(vrt_server* s) {
    if (!s) return;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (auto& sl : s->slots)
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (auto& sl : s->slots) sl.busy = false;
    // the context's render copy takes the edits only the server's copy has seen
    if (s->edit_stream && s->own_copy && s->copy_version[0] != s->version_applied) (void)catch_up(s, 0);
    if (s->edit_stream) (void)hipStreamSynchronize(s->edit_stream);
    for (auto& sl : s->slots) {
        if (sl.exec) (void)hipEventDestroy(sl.exec);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.views_dev) (void)hipFree(sl.views_dev);
        if (sl.views_host) (void)hipHostFree(sl.views_host);
        if (sl.queue) (void)hipFree(sl.queue);
        if (sl.parts) (void)hipFree(sl.parts);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->arena.base) (void)hipFree(s->arena.base);
    if (s->own_copy) (void)hipFree(s->own_copy);
    if (s->edit_done) (void)hipEventDestroy(s->edit_done);
    if (s->ref) (void)hipEventDestroy(s->ref);
    if (s->edit_stream) (void)hipStreamDestroy(s->edit_stream);
    delete s;
}

}  // namespace
}  // namespace vrt

using namespace vrt;

extern "C" {

int vrt_server_create(vrt_ctx* ctx, const vrt_server_config* cfg, vrt_server** out) {
    if (!ctx || !ctx->grid.vox || !cfg || !out) return fail(VRT_ERR_INVALID, "invalid arguments");
    *out = nullptr;
    if (cfg->max_batch_views < 1 || cfg->max_batch_views > 4096 || cfg->streams < 1 || cfg->streams > 16 ||
        cfg->framebuffer_budget == 0 || cfg->max_batch_pixels < 0)
        return fail(VRT_ERR_INVALID, "server config out of range");
    // tile_w == 0 takes the context's default (already checked when tuned)
    if (cfg->batch.tile_w != 0)
        if (int rc = check_batch_config(cfg->batch)) return rc;
    VRT_HIP(hipSetDevice(ctx->device));
    vrt_server* s = new vrt_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->t0 = now_seconds();
    auto bail = [&](int rc) {
        destroy_server(s);
        return rc;
    };
    hipError_t e;
    s->arena.size = align256(cfg->framebuffer_budget);
    if ((e = hipMalloc(&s->arena.base, s->arena.size)) != hipSuccess) {
        s->arena.base = nullptr;
        return bail(hip_fail(e, "hipMalloc of the framebuffer budget"));
    }
    s->arena.free_blocks.emplace(0, s->arena.size);
    if ((e = hipStreamCreateWithFlags(&s->edit_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&s->edit_done, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreate(&s->ref)) != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
    // edits and batches must also follow whatever the context queued before
    if ((e = hipEventRecord(s->edit_done, ctx->stream)) != hipSuccess) return bail(hip_fail(e, "hipEventRecord"));
    if ((e = hipStreamWaitEvent(s->edit_stream, s->edit_done, 0)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamWaitEvent"));
    s->have_edit = true;
    // the second render copy, equal to the context's
    s->copy[0] = ctx->rgrid;
    s->copy[1] = ctx->rgrid;
    if ((e = hipMalloc(&s->own_copy, ctx->rgrid.count * sizeof(int32_t))) != hipSuccess) {
        s->own_copy = nullptr;
        return bail(hip_fail(e, "hipMalloc of the second render copy"));
    }
    s->copy[1].vox = s->own_copy;
    if (int rc = launch_copy(s->own_copy, ctx->rgrid.vox, ctx->rgrid.count, s->edit_stream)) return bail(rc);
    // the clock anchor: an event the host waits for, stamped with the host time
    if ((e = hipEventRecord(s->ref, s->edit_stream)) != hipSuccess) return bail(hip_fail(e, "hipEventRecord"));
    if ((e = hipEventSynchronize(s->ref)) != hipSuccess) return bail(hip_fail(e, "hipEventSynchronize"));
    s->t_ref = clock_of(s);
    s->slots.resize(cfg->streams);
    const size_t vbytes = (size_t)cfg->max_batch_views * sizeof(ViewDesc);
    for (auto& sl : s->slots) {
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess)
            return bail(hip_fail(e, "hipStreamCreate"));
        if ((e = hipEventCreate(&sl.exec)) != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
        if ((e = hipEventCreate(&sl.done)) != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
        if ((e = hipMalloc(&sl.views_dev, vbytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc"));
        if ((e = hipHostMalloc(&sl.views_host, vbytes, hipHostMallocDefault)) != hipSuccess)
            return bail(hip_fail(e, "hipHostMalloc"));
        if ((e = hipMalloc(&sl.queue, kQueueBytes)) != hipSuccess) return bail(hip_fail(e, "hipMalloc"));
        if ((e = hipMalloc(&sl.parts, (size_t)kSlotParts * sizeof(unsigned long long))) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc"));
    }
    *out = s;
    return VRT_OK;
}

int vrt_server_destroy(vrt_server* s) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    destroy_server(s);
    return VRT_OK;
}

double vrt_server_clock(vrt_server* s) { return s ? clock_of(s) : 0.0; }

int vrt_server_submit(vrt_server* s, const vrt_frame* f, int32_t width, int32_t height, double t_submit,
                      uint64_t* id) {
    if (!s || !f) return fail(VRT_ERR_INVALID, "invalid arguments");
    if (int rc = check_frame(*f)) return rc;
    if (width <= 0 || height <= 0 || (int64_t)width * height > ((int64_t)1 << 31))
        return fail(VRT_ERR_INVALID, "invalid frame size");
    Request r;
    r.id = s->next_id++;
    r.frame = *f;
    r.width = width;
    r.height = height;
    r.bytes = frame_bytes(s->cfg, width, height);
    if (r.bytes > s->cfg.framebuffer_budget) return fail(VRT_ERR_BUDGET, "request exceeds the framebuffer budget");
    r.t_submit = t_submit;
    s->reqs[r.id] = r;
    Item it;
    it.id = r.id;
    s->fifo.push_back(it);
    s->queued++;
    if (id) *id = r.id;
    return VRT_OK;
}

int vrt_server_submit_edit(vrt_server* s, int32_t cx, int32_t cy, int32_t cz, int32_t radius, uint64_t* version) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    if (radius < -3 || radius > 4096) return fail(VRT_ERR_INVALID, "radius must be in [-3, 4096]");
    Item it;
    it.is_edit = true;
    it.cx = cx;
    it.cy = cy;
    it.cz = cz;
    it.r = radius;
    s->fifo.push_back(it);
    s->version_queued++;
    if (version) *version = s->version_queued;
    return VRT_OK;
}

int vrt_server_cancel(vrt_server* s, uint64_t id) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    for (auto it = s->fifo.begin(); it != s->fifo.end(); ++it) {
        if (!it->is_edit && it->id == id) {
            s->fifo.erase(it);
            s->reqs.erase(id);
            s->queued--;
            return VRT_OK;
        }
    }
    return fail(VRT_ERR_NOTFOUND, "no queued request with this id (admitted requests run to completion)");
}

int vrt_server_pump(vrt_server* s, int32_t wait) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    VRT_HIP(hipSetDevice(s->ctx->device));
    if (int rc = retire(s)) return rc;
    int launched = 0;
    if (int rc = admit(s, &launched)) return rc;
    if (wait && launched == 0 && s->completed.empty()) {
        // nothing moved: block on the oldest in-flight batch
        Slot* oldest = nullptr;
        for (auto& sl : s->slots)
            if (sl.busy && (!oldest || s->reqs[sl.ids.front()].t_start < s->reqs[oldest->ids.front()].t_start))
                oldest = &sl;
        if (oldest) {
            VRT_HIP(hipEventSynchronize(oldest->done));
            if (int rc = retire(s)) return rc;
            if (int rc = admit(s, &launched)) return rc;
        }
    }
    return VRT_OK;
}

int vrt_server_poll(vrt_server* s, vrt_result* out, int32_t max, int32_t* n) {
    if (!s || (!out && max > 0)) return fail(VRT_ERR_INVALID, "invalid arguments");
    int32_t k = 0;
    while (k < max && !s->completed.empty()) {
        const uint64_t id = s->completed.front();
        s->completed.pop_front();
        const Request& r = s->reqs[id];
        vrt_result& o = out[k++];
        o.id = r.id;
        o.world_version = r.world_version;
        o.width = r.width;
        o.height = r.height;
        o.t_submit = r.t_submit;
        o.t_start = r.t_start;
        o.t_exec = r.t_exec;
        o.t_gpu_done = r.t_gpu_done;
        o.t_done = r.t_done;
        o.target = r.target;
    }
    if (n) *n = k;
    return VRT_OK;
}

int vrt_server_release(vrt_server* s, uint64_t id) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second.has_mem || !it->second.done)
        return fail(VRT_ERR_NOTFOUND, "no completed request with this id");
    s->arena.give(it->second.offset, it->second.span);
    s->bytes_in_use -= it->second.bytes;
    s->reqs.erase(it);
    return VRT_OK;
}

int vrt_server_status(vrt_server* s, int32_t* queued, int32_t* inflight, int32_t* done, uint64_t* bytes_in_use) {
    if (!s) return fail(VRT_ERR_INVALID, "invalid server");
    int32_t fl = 0;
    for (auto& sl : s->slots) fl += sl.busy ? (int32_t)sl.ids.size() : 0;
    if (queued) *queued = s->queued;
    if (inflight) *inflight = fl;
    if (done) *done = (int32_t)s->completed.size();
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
    return VRT_OK;
}

}  // extern "C"
