"""Benchmark: Mrays/s (primary + shadow) of the plain drop-in frame at 1080p
with the sun and 16 point lights (512x256x512 grid of 4-byte voxels), one
frame per step, inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
                    [--only plain|batched|served|large]

Prints one JSON line.  Besides the headline value it reports the dominant
kernel's roofline (algorithmic bytes per launch / its average duration from
HIP events on the launch stream, against 8 TB/s; and its VALU lane-operations
per launch, from the committed counter summary, against the nominal VALU
rate), the same frames with the cutoff off, the CPU oracle timed on
this host as a baseline, and, unless --no-extra, the batched form (64 views of
mixed resolution in one persistent launch vs 64 plain launches), the served
form (p50/p99 latency under a Poisson request stream with edits) and the
large world (8 GiB grid, 4K frames, shaded and depth view).  --only plain
runs the headline frames alone (the command the committed profiles use).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# 256 CUs x 64 lanes per clock (a wave64 VALU instruction every 4 clocks per
# SIMD) x 2.4 GHz: the nominal non-packed VALU rate
VALU_PEAK_T = 256 * 64 * 2.4e9 / 1e12
METRIC = "Mrays/s (primary+shadow) at 1080p, 16 lights; p99 frame latency batched vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load():
    import __graft_entry__ as ge
    vrt = ge.load_package()
    if not os.path.exists(vrt.LIB_PATH):
        ge.build_lib()
    vrt.load_library()
    return vrt


# the sources k_render_plain is compiled from: a counter summary applies to
# the kernel only while these are unchanged
KERNEL_SOURCES = ["voxel-rt_amd/csrc/vrt_device.h", "voxel-rt_amd/csrc/vrt_render.hip",
                  "voxel-rt_amd/csrc/vrt_internal.h"]
COUNTERS = os.path.join(ROOT, "profiles", "k_render_plain_1080p_16lights_counters.md")


def kernel_source_hash():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update(open(os.path.join(ROOT, f), "rb").read())
    return h.hexdigest()[:16]


def profiled(key):
    """A per-launch figure of k_render_plain measured with rocprofv3 --pmc, from
    the committed counter summary: "traffic_bytes_per_launch" (calibrated
    FETCH_SIZE + WRITE_SIZE), "valu_lane_ops_per_launch"
    (SQ_THREAD_CYCLES_VALU, one per active lane per VALU instruction) or
    "valu_issue_frac" (4 x SQ_INSTS_VALU over 1024 SIMDs x the launch's
    cycles).  None when the summary was
    taken on other kernel sources (its kernel_source_sha16 line differs from
    the current sources' hash): a stale figure is not reported."""
    vals = {}
    try:
        for line in open(COUNTERS):
            if ":" in line and not line.startswith(("|", " ", "#")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    except OSError:
        return None
    if vals.get("kernel_source_sha16") != kernel_source_hash() or key not in vals:
        return None
    v = vals[key]
    return float(v) if "." in v else int(v)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def grid_ptr(vrt, ctx):
    p = vrt.P()
    vrt.call("vrt_grid_device_ptr", ctx.h, vrt.C.byref(p))
    return p.value


def plain_bench(vrt, steps, warmup, cutoff_off_leg=True):
    sc = vrt.make_scene("frame1080", seed=1)
    ctx = vrt.Context(sc.dims)
    ctx.upload(sc.world)
    t0 = time.time()
    ctx.build_depth_field()
    ctx.sync()
    log("1080p world %s, depth field %.3f s" % (sc.dims, time.time() - t0))
    # One device-to-device copy of the whole grid (256 MiB read + 256 MiB
    # written by k_copy): a plain copy of known size that calibrates the
    # profiler's FETCH_SIZE / WRITE_SIZE counters for this run.
    snapshot = vrt.DeviceBuffer(4 * sc.world.size)
    vrt.hip_check(vrt.hip().hipMemcpy(snapshot.ptr, vrt.C.c_void_p(grid_ptr(vrt, ctx)), snapshot.nbytes, vrt.D2D),
                  "hipMemcpy D2D")
    # a device-to-device hipMemcpy may return before the copy is done, and the
    # context's stream does not wait for the null stream
    vrt.hip_check(vrt.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
    vrt.call("vrt_upload_device", ctx.h, snapshot.addr)
    ctx.sync()
    snapshot.free()
    tgt = vrt.Target(sc.width, sc.height, depth=True, aux=False)
    ts = tgt.struct()
    stream = ctx.stream
    for _ in range(warmup):
        ctx.draw(ts, frame=sc.frame)
    ctx.sync()
    ctx.stats(reset=True)
    evs = [vrt.Event() for _ in range(steps + 1)]
    ctx.sync()
    t_start = time.perf_counter()
    evs[0].record(stream)
    for k in range(steps):
        ctx.draw(ts, frame=sc.frame)
        evs[k + 1].record(stream)
    ctx.sync()
    wall = time.perf_counter() - t_start
    per = [evs[k + 1].ms_since(evs[k]) for k in range(steps)]
    st = ctx.stats(reset=True)
    rays_per_frame = st["rays"] // steps
    steps_per_frame = st["dda_steps"] // steps
    pixels = sc.width * sc.height
    kernel_ms = float(np.mean(per))
    total_ms = float(np.sum(per))
    value = rays_per_frame * steps / (total_ms * 1e-3) / 1e6
    bytes_per_launch = 4 * steps_per_frame + pixels * (3 + 4)
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    lane_ops = profiled("valu_lane_ops_per_launch")
    valu = None
    if lane_ops:
        valu_achieved = lane_ops / (kernel_ms * 1e-3) / 1e12
        valu = {"achieved": valu_achieved, "peak": VALU_PEAK_T, "unit": "T lane-ops/s",
                "frac": valu_achieved / VALU_PEAK_T, "lane_ops_per_launch": lane_ops, "limiter": True}
    # The same frames with the cutoff off: every output is identical,
    # the DDA walks are the reference's full length (its iteration count).
    full = None
    if cutoff_off_leg:
        full = plain_cutoff_off(vrt, ctx, sc, ts, steps)
    out = {
        "value": value, "ms_per_step": wall * 1e3 / steps, "kernel_ms": kernel_ms,
        "kernel_ms_p50": float(np.percentile(per, 50)), "kernel_ms_p99": float(np.percentile(per, 99)),
        "rays_per_frame": int(rays_per_frame), "dda_steps_per_frame": int(steps_per_frame),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled("traffic_bytes_per_launch"),
                     "bytes_per_launch": int(bytes_per_launch), "kernel": "k_render_plain",
                     "note": "achieved = algorithmic bytes (4 B per DDA iteration + 7 B per pixel) over the "
                             "kernel time; the measured HBM traffic is below them (the grid is served from L2) "
                             "and the kernel is bound by VALU throughput (valu: lane-ops against 16 lanes per SIMD "
                             "per cycle)", "valu": valu},
        "shadow_cutoff_off": full,
    }
    return sc, ctx, out


def plain_cutoff_off(vrt, ctx, sc, ts, steps):
    """The plain frame again with the cutoff off (shadow and primary rays walk
    the reference's full length): the outputs are the same, the DDA-iteration
    total is the reference's."""
    ctx.set_shadow_cutoff(False)
    ctx.draw(ts, frame=sc.frame)
    ctx.sync()
    ctx.stats(reset=True)
    evs = [vrt.Event() for _ in range(steps + 1)]
    evs[0].record(ctx.stream)
    for k in range(steps):
        ctx.draw(ts, frame=sc.frame)
        evs[k + 1].record(ctx.stream)
    ctx.sync()
    per = [evs[k + 1].ms_since(evs[k]) for k in range(steps)]
    st = ctx.stats(reset=True)
    ctx.set_shadow_cutoff(True)
    return {"value": st["rays"] / (float(np.sum(per)) * 1e-3) / 1e6, "kernel_ms": float(np.mean(per)),
            "dda_steps_per_frame": int(st["dda_steps"] // steps)}


def world_ctx(vrt, name, seed):
    cfg = vrt.CONFIGS[name]
    dims = cfg["dims"]
    params = vrt.world_params(dims, seed, caves=cfg["caves"], boxes=cfg["boxes"])
    world = vrt.generate_world(params)
    ctx = vrt.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    ctx.sync()
    return params, dims, world, ctx


def percentiles(xs):
    return {"p50": float(np.percentile(xs, 50)), "p99": float(np.percentile(xs, 99)), "max": float(np.max(xs))}


def batched_bench(vrt, reps=5):
    """64 views of resolution uniform in {320x180 .. 3840x2160}, random poses
    over a 1024x256x1024 world, 0-16 lights each: 64 plain launches vs one
    batch launch (autotuned tile shape and scheduling -- hardware-scheduled
    workgroups or persistent waves with a given queue chunk -- as a hipGraph)."""
    C = vrt.C
    params, dims, world, ctx = world_ctx(vrt, "views64", 4)
    views = vrt.make_views(params, dims, world, 44, 64)
    del world
    targets = [vrt.Target(w, h, depth=True) for (_, w, h) in views]
    arr = (vrt.vrt_view * len(views))()
    for i, ((f, w, h), t) in enumerate(zip(views, targets)):
        arr[i].frame = f
        arr[i].target = t.struct()
    stream = ctx.stream
    pixels = sum(w * h for (_, w, h) in views)
    # plain: one launch per view, back to back; per-view completion times
    vrt.call("vrt_draw_views_plain", ctx.h, arr, len(views))
    ctx.sync()
    ctx.stats(reset=True)
    plain_ms, plain_done = [], []
    for _ in range(reps):
        evs = [vrt.Event() for _ in range(len(views) + 1)]
        evs[0].record(stream)
        for i in range(len(views)):
            vrt.call("vrt_draw_frame", ctx.h, C.byref(arr[i].frame), C.byref(arr[i].target))
            evs[i + 1].record(stream)
        ctx.sync()
        plain_ms.append(evs[-1].ms_since(evs[0]))
        plain_done.append([evs[i + 1].ms_since(evs[0]) for i in range(len(views))])
    st = ctx.stats(reset=True)
    rays = st["rays"] // reps
    steps = st["dda_steps"] // reps
    # autotune, then the tuned launch as a graph
    best = vrt.vrt_batch_config()
    res = (vrt.vrt_tune_entry * 32)()
    n = C.c_int32(0)
    t0 = time.time()
    vrt.call("vrt_autotune_batch", ctx.h, arr, len(views), 3, C.byref(best), res, 32, C.byref(n))
    tune_s = time.time() - t0
    b = C.c_void_p()
    vrt.call("vrt_batch_create", ctx.h, arr, len(views), C.byref(best), C.byref(b))
    vrt.call("vrt_batch_launch", b)
    ctx.sync()
    ctx.stats(reset=True)
    batch_ms = []
    for _ in range(reps):
        e0, e1 = vrt.Event(), vrt.Event()
        e0.record(stream)
        vrt.call("vrt_batch_launch", b)
        e1.record(stream)
        ctx.sync()
        batch_ms.append(e1.ms_since(e0))
    st2 = ctx.stats(reset=True)
    assert st2["rays"] // reps == rays, "batched and plain cast different ray counts"
    vrt.call("vrt_batch_destroy", b)
    # Per resolution class: the class's plain launches (from the runs above)
    # against one tuned batch of that class alone.  A batch of all 64 views
    # cannot take less time than its largest class alone, so plain_ms over
    # the 4K class's batched time bounds the speedup any scheduling can reach.
    per_view = np.median(np.array([np.diff([0.0] + d) for d in plain_done]), axis=0)
    classes = {}
    for i, (_, w, h) in enumerate(views):
        classes.setdefault((w, h), []).append(i)
    by_class = []
    for (w, h), idx in sorted(classes.items(), key=lambda kv: kv[0][0] * kv[0][1]):
        sub = (vrt.vrt_view * len(idx))()
        for j, i in enumerate(idx):
            sub[j] = arr[i]
        bc = C.c_void_p()
        vrt.call("vrt_batch_create", ctx.h, sub, len(idx), C.byref(best), C.byref(bc))
        vrt.call("vrt_batch_launch", bc)
        ctx.sync()
        ms = []
        for _ in range(reps):
            e0, e1 = vrt.Event(), vrt.Event()
            e0.record(stream)
            vrt.call("vrt_batch_launch", bc)
            e1.record(stream)
            ctx.sync()
            ms.append(e1.ms_since(e0))
        vrt.call("vrt_batch_destroy", bc)
        pms = float(np.sum(per_view[idx]))
        by_class.append({"resolution": [w, h], "views": len(idx), "pixels": int(len(idx) * w * h),
                         "plain_ms": pms, "batched_alone_ms": float(np.median(ms)),
                         "speedup": pms / float(np.median(ms))})
    ctx.stats(reset=True)
    ctx.close()
    largest = by_class[-1]
    pm, bm = float(np.median(plain_ms)), float(np.median(batch_ms))
    done = np.median(np.array(plain_done), axis=0)
    # algorithmic bytes of the set: 4 B per DDA iteration, RGB8 + f32 depth per pixel
    alg = 4 * steps + 7 * pixels
    return {
        "workload": "64 views, resolutions uniform in {320x180,640x360,1280x720,1920x1080,3840x2160}, "
                    "random poses over a 1024x256x1024 grid (1 GiB), 0-16 lights per view",
        "pixels": int(pixels), "rays": int(rays), "dda_steps": int(steps),
        "plain_ms": pm, "batched_ms": bm, "speedup": pm / bm,
        "plain_mrays_s": rays / (pm * 1e-3) / 1e6, "batched_mrays_s": rays / (bm * 1e-3) / 1e6,
        "frame_latency_ms": {"plain": percentiles(done), "batched": percentiles([bm] * len(views))},
        "roofline": {"bound": "hbm", "bytes": int(alg), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "plain_achieved": alg / (pm * 1e-3) / 1e9, "plain_frac": alg / (pm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "batched_achieved": alg / (bm * 1e-3) / 1e9,
                     "batched_frac": alg / (bm * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "by_class": by_class,
        "speedup_bound": {"value": pm / largest["batched_alone_ms"],
                          "note": "64 plain launches over one batch of the %dx%d views alone (%d of %d pixels): "
                                  "the whole set batched cannot be faster than that class batched"
                                  % (largest["resolution"][0], largest["resolution"][1], largest["pixels"], pixels)},
        "tuned": {"tile": [best.tile_w, best.tile_h], "chunk_tiles": best.chunk_tiles, "tune_s": tune_s,
                  "sweep": [{"tile": [res[i].config.tile_w, res[i].config.tile_h],
                             "chunk": res[i].config.chunk_tiles, "ms": res[i].ms} for i in range(n.value)]},
    }


def views64_profile(vrt, launches=3):
    """For counter runs: the 64-view set drawn once as 64 plain launches (the
    first 64 k_render_plain dispatches), the autotuner's quick pass, then
    `launches` launches of the tuned batch (the last batch-kernel dispatches)."""
    C = vrt.C
    params, dims, world, ctx = world_ctx(vrt, "views64", 4)
    views = vrt.make_views(params, dims, world, 44, 64)
    del world
    targets = [vrt.Target(w, h, depth=True) for (_, w, h) in views]
    arr = (vrt.vrt_view * len(views))()
    for i, ((f, w, h), t) in enumerate(zip(views, targets)):
        arr[i].frame = f
        arr[i].target = t.struct()
    vrt.call("vrt_draw_views_plain", ctx.h, arr, len(views))
    ctx.sync()
    best = vrt.vrt_batch_config()
    vrt.call("vrt_autotune_batch", ctx.h, arr, len(views), 1, C.byref(best), None, 0, None)
    b = C.c_void_p()
    vrt.call("vrt_batch_create", ctx.h, arr, len(views), C.byref(best), C.byref(b))
    ctx.sync()
    ctx.stats(reset=True)
    for _ in range(launches):
        vrt.call("vrt_batch_launch", b)
    ctx.sync()
    st = ctx.stats(reset=True)
    vrt.call("vrt_batch_destroy", b)
    ctx.close()
    return {"tuned": {"tile": [best.tile_w, best.tile_h], "chunk_tiles": best.chunk_tiles},
            "rays_per_launch": st["rays"] // launches, "dda_steps_per_launch": st["dda_steps"] // launches,
            "pixels": int(sum(w * h for (_, w, h) in views))}


# served batches: up to 16 views and one 3840x2160 frame's pixels each, on 4 slots
SERVED_BATCH_PIXELS = 3840 * 2160
SERVED_SLOTS = 4


def served_stream(vrt, ctx, params, dims, views, rate, edit_seed):
    """One request stream through a fresh server: requests arrive at Poisson
    times of `rate` per second (rate None: all at once, the saturation run),
    a removeSphere edit (radius 1-8 at a surface point) before every tenth.
    Returns per-request times on the server's clock and the stream's totals."""
    C = vrt.C
    X, Y, Z = dims
    n = len(views)
    rng = np.random.default_rng(edit_seed)
    arrivals = np.zeros(n) if rate is None else np.cumsum(rng.exponential(1.0 / rate, n))
    cfg = vrt.vrt_server_config()
    cfg.framebuffer_budget = 32 << 30
    cfg.max_batch_views = 16
    cfg.max_batch_pixels = SERVED_BATCH_PIXELS
    cfg.streams = SERVED_SLOTS
    cfg.want_depth = 1
    cfg.want_aux = 0
    s = C.c_void_p()
    vrt.call("vrt_server_create", ctx.h, C.byref(cfg), C.byref(s))
    ctx.stats(reset=True)
    res = (vrt.vrt_result * 64)()
    got = C.c_int32(0)
    rows, edits, nxt = [], 0, 0
    after_edit = set()
    t_start = vrt.lib().vrt_server_clock(s)
    while len(rows) < n:
        now = vrt.lib().vrt_server_clock(s) - t_start
        while nxt < n and arrivals[nxt] <= now:
            if nxt % 10 == 9:  # 10 edits per 100 requests
                x, z = int(rng.integers(0, X)), int(rng.integers(0, Z))
                vrt.call("vrt_server_submit_edit", s, x, vrt.surface(params, x, z), z, int(rng.integers(1, 9)),
                         C.byref(C.c_uint64()))
                edits += 1
            f, w, h = views[nxt]
            rid = C.c_uint64()
            vrt.call("vrt_server_submit", s, C.byref(f), w, h, float(t_start + arrivals[nxt]), C.byref(rid))
            if nxt % 10 == 9:
                after_edit.add(rid.value)
            nxt += 1
        vrt.call("vrt_server_pump", s, 1 if nxt == n else 0)
        vrt.call("vrt_server_poll", s, res, 64, C.byref(got))
        for i in range(got.value):
            r = res[i]
            rows.append((r.t_submit, r.t_start, r.t_exec, r.t_gpu_done, r.t_done, r.width * r.height,
                         r.id in after_edit))
            vrt.call("vrt_server_release", s, r.id)
    wall = vrt.lib().vrt_server_clock(s) - t_start
    st = ctx.stats(reset=True)
    vrt.call("vrt_server_destroy", s)
    t = np.array([r[:5] for r in rows])
    return {"t": t, "pixels": np.array([r[5] for r in rows]), "after_edit": np.array([r[6] for r in rows]),
            "wall": wall, "rays": st["rays"], "dda_steps": st["dda_steps"], "edits": edits}


def served_summary(run, label):
    """Latency percentiles and, for the slowest 1% of requests, where their
    time went: queue (submitted -> launched by the host), device wait
    (launched -> its batch began: an edit's render-copy barrier and the
    stream's earlier work), render (on the device) and notice (device done ->
    the host saw it)."""
    t = run["t"]
    lat = (t[:, 4] - t[:, 0]) * 1e3
    parts = {"queue": t[:, 1] - t[:, 0], "device_wait": t[:, 2] - t[:, 1], "render": t[:, 3] - t[:, 2],
             "notice": t[:, 4] - t[:, 3]}
    slow = lat >= np.percentile(lat, 99)
    busy = t[:, 3].max() - t[:, 2].min()
    return {
        "load": label, "requests": int(len(lat)), "edits": int(run["edits"]),
        "latency_ms": percentiles(lat),
        "parts_ms": {k: {"p50": float(np.percentile(v, 50) * 1e3), "p99": float(np.percentile(v, 99) * 1e3)}
                     for k, v in parts.items()},
        "slowest_1pct_mean_ms": {k: float(v[slow].mean() * 1e3) for k, v in parts.items()},
        "slowest_1pct_after_edit": int(run["after_edit"][slow].sum()),
        "mrays_s": run["rays"] / run["wall"] / 1e6, "wall_s": run["wall"],
        "device_span_s": float(busy), "requests_per_s": len(lat) / run["wall"],
    }


def served_bench(vrt, n_requests=240, loads=(0.5, 0.9, 1.2, 2.0)):
    """Resident 1024x512x1024 world; render requests with the batched set's
    resolution mix, 10 removeSphere edits (radius 1-8 at surface points) per
    100 requests, 32 GB framebuffer budget.  First all requests at once (the
    server at saturation: its capacity in requests/s and Mrays/s), then
    Poisson streams offered at fractions and multiples of that capacity."""
    params, dims, world, ctx = world_ctx(vrt, "served2g", 5)
    views = vrt.make_views(params, dims, world, 55, n_requests)
    del world
    sat = served_stream(vrt, ctx, params, dims, views, None, 56)
    cap_rps = len(views) / sat["wall"]
    out = {"workload": "resident 1024x512x1024 grid (2 GiB), %d requests per stream, resolutions as the batched "
                       "set, 10 removeSphere edits (r 1-8) per 100 requests, 32 GB framebuffer budget, %d slots of up "
                       "to 16 views and %d pixels" % (n_requests, SERVED_SLOTS, SERVED_BATCH_PIXELS),
           "saturation": served_summary(sat, "all at once")}
    out["capacity"] = {"requests_per_s": cap_rps, "mrays_s": sat["rays"] / sat["wall"] / 1e6,
                       "mrays_s_device_span": sat["rays"] / out["saturation"]["device_span_s"] / 1e6}
    alg = 4 * sat["dda_steps"] + 7 * int(sat["pixels"].sum())
    out["roofline"] = {"bound": "hbm", "bytes": int(alg), "achieved": alg / sat["wall"] / 1e9, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": alg / sat["wall"] / 1e9 / HBM_PEAK_GBS,
                       "note": "algorithmic bytes of the saturation run over its wall clock"}
    out["loads"] = []
    for k, load in enumerate(loads):
        run = served_stream(vrt, ctx, params, dims, views, load * cap_rps, 57 + k)
        out["loads"].append(served_summary(run, "%.1fx capacity (%.0f req/s)" % (load, load * cap_rps)))
    # one removeSphere (carve + depth-field repair + render-copy box) on its
    # own, device time between events on the context's stream
    rng = np.random.default_rng(58)
    X, Y, Z = dims
    e0, e1 = vrt.Event(), vrt.Event()
    edit_us = []
    for _ in range(40):
        x, z = int(rng.integers(8, X - 8)), int(rng.integers(8, Z - 8))
        y = vrt.surface(params, x, z)
        e0.record(ctx.stream)
        ctx.remove_sphere(x, y, z, int(rng.integers(1, 9)))
        e1.record(ctx.stream)
        ctx.sync()
        edit_us.append(e1.ms_since(e0) * 1e3)
    ctx.close()
    out["remove_sphere_us"] = {"p50": float(np.percentile(edit_us, 50)), "p99": float(np.percentile(edit_us, 99)),
                               "sample": "40 edits, radius 1-8 at surface points, carve + repair + copy"}
    return out


def large_bench(vrt, frames=10):
    """The large world (2048x512x2048, 8 GiB, 2^31 voxels, ~35% carved out
    below the surface) generated on the device, its depth field, and 4K
    frames from a grazing camera with a 4096-step cap, shaded and in the
    depth-field view (the same scene the large-world parity test checks)."""
    sc = vrt.make_scene("large4k", seed=3)
    ctx = vrt.Context(sc.dims)
    t0 = time.perf_counter()
    ctx.generate(sc.params)
    ctx.sync()
    gen_ms = (time.perf_counter() - t0) * 1e3
    del sc.world
    t0 = time.perf_counter()
    ctx.build_depth_field()
    ctx.sync()
    df_ms = (time.perf_counter() - t0) * 1e3
    tgt = vrt.Target(sc.width, sc.height, depth=True, aux=False)
    ts = tgt.struct()
    out = {"workload": "2048x512x2048 grid (8 GiB) generated on the device, 3840x2160, sun + 16 lights, "
                       "render distance 4096, grazing camera",
           "generate_ms": gen_ms, "depth_field_ms": df_ms}
    for mode, key in ((0, "shaded"), (1, "depth_view")):
        sc.frame.view_depth_field = mode
        ctx.draw(ts, frame=sc.frame)
        ctx.sync()
        ctx.stats(reset=True)
        evs = [vrt.Event() for _ in range(frames + 1)]
        evs[0].record(ctx.stream)
        for k in range(frames):
            ctx.draw(ts, frame=sc.frame)
            evs[k + 1].record(ctx.stream)
        ctx.sync()
        per = [evs[k + 1].ms_since(evs[k]) for k in range(frames)]
        st = ctx.stats(reset=True)
        ms = float(np.mean(per))
        out[key] = {"kernel_ms": ms, "rays_per_frame": st["rays"] // frames,
                    "dda_steps_per_frame": st["dda_steps"] // frames,
                    "mrays_s": st["rays"] / (float(np.sum(per)) * 1e-3) / 1e6}
    ctx.close()
    del ts, tgt
    return out


def cpu_baseline(vrt, sc, ctx):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as orc  # the oracle, timed as the CPU baseline only
    wdf = ctx.download()
    n = orc.threads()
    orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height, nthreads=n)  # warm-up
    # whole frames until about 10 s of wall time (at least 3)
    times, rays = [], 0
    t_start = time.perf_counter()
    while len(times) < 3 or time.perf_counter() - t_start < 10.0:
        t = time.perf_counter()
        r = orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height, nthreads=n)
        times.append(time.perf_counter() - t)
        rays = r["rays"]
    med = float(np.median(times))
    return {"value": rays / med / 1e6, "unit": "Mrays/s", "cores": n, "kind": "port", "cpu_model": cpu_model(),
            "sample": "full 1920x1080 frames of the same scene for ~10 s after 1 warm-up: median of %d "
                      "(%.3f s/frame)" % (len(times), med)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--only", choices=["plain", "batched", "served", "large", "views64_profile"], default=None,
                    help="run only this leg (for profiling); prints its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single MI355X (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    vrt = load()
    if args.only == "views64_profile":
        print(json.dumps(views64_profile(vrt)), flush=True)
        return
    if args.only == "batched":
        print(json.dumps(batched_bench(vrt, reps=2)), flush=True)
        return
    if args.only == "served":
        print(json.dumps(served_bench(vrt)), flush=True)
        return
    if args.only == "large":
        print(json.dumps(large_bench(vrt)), flush=True)
        return
    # --only plain: the headline frames alone (what the committed profiles measure)
    sc, ctx, plain = plain_bench(vrt, args.steps, args.warmup, cutoff_off_leg=args.only != "plain")
    log("plain: %.1f Mrays/s, %.3f ms/frame" % (plain["value"], plain["kernel_ms"]))
    if plain["shadow_cutoff_off"]:
        log("plain with the cutoff off: %.1f Mrays/s, %.3f ms/frame" % (
            plain["shadow_cutoff_off"]["value"], plain["shadow_cutoff_off"]["kernel_ms"]))
    res = {
        "metric": METRIC, "value": plain["value"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": plain["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain drop-in frame: 512x256x512 grid (256 MiB) value-noise terrain + 512 caves "
                               "+ 2000 boxes, 1920x1080, sun + 16 point lights, RGB8 + f32 depth",
                   "grid": list(sc.dims), "width": sc.width, "height": sc.height, "lights": 16,
                   "render_dist": int(sc.frame.render_dist), "parallelism": "single GPU"},
        "kernel_ms": plain["kernel_ms"], "kernel_ms_p50": plain["kernel_ms_p50"],
        "kernel_ms_p99": plain["kernel_ms_p99"], "rays_per_frame": plain["rays_per_frame"],
        "dda_steps_per_frame": plain["dda_steps_per_frame"], "roofline": plain["roofline"],
        "shadow_cutoff_off": plain["shadow_cutoff_off"],
        "cpu_baseline": None,
    }
    if not args.no_cpu and args.only != "plain":
        res["cpu_baseline"] = cpu_baseline(vrt, sc, ctx)
    ctx.close()
    del sc
    if not args.no_extra and args.only != "plain":
        res["batched"] = batched_bench(vrt)
        log("batched: plain %.2f ms, batched %.2f ms (x%.2f); HBM roofline %.2f vs %.2f" % (
            res["batched"]["plain_ms"], res["batched"]["batched_ms"], res["batched"]["speedup"],
            res["batched"]["roofline"]["plain_frac"], res["batched"]["roofline"]["batched_frac"]))
        res["served"] = served_bench(vrt)
        sv = res["served"]
        log("served: capacity %.0f req/s, %.0f Mrays/s at saturation; " % (
            sv["capacity"]["requests_per_s"], sv["capacity"]["mrays_s"]) + ", ".join(
            "%s p50 %.2f p99 %.2f ms" % (L["load"], L["latency_ms"]["p50"], L["latency_ms"]["p99"])
            for L in sv["loads"]))
        res["large"] = large_bench(vrt)
        log("large: shaded %.2f ms (%.0f Mrays/s), depth view %.2f ms, depth field %.0f ms" % (
            res["large"]["shaded"]["kernel_ms"], res["large"]["shaded"]["mrays_s"],
            res["large"]["depth_view"]["kernel_ms"], res["large"]["depth_field_ms"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
