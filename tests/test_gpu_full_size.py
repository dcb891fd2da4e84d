"""The batched and served forms at their own sizes.

views64: the 64-view mixed-resolution set (320x180 .. 3840x2160, random poses,
0-16 lights each) over the 1024x256x1024 world in one tuned batch launch:
every view's outputs byte-equal to its own plain launch, and sampled pixels
(random and the view's longest primary walks) equal to the oracle.

served2g: the 1024x512x1024 world resident on the device, a Poisson stream of
requests with that resolution mix and 10 removeSphere edits (radius 1-8 at
surface points) per 100 requests, all maps kept (want_aux): every frame saw
the world version its submission order says, and its sampled pixels equal
the oracle's replay of the same edit log on the raw world; afterwards the
device grid equals the replayed grid on every cell the edits repaired.

The oracle renders from the raw world with its depth-field words computed on
demand (oracle_ctypes.LazyWorld), so nothing the GPU built enters the
expected values."""
import ctypes as C
import time

import numpy as np
import pytest

from test_gpu_large import compare_sample, sampled_pixels

pytestmark = pytest.mark.gpu

KEYS = ("rgb", "depth", "hit", "face", "shadow_eval", "shadow_occ", "steps")


def world_of(vrt, name, seed):
    cfg = vrt.CONFIGS[name]
    params = vrt.world_params(cfg["dims"], seed, caves=cfg["caves"], boxes=cfg["boxes"])
    return params, cfg["dims"], vrt.generate_world(params)


def first_difference(a, b):
    for k in KEYS:
        if k == "depth":
            ok = np.all((a[k].view(np.uint32) == b[k].view(np.uint32)) | (np.isnan(a[k]) & np.isnan(b[k])))
        else:
            ok = np.array_equal(a[k], b[k])
        if not ok:
            return k
    return None


def test_views64_batch_equals_plain_and_oracle(gpu, orc):
    params, dims, world = world_of(gpu, "views64", 4)
    views = gpu.make_views(params, dims, world, 44, 64)  # the benchmark's set
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    print("views64 world on the GPU, %d views, %d pixels" % (len(views), sum(w * h for _, w, h in views)),
          flush=True)
    plain = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]
    batched = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]

    def arr_of(targets):
        arr = (gpu.vrt_view * len(views))()
        for i, ((f, w, h), t) in enumerate(zip(views, targets)):
            arr[i].frame = f
            arr[i].target = t.struct()
        return arr

    ctx.stats(reset=True)
    gpu.call("vrt_draw_views_plain", ctx.h, arr_of(plain), len(views))
    st_plain = ctx.stats(reset=True)
    # the autotuner picks the config on this very set, then one graph launch
    best = gpu.vrt_batch_config()
    res = (gpu.vrt_tune_entry * 32)()
    n = C.c_int32(0)
    gpu.call("vrt_autotune_batch", ctx.h, arr_of(batched), len(views), 1, C.byref(best), res, 32, C.byref(n))
    print("tuned: tile %dx%d chunk %d" % (best.tile_w, best.tile_h, best.chunk_tiles), flush=True)
    for t in batched:
        t.rgb.fill(0x5A)
    b = C.c_void_p()
    gpu.call("vrt_batch_create", ctx.h, arr_of(batched), len(views), C.byref(best), C.byref(b))
    ctx.stats(reset=True)
    gpu.call("vrt_batch_launch", b)
    st_batch = ctx.stats(reset=True)
    gpu.call("vrt_batch_destroy", b)
    assert st_batch["rays"] == st_plain["rays"] and st_batch["pixels"] == st_plain["pixels"]
    lazy = orc.LazyWorld(world, dims)
    t0 = time.time()
    for i, ((f, w, h), tp, tb) in enumerate(zip(views, plain, batched)):
        gp, gb = tp.download(), tb.download()
        bad = first_difference(gb, gp)
        assert bad is None, "view %d (%dx%d): batched differs from plain in %s" % (i, w, h, bad)
        pix = sampled_pixels(gp, 2000, 100, seed=i)
        r = lazy.render(f, w, h, pixels=pix)
        compare_sample(gp, r, pix, "views64 view %d (%dx%d)" % (i, w, h))
        tp.rgb.free()
        tb.rgb.free()
    print("64 views equal plain and the oracle (%.1f s of checks)" % (time.time() - t0), flush=True)
    ctx.close()


def test_served2g_stream_with_edits_matches_oracle_replay(gpu, orc):
    params, dims, world = world_of(gpu, "served2g", 5)
    X, Y, Z = dims
    n_req = 120
    views = gpu.make_views(params, dims, world, 55, n_req)
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    print("served2g world on the GPU", flush=True)
    cfg = gpu.vrt_server_config()
    cfg.framebuffer_budget = 32 << 30
    cfg.max_batch_views = 16
    cfg.streams = 3
    cfg.want_depth = 1
    cfg.want_aux = 1
    s = C.c_void_p()
    gpu.call("vrt_server_create", ctx.h, C.byref(cfg), C.byref(s))
    rng = np.random.default_rng(56)
    # Poisson arrivals at about the server's pace (one 1080p frame ~0.2 ms)
    arrivals = np.cumsum(rng.exponential(0.5e-3, n_req))
    edits, expect_version, ids = [], {}, {}
    res = (gpu.vrt_result * 64)()
    n = C.c_int32(0)
    got = {}
    t_start = gpu.lib().vrt_server_clock(s)
    nxt = 0

    def collect():
        gpu.call("vrt_server_poll", s, res, 64, C.byref(n))
        for i in range(n.value):
            r = res[i]
            out = gpu.download_target(r.target)
            k = ids[r.id]
            fr, w, h = views[k]
            pix = sampled_pixels(out, 1000, 50, seed=k)
            got[k] = (r.world_version, pix, {key: out[key].reshape((w * h,) + out[key].shape[2:])[pix]
                                           for key in KEYS})
            gpu.call("vrt_server_release", s, r.id)

    while len(got) < n_req:
        now = gpu.lib().vrt_server_clock(s) - t_start
        while nxt < n_req and arrivals[nxt] <= now:
            if nxt % 10 == 9:  # 10 edits per 100 requests
                x, z = int(rng.integers(0, X)), int(rng.integers(0, Z))
                e = (x, gpu.surface(params, x, z), z, int(rng.integers(1, 9)))
                ver = C.c_uint64(0)
                gpu.call("vrt_server_submit_edit", s, *e, C.byref(ver))
                edits.append(e)
                assert ver.value == len(edits)
            f, w, h = views[nxt]
            rid = C.c_uint64(0)
            gpu.call("vrt_server_submit", s, C.byref(f), w, h, float(t_start + arrivals[nxt]), C.byref(rid))
            ids[rid.value] = nxt
            expect_version[nxt] = len(edits)
            nxt += 1
        gpu.call("vrt_server_pump", s, 1 if nxt == n_req else 0)
        collect()
    gpu.call("vrt_server_destroy", s)
    assert len(edits) == n_req // 10
    print("%d requests served, %d edits; replaying on the oracle" % (n_req, len(edits)), flush=True)
    lazy = orc.LazyWorld(world, dims, edits)
    for k in range(n_req):
        wv, pix, g = got[k]
        assert wv == expect_version[k], "request %d saw version %d, submitted after %d edits" % (k, wv,
                                                                                               expect_version[k])
        f, w, h = views[k]
        r = lazy.render(f, w, h, version=wv, pixels=pix)
        for key in ("hit", "face", "shadow_eval", "shadow_occ", "steps"):
            assert np.array_equal(g[key], r[key]), "request %d (%dx%d, version %d): %s differs" % (k, w, h, wv, key)
        assert np.all((g["depth"].view(np.uint32) == r["depth"].view(np.uint32)) |
                      (np.isnan(g["depth"]) & np.isnan(r["depth"]))), "request %d depth" % k
        assert int(np.abs(g["rgb"].astype(np.int16) - r["rgb"].astype(np.int16)).max()) <= 1, "request %d rgb" % k
    # the device grid after the whole log: every cell of every repair box, and random cells
    dev = ctx.download()
    cells = []
    for (cx, cy, cz, r) in edits:
        R = r + 3
        zz, yy, xx = np.meshgrid(np.arange(cz - R, cz + R), np.arange(cy - R, cy + R), np.arange(cx - R, cx + R),
                                 indexing="ij")
        ok = (xx >= 0) & (xx < X) & (yy >= 0) & (yy < Y) & (zz >= 0) & (zz < Z)
        cells.append((xx[ok] + X * yy[ok] + X * Y * zz[ok].astype(np.int64)).astype(np.int64))
    cells.append(rng.integers(0, X * Y * Z, 200000))
    idx = np.unique(np.concatenate(cells))
    exp = lazy.words(idx)
    bad = np.nonzero(dev[idx] != exp)[0]
    assert bad.size == 0, "device grid differs from the replayed log at %d of %d cells" % (bad.size, idx.size)
    ctx.close()
