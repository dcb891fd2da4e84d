"""Batched and served forms on the GPU: every view's outputs are bit-identical
to its plain launch (and to the oracle), whatever the tile shape, queue chunk,
graph replay or batch it ran in; the server applies edits between frames in
submission order and admits requests under its framebuffer budget."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_parity import assert_frames_match

pytestmark = pytest.mark.gpu

KEYS = ("rgb", "depth", "hit", "face", "shadow_eval", "shadow_occ", "steps")


def small_views(vrt, sc, n, seed):
    res = [(37, 23), (64, 64), (160, 90), (200, 120), (8, 8), (1, 1), (96, 33)]
    return vrt.make_views(sc.params, sc.dims, sc.world, seed, n, resolutions=res)


def view_array(vrt, views, targets):
    arr = (vrt.vrt_view * len(views))()
    for i, ((f, w, h), t) in enumerate(zip(views, targets)):
        arr[i].frame = f
        arr[i].target = t.struct()
    return arr


def same_outputs(a, b):
    """Name of the first output that differs (floats: bitwise or both NaN)."""
    for k in KEYS:
        if k == "depth":
            ok = np.all((a[k].view(np.uint32) == b[k].view(np.uint32)) | (np.isnan(a[k]) & np.isnan(b[k])))
        else:
            ok = np.array_equal(a[k], b[k])
        if not ok:
            return k
    return None


@pytest.fixture(scope="module")
def scene(gpu, orc):
    sc = gpu.make_scene("terrain128", seed=21)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    wdf = orc.depth_field(sc.world, sc.dims)  # the oracle's own grid
    assert np.array_equal(ctx.download(), wdf)
    yield sc, ctx, wdf
    ctx.close()


def plain_outputs(gpu, ctx, views):
    outs = []
    for (f, w, h) in views:
        t = gpu.Target(w, h, depth=True, aux=True)
        ctx.draw(t, frame=f)
        ctx.sync()
        outs.append(t.download())
    return outs


@pytest.mark.parametrize("tw,th", [(8, 8), (16, 4), (32, 2), (16, 16)])
@pytest.mark.parametrize("chunk", [0, 1, 3, 16])
def test_batch_equals_plain(gpu, scene, tw, th, chunk):
    sc, ctx, _ = scene
    views = small_views(gpu, sc, 12, seed=tw * 100 + chunk)
    ref = plain_outputs(gpu, ctx, views)
    targets = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]
    arr = view_array(gpu, views, targets)
    cfg = gpu.vrt_batch_config(tw, th, chunk, 0)
    b = C.c_void_p()
    gpu.call("vrt_batch_create", ctx.h, arr, len(views), C.byref(cfg), C.byref(b))
    # replay the captured graph several times with other work in between
    scratch = gpu.Target(views[0][1], views[0][2], depth=True, aux=True)
    for rep in range(3):
        for t in targets:
            t.rgb.fill(0x5A)
        gpu.call("vrt_batch_launch", b)
        ctx.sync()
        for i, t in enumerate(targets):
            bad = same_outputs(t.download(), ref[i])
            assert bad is None, "view %d (%dx%d) replay %d differs in %s" % (i, t.width, t.height, rep, bad)
        ctx.draw(scratch, frame=views[0][0])  # eager work between replays
        ctx.sync()
    gpu.call("vrt_batch_destroy", b)


def test_batch_views_match_oracle(gpu, orc, scene):
    sc, ctx, wdf = scene
    views = small_views(gpu, sc, 5, seed=3)
    targets = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]
    arr = view_array(gpu, views, targets)
    ctx.stats(reset=True)
    b = C.c_void_p()
    ctx.set_shadow_cutoff(False)  # so that the DDA-step total is the reference's
    try:
        gpu.call("vrt_batch_create", ctx.h, arr, len(views), None, C.byref(b))
    finally:
        ctx.set_shadow_cutoff(True)
    gpu.call("vrt_batch_launch", b)
    st = ctx.stats(reset=True)
    rays = steps = 0
    for (f, w, h), t in zip(views, targets):
        g = t.download()
        r = orc.render(wdf, sc.dims, f, w, h)
        g["rays"], g["dda_steps"] = r["rays"], r["dda_steps"]
        assert_frames_match(g, r, "batch view %dx%d" % (w, h))
        rays += r["rays"]
        steps += r["dda_steps"]
    assert st["rays"] == rays and st["dda_steps"] == steps
    gpu.call("vrt_batch_destroy", b)


def test_autotune_picks_a_valid_config(gpu, scene):
    sc, ctx, _ = scene
    views = small_views(gpu, sc, 8, seed=9)
    targets = [gpu.Target(w, h, depth=True) for (_, w, h) in views]
    arr = view_array(gpu, views, targets)
    best = gpu.vrt_batch_config()
    res = (gpu.vrt_tune_entry * 32)()
    n = C.c_int32(0)
    gpu.call("vrt_autotune_batch", ctx.h, arr, len(views), 2, C.byref(best), res, 32, C.byref(n))
    assert n.value == 24  # 4 tile shapes x chunk 0 (hardware-scheduled), 1, 2, 4, 8, 16
    shapes = {(res[i].config.tile_w, res[i].config.tile_h) for i in range(n.value)}
    assert shapes == {(8, 8), (16, 4), (32, 2), (16, 16)}
    assert min(res[i].ms for i in range(n.value)) == min(r.ms for r in res[:n.value])
    assert (best.tile_w, best.tile_h) in shapes and best.chunk_tiles in (0, 1, 2, 4, 8, 16)


# ------------------------------------------------------------------ server
def server_config(gpu, budget, max_views=4, streams=2, aux=1):
    cfg = gpu.vrt_server_config()
    cfg.framebuffer_budget = budget
    cfg.max_batch_views = max_views
    cfg.streams = streams
    cfg.want_depth = 1
    cfg.want_aux = aux
    return cfg


def drain(gpu, s, expect):
    got = {}
    res = (gpu.vrt_result * 64)()
    n = C.c_int32(0)
    for _ in range(100000):
        gpu.call("vrt_server_pump", s, 1)
        gpu.call("vrt_server_poll", s, res, 64, C.byref(n))
        for i in range(n.value):
            r = res[i]
            got[r.id] = (r.world_version, gpu.download_target(r.target), r)
            gpu.call("vrt_server_release", s, r.id)
        if len(got) == expect:
            break
    return got


def test_server_frames_equal_oracle_replaying_the_edit_log(gpu, orc):
    sc = gpu.make_scene("terrain128", seed=31)
    X, Y, Z = sc.dims
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    wdf = orc.depth_field(sc.world, sc.dims)  # the oracle replays the log on its own grid
    views = small_views(gpu, sc, 15, seed=5)  # 5 edits: odd, so the server's own copy ends ahead
    rng = np.random.default_rng(6)
    ctx.stats(reset=True)
    s = C.c_void_p()
    gpu.call("vrt_server_create", ctx.h, C.byref(server_config(gpu, 1 << 30)), C.byref(s))
    log = []       # ("req", id, view) or ("edit", (cx,cy,cz,r))
    ids = {}
    for i, v in enumerate(views):
        if i % 3 == 2:  # an edit every third request
            e = (int(rng.integers(0, X)), int(rng.integers(20, 70)), int(rng.integers(0, Z)), int(rng.integers(1, 9)))
            ver = C.c_uint64(0)
            gpu.call("vrt_server_submit_edit", s, *e, C.byref(ver))
            log.append(("edit", e))
        rid = C.c_uint64(0)
        gpu.call("vrt_server_submit", s, C.byref(v[0]), v[1], v[2], 0.0, C.byref(rid))
        ids[rid.value] = len(log)
        log.append(("req", rid.value, v))
        if i % 4 == 0:
            gpu.call("vrt_server_pump", s, 0)  # admit part of the stream while the rest arrives
    got = drain(gpu, s, len(views))
    assert len(got) == len(views)
    served = ctx.stats(reset=True)  # every batch (and its count fold) has ended
    for rid, (_, _, r) in got.items():  # device times sit between the host's (1 ms for the clock anchor)
        assert r.t_start <= r.t_exec + 1e-3 and r.t_exec <= r.t_gpu_done and r.t_gpu_done <= r.t_done + 1e-3, rid
    # the oracle replays the same log
    world = wdf.copy()
    version = 0
    rays = steps = pixels = 0
    for entry in log:
        if entry[0] == "edit":
            orc.remove_sphere(world, sc.dims, entry[1][:3], entry[1][3])
            version += 1
            continue
        _, rid, (f, w, h) = entry
        wv, out, _ = got[rid]
        assert wv == version, "request %d saw world version %d, expected %d" % (rid, wv, version)
        r = orc.render(world, sc.dims, f, w, h)
        out["rays"], out["dda_steps"] = r["rays"], r["dda_steps"]
        assert_frames_match(out, r, "served request %d" % rid)
        rays, steps, pixels = rays + r["rays"], steps + r["dda_steps"], pixels + w * h
    # the served batches' statistics: every ray counted once, no iteration more than the oracle's
    assert served["rays"] == rays and served["pixels"] == pixels and 0 < served["dda_steps"] <= steps
    # the server's world is the oracle's after the whole log
    gpu.call("vrt_server_destroy", s)
    assert np.array_equal(ctx.download(), world)
    # and the context's own draws see every edit once the server is gone (the
    # server renders from two alternating copies of the grid)
    f, w, h = views[0]
    t = gpu.Target(w, h, depth=True, aux=True)
    ctx.stats(reset=True)
    ctx.draw(t, frame=f)
    g = t.download()
    g.update(ctx.stats(reset=True))
    assert_frames_match(g, orc.render(world, sc.dims, f, w, h), "context draw after the server")
    ctx.close()


def test_server_same_frame_alone_and_in_the_mix(gpu):
    sc = gpu.make_scene("terrain128", seed=41)
    views = small_views(gpu, sc, 10, seed=8)
    outs = []
    for mix in (False, True):
        ctx = gpu.Context(sc.dims)
        ctx.upload(sc.world)
        ctx.build_depth_field()
        s = C.c_void_p()
        gpu.call("vrt_server_create", ctx.h, C.byref(server_config(gpu, 1 << 30, max_views=8, streams=3)), C.byref(s))
        got = {}
        if mix:
            for v in views:
                rid = C.c_uint64(0)
                gpu.call("vrt_server_submit", s, C.byref(v[0]), v[1], v[2], 0.0, C.byref(rid))
            got = drain(gpu, s, len(views))
            outs.append([got[k][1] for k in sorted(got)])
        else:
            alone = []
            for v in views:
                rid = C.c_uint64(0)
                gpu.call("vrt_server_submit", s, C.byref(v[0]), v[1], v[2], 0.0, C.byref(rid))
                g = drain(gpu, s, 1)
                alone.append(g[rid.value][1])
            outs.append(alone)
        gpu.call("vrt_server_destroy", s)
        ctx.close()
    for i, (a, b) in enumerate(zip(*outs)):
        assert same_outputs(a, b) is None, "request %d differs alone vs in the mix" % i


def test_server_admission_under_budget(gpu):
    sc = gpu.make_scene("tiny", seed=2)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    f = sc.frame
    per = 64 * 64 * (3 + 4) + 6 * 256  # rgb + depth per 64x64 request, buffers 256-aligned
    s = C.c_void_p()
    gpu.call("vrt_server_create", ctx.h, C.byref(server_config(gpu, 3 * per, max_views=8, aux=0)), C.byref(s))
    rid = C.c_uint64(0)
    assert gpu.lib().vrt_server_submit(s, C.byref(f), 128, 128, 0.0, C.byref(rid)) == gpu.VRT_ERR_BUDGET
    for _ in range(5):
        gpu.call("vrt_server_submit", s, C.byref(f), 64, 64, 0.0, C.byref(rid))
    gpu.call("vrt_server_pump", s, 0)
    q, fl, d, used = C.c_int32(), C.c_int32(), C.c_int32(), C.c_uint64()
    gpu.call("vrt_server_status", s, C.byref(q), C.byref(fl), C.byref(d), C.byref(used))
    assert q.value == 2 and fl.value + d.value == 3 and used.value == 3 * per  # three admitted, two wait
    # nothing more is admitted until results are released
    res = (gpu.vrt_result * 8)()
    n = C.c_int32(0)
    done = []
    for _ in range(1000):
        gpu.call("vrt_server_pump", s, 1)
        gpu.call("vrt_server_poll", s, res, 8, C.byref(n))
        done += [res[i].id for i in range(n.value)]
        if len(done) == 3:
            break
    gpu.call("vrt_server_status", s, C.byref(q), C.byref(fl), C.byref(d), C.byref(used))
    assert q.value == 2 and used.value == 3 * per
    assert gpu.lib().vrt_server_release(s, 999) == gpu.VRT_ERR_NOTFOUND
    gpu.call("vrt_server_release", s, done[0])
    gpu.call("vrt_server_pump", s, 0)
    gpu.call("vrt_server_status", s, C.byref(q), C.byref(fl), C.byref(d), C.byref(used))
    assert q.value == 1 and used.value == 3 * per
    gpu.call("vrt_server_destroy", s)
    ctx.close()


def test_server_cancel_queued_request(gpu):
    sc = gpu.make_scene("tiny", seed=3)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    per = 64 * 64 * 7 + 6 * 256
    s = C.c_void_p()
    gpu.call("vrt_server_create", ctx.h, C.byref(server_config(gpu, per, max_views=4, aux=0)), C.byref(s))
    ids = []
    for _ in range(3):
        rid = C.c_uint64(0)
        gpu.call("vrt_server_submit", s, C.byref(sc.frame), 64, 64, 0.0, C.byref(rid))
        ids.append(rid.value)
    gpu.call("vrt_server_pump", s, 0)  # the budget admits one
    assert gpu.lib().vrt_server_cancel(s, ids[0]) == gpu.VRT_ERR_NOTFOUND  # already admitted
    gpu.call("vrt_server_cancel", s, ids[1])
    got = drain(gpu, s, 2)
    assert sorted(got) == [ids[0], ids[2]]
    q, fl, d, used = C.c_int32(), C.c_int32(), C.c_int32(), C.c_uint64()
    gpu.call("vrt_server_status", s, C.byref(q), C.byref(fl), C.byref(d), C.byref(used))
    assert q.value == 0 and fl.value == 0 and used.value == 0
    gpu.call("vrt_server_destroy", s)
    ctx.close()


def test_failures_are_reported_not_fatal(gpu):
    """Allocation failure, invalid frames, targets and edits return errors and
    leave the library usable."""
    h = C.c_void_p()
    # 2^20 x 2^20 x 2^10 voxels = 4 PiB: the grid allocation must fail cleanly
    assert gpu.lib().vrt_create(C.byref(h), 1 << 20, 1 << 20, 1 << 10, 0) == gpu.VRT_ERR_NOMEM and not h.value
    sc = gpu.make_scene("tiny", seed=5)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    f = sc.frame
    t = gpu.Target(16, 16)
    bad = gpu.vrt_frame.from_buffer_copy(f)
    bad.render_dist = -1
    assert gpu.lib().vrt_draw_frame(ctx.h, C.byref(bad), C.byref(t.struct())) == gpu.VRT_ERR_INVALID
    bad.render_dist = (1 << 24) + 1
    assert gpu.lib().vrt_set_frame(ctx.h, C.byref(bad)) == gpu.VRT_ERR_INVALID
    ts = t.struct()
    ts.rgb = None
    assert gpu.lib().vrt_draw_frame(ctx.h, C.byref(f), C.byref(ts)) == gpu.VRT_ERR_INVALID
    ts = t.struct()
    ts.width = 0
    assert gpu.lib().vrt_draw_frame(ctx.h, C.byref(f), C.byref(ts)) == gpu.VRT_ERR_INVALID
    assert gpu.lib().vrt_remove_sphere(ctx.h, 5, 5, 5, 100000) == gpu.VRT_ERR_INVALID
    cfg = gpu.vrt_batch_config(8, 16, 1, 0)  # not a supported tile shape
    arr = (gpu.vrt_view * 1)()
    arr[0].frame = f
    arr[0].target = t.struct()
    b = C.c_void_p()
    assert gpu.lib().vrt_batch_create(ctx.h, arr, 1, C.byref(cfg), C.byref(b)) == gpu.VRT_ERR_INVALID
    # a server's batch config is checked the same way (8x4 would shade no row,
    # 128x1 has no whole row per wave pass)
    for tw, th, ch in ((8, 4, 1), (128, 1, 0), (8, 8, -1)):
        sc_cfg = server_config(gpu, 1 << 24)
        sc_cfg.batch = gpu.vrt_batch_config(tw, th, ch, 0)
        srv = C.c_void_p()
        assert gpu.lib().vrt_server_create(ctx.h, C.byref(sc_cfg), C.byref(srv)) == gpu.VRT_ERR_INVALID
        assert not srv.value
    # still usable
    ctx.draw(t, frame=f)
    assert t.download()["rgb"].shape == (16, 16, 3)
    ctx.close()
