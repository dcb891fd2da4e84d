"""Ray/step statistics of plain draws: the per-tile partial counts, their fold
into the counters (when the partial-count array fills up and when stats are
read), and the atomic path a draw takes under stream capture."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARTS_CAP = 1 << 22  # entries of the context's partial-count array (kPartsCap)


def test_stats_across_partial_count_folds(gpu, orc):
    """70 draws of a 2048x2048 frame (65,536 tiles each) fill the 2^22-entry
    array after 64 draws, so the 65th folds it in the middle of the run; the
    totals are 70 times one draw's, and one draw's equal the oracle's."""
    sc = gpu.make_scene("tiny", seed=1)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    ctx.set_shadow_cutoff(False)  # the oracle's iteration totals
    W = H = 2048
    assert (W // 8) * (H // 8) * 64 == PARTS_CAP
    t = gpu.Target(W, H, depth=False, aux=False)
    ctx.stats(reset=True)
    ctx.draw(t, frame=sc.frame)
    one = ctx.stats(reset=True)
    wdf = orc.depth_field(sc.world, sc.dims)
    r = orc.render(wdf, sc.dims, sc.frame, 256, 256)  # rays of a small frame, below
    assert one["pixels"] == W * H
    n = 70
    for _ in range(n):
        ctx.draw(t, frame=sc.frame)
    many = ctx.stats(reset=True)
    for k in ("rays", "dda_steps", "pixels"):
        assert many[k] == n * one[k], (k, many[k], one[k])
    # a frame the oracle can render: totals through the partial counts equal its own
    s = gpu.Target(256, 256, depth=False, aux=False)
    ctx.draw(s, frame=sc.frame)
    st = ctx.stats(reset=True)
    assert st["rays"] == r["rays"] and st["dda_steps"] == r["dda_steps"] and st["pixels"] == 256 * 256
    ctx.close()


def _graph_api(h):
    P = C.c_void_p
    h.hipStreamBeginCapture.argtypes = [P, C.c_int]
    h.hipStreamEndCapture.argtypes = [P, C.POINTER(P)]
    h.hipGraphInstantiate.argtypes = [C.POINTER(P), P, P, P, C.c_size_t]
    h.hipGraphLaunch.argtypes = [P, P]
    h.hipGraphExecDestroy.argtypes = [P]
    h.hipGraphDestroy.argtypes = [P]
    h.hipStreamSynchronize.argtypes = [P]
    return h


def test_plain_draw_captured_in_a_graph(gpu):
    """A plain draw captured into a HIP graph counts its rays on every replay
    (the captured draw adds into the counters itself), and every replay's
    frame equals the eager draw's; eager draws between the replays go through
    the partial counts and are counted too."""
    h = _graph_api(gpu.hip())
    sc = gpu.make_scene("terrain128", seed=4)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    t = gpu.Target(sc.width, sc.height, depth=True, aux=True)
    ctx.stats(reset=True)
    ctx.draw(t, frame=sc.frame)
    one = ctx.stats(reset=True)
    ref = t.download()
    stream = C.c_void_p(ctx.stream)
    graph, exe = C.c_void_p(), C.c_void_p()
    gpu.hip_check(h.hipStreamBeginCapture(stream, 2), "hipStreamBeginCapture")  # relaxed
    ctx.draw(t, frame=sc.frame)
    gpu.hip_check(h.hipStreamEndCapture(stream, C.byref(graph)), "hipStreamEndCapture")
    gpu.hip_check(h.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "hipGraphInstantiate")
    other = gpu.Target(97, 61, depth=False, aux=False)
    replays = 4
    for k in range(replays):
        t.rgb.fill(0)
        gpu.hip_check(h.hipGraphLaunch(exe, stream), "hipGraphLaunch")
        gpu.hip_check(h.hipStreamSynchronize(stream), "hipStreamSynchronize")
        got = t.download()
        for key in ("rgb", "hit", "face", "shadow_eval", "shadow_occ", "steps"):
            assert np.array_equal(got[key], ref[key]), (k, key)
        ctx.draw(other, frame=sc.frame)  # eager work between replays
    st = ctx.stats(reset=True)
    ctx.draw(other, frame=sc.frame)
    small = ctx.stats(reset=True)
    for key in ("rays", "dda_steps", "pixels"):
        assert st[key] == replays * (one[key] + small[key]), (key, st[key], one[key], small[key])
    gpu.hip_check(h.hipGraphExecDestroy(exe), "hipGraphExecDestroy")
    gpu.hip_check(h.hipGraphDestroy(graph), "hipGraphDestroy")
    ctx.close()


def test_stats_follow_a_stream_switch(gpu):
    """Draws queued on the context's own stream, then more on a caller's
    stream after vrt_set_stream: the totals read on the new stream count all
    of them (the new stream waits for the old one's work, folds included)."""
    h = gpu.hip()
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    sc = gpu.make_scene("terrain128", seed=1)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    t = gpu.Target(sc.width, sc.height, depth=False, aux=False)
    ctx.stats(reset=True)
    ctx.draw(t, frame=sc.frame)
    one = ctx.stats(reset=True)
    user = C.c_void_p()
    gpu.hip_check(h.hipStreamCreate(C.byref(user)), "hipStreamCreate")
    for _ in range(5):
        ctx.draw(t, frame=sc.frame)
    gpu.call("vrt_set_stream", ctx.h, user)
    assert ctx.stream == user.value
    for _ in range(3):
        ctx.draw(t, frame=sc.frame)
    st = ctx.stats(reset=True)
    for k in ("rays", "dda_steps", "pixels"):
        assert st[k] == 8 * one[k], (k, st[k], one[k])
    gpu.call("vrt_set_stream", ctx.h, None)
    gpu.hip_check(h.hipStreamSynchronize(user), "hipStreamSynchronize")
    ctx.close()
    gpu.hip_check(h.hipStreamDestroy(user), "hipStreamDestroy")


def test_stats_with_a_render_distance_beyond_the_packed_sums(gpu, orc):
    """render_dist 4,000,000 >= the packed sums' bound (64 x 18 x lim < 2^32):
    the plain draw and a hardware-scheduled batch add their counts by atomics
    instead, and the totals still equal the oracle's."""
    sc = gpu.make_scene("tiny", seed=2)
    sc.frame.render_dist = 4_000_000
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    ctx.set_shadow_cutoff(False)
    w, h = 96, 64
    r = orc.render(orc.depth_field(sc.world, sc.dims), sc.dims, sc.frame, w, h)
    t = gpu.Target(w, h, depth=True, aux=True)
    ctx.stats(reset=True)
    ctx.draw(t, frame=sc.frame)
    g = t.download()
    g.update(ctx.stats(reset=True))
    assert g["rays"] == r["rays"] and g["dda_steps"] == r["dda_steps"] and g["pixels"] == w * h
    assert np.array_equal(g["hit"], r["hit"]) and np.array_equal(g["steps"], r["steps"])
    # the same frame twice in one hardware-scheduled batch (8x8 tiles, chunk 0)
    arr = (gpu.vrt_view * 2)()
    targets = [gpu.Target(w, h, depth=False, aux=False) for _ in range(2)]
    for i in range(2):
        arr[i].frame = sc.frame
        arr[i].target = targets[i].struct()
    cfg = gpu.vrt_batch_config()
    cfg.tile_w, cfg.tile_h, cfg.chunk_tiles = 8, 8, 0
    b = C.c_void_p()
    gpu.call("vrt_batch_create", ctx.h, arr, 2, C.byref(cfg), C.byref(b))
    ctx.stats(reset=True)
    gpu.call("vrt_batch_launch", b)
    st = ctx.stats(reset=True)
    gpu.call("vrt_batch_destroy", b)
    assert st["rays"] == 2 * r["rays"] and st["dda_steps"] == 2 * r["dda_steps"]
    ctx.close()
