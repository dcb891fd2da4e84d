"""vrt_draw_views_noalloc (the multi-view draw that allocates nothing: its
descriptors and queue counters live in a fixed per-GPU array inside the
library) against the allocating entry point it varies, vrt_batch_create +
vrt_batch_launch with the same config: every view's outputs bit-identical,
on two calls in a row, the second with fewer and smaller views (so stale
descriptors or queue counters left by the first call would show)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("rgb", "depth", "hit", "face", "shadow_eval", "shadow_occ", "steps")
RES = [(37, 23), (64, 64), (160, 90), (200, 120), (8, 8), (1, 1), (96, 33), (320, 180)]


def _view_array(vrt, views, targets):
    arr = (vrt.vrt_view * len(views))()
    for i, ((f, w, h), t) in enumerate(zip(views, targets)):
        arr[i].frame = f
        arr[i].target = t.struct()
    return arr


def _same(a, b):
    for k in KEYS:
        if k == "depth":
            ok = np.all((a[k].view(np.uint32) == b[k].view(np.uint32)) | (np.isnan(a[k]) & np.isnan(b[k])))
        else:
            ok = np.array_equal(a[k], b[k])
        if not ok:
            return k
    return None


@pytest.fixture(scope="module")
def world(gpu):
    sc = gpu.make_scene("terrain128", seed=61)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    yield sc, ctx
    ctx.close()


@pytest.mark.parametrize("tw,th,chunk", [(8, 8, 0), (16, 4, 2), (16, 16, 1)])
def test_noalloc_views_equal_batch_launch(gpu, world, tw, th, chunk):
    sc, ctx = world
    cfg = gpu.vrt_batch_config(tw, th, chunk, 0)
    for call, n in enumerate((40, 9)):  # the second call is smaller
        print("no-alloc call %d: %d views, tile %dx%d chunk %d" % (call, n, tw, th, chunk), flush=True)
        views = gpu.make_views(sc.params, sc.dims, sc.world, 100 * call + tw + chunk, n, resolutions=RES)
        if call == 1:  # smaller views too
            views = [(f, max(1, w // 2), max(1, h // 2)) for (f, w, h) in views]
        ref_t = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]
        new_t = [gpu.Target(w, h, depth=True, aux=True) for (_, w, h) in views]
        for t in new_t:
            t.rgb.fill(0x5A)
        ref_arr = _view_array(gpu, views, ref_t)
        b = C.c_void_p()
        gpu.call("vrt_batch_create", ctx.h, ref_arr, len(views), C.byref(cfg), C.byref(b))
        gpu.call("vrt_batch_launch", b)
        ctx.sync()
        gpu.call("vrt_batch_destroy", b)
        ctx.stats(reset=True)
        gpu.call("vrt_draw_views_noalloc", ctx.h, _view_array(gpu, views, new_t), len(views), C.byref(cfg))
        ctx.sync()
        st = ctx.stats(reset=True)
        assert st["pixels"] == sum(w * h for (_, w, h) in views)
        for i, (a, r) in enumerate(zip(new_t, ref_t)):
            bad = _same(a.download(), r.download())
            assert bad is None, "call %d view %d (%dx%d) differs in %s" % (call, i, a.width, a.height, bad)


def test_noalloc_rejects_too_many_views(gpu, world):
    sc, ctx = world
    f = sc.frame
    t = gpu.Target(8, 8)
    arr = (gpu.vrt_view * 257)()
    for i in range(257):
        arr[i].frame = f
        arr[i].target = t.struct()
    assert gpu.lib().vrt_draw_views_noalloc(ctx.h, arr, 257, None) == gpu.VRT_ERR_INVALID
    assert b"256" in gpu.lib().vrt_error_string()
    bad = gpu.vrt_batch_config(8, 16, 1, 0)  # not a tile shape the kernels take
    assert gpu.lib().vrt_draw_views_noalloc(ctx.h, arr, 1, C.byref(bad)) == gpu.VRT_ERR_INVALID
