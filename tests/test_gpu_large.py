"""Full-size parity through sampling and size-independent properties.

The large world (2048x512x2048, 8 GiB, 2^31 voxels, ~35% of cells below the
surface carved out) at 3840x2160 with grazing cameras and a 4096-step cap,
shaded and depth-field modes; and a world whose X*Y exceeds 2^24, which takes
the kernels' 64-bit index path."""
import numpy as np
import pytest

from test_gpu_parity import assert_frames_match, same_floats

pytestmark = pytest.mark.gpu


def sampled_pixels(g, n_random, n_longest, seed):
    """Random pixels plus the pixels with the longest primary walks."""
    H, W = g["steps"].shape
    rng = np.random.default_rng(seed)
    rand = rng.integers(0, H * W, n_random)
    longest = np.argsort(g["steps"].ravel())[-n_longest:]
    return np.unique(np.concatenate([rand, longest]))


def compare_sample(g, r, pix, label):
    for k in ("hit", "face", "shadow_eval", "shadow_occ", "steps"):
        a = g[k].reshape(-1)[pix]
        assert np.array_equal(a, r[k]), "%s: %s differs at %d sampled pixels" % (label, k, int((a != r[k]).sum()))
    assert same_floats(g["depth"].reshape(-1)[pix], r["depth"]), label + " depth"
    d = np.abs(g["rgb"].reshape(-1, 3)[pix].astype(np.int16) - r["rgb"].astype(np.int16))
    assert int(d.max()) <= 1, label + " rgb"


def test_large_world_4k_shaded_and_depth_view(gpu, orc):
    """The 8 GiB world: the device generator equals the host generator; the
    GPU depth field equals the oracle's on whole slabs through the camera and
    at the surface plus random voxels; sampled pixels (random and the longest
    walks) of both modes equal the oracle's, which renders from the raw world
    with its depth-field words computed on demand (never from the GPU's grid)."""
    sc = gpu.make_scene("large4k", seed=3)
    X, Y, Z = sc.dims
    print("large world generated", flush=True)
    # about 35% of the cells below the terrain surface are carved out (caves)
    rng = np.random.default_rng(4)
    cols = rng.integers(0, [X, Z], size=(4000, 2))
    below = empty = 0
    for x, z in cols:
        h = gpu.surface(sc.params, int(x), int(z))
        col = sc.world[int(x) + X * np.arange(h, dtype=np.int64) + X * Y * int(z)]
        below += h
        empty += int((col < 0).sum())
    frac = empty / below
    print("empty below the surface: %.3f" % frac, flush=True)
    assert 0.30 <= frac <= 0.40, frac
    ctx = gpu.Context(sc.dims)
    ctx.generate(sc.params)  # device generator
    dev = ctx.download()
    assert np.array_equal(dev, sc.world), "device world differs from host world"
    del dev
    ctx.build_depth_field()
    lazy = orc.LazyWorld(sc.world, sc.dims)
    gdf = ctx.download()
    # whole slabs: the z-slab and the x-row band through the camera, the
    # y-slab just above the camera's surface height, and 200k random voxels
    cz, cy = int(sc.frame.cam_pos[2]), int(sc.frame.cam_pos[1])
    zs = X * Y * cz + np.arange(X * Y, dtype=np.int64)
    ys = (np.arange(X, dtype=np.int64)[None, :] + X * cy + X * Y * np.arange(Z, dtype=np.int64)[:, None]).ravel()
    rnd = rng.integers(0, X * Y * Z, 200000)
    for name, idx in (("z-slab %d" % cz, zs), ("y-slab %d" % cy, ys), ("random", rnd)):
        exp = lazy.words(idx)
        bad = np.nonzero(gdf[idx] != exp)[0]
        assert bad.size == 0, "depth field differs on the %s at %d voxels" % (name, bad.size)
    del gdf
    print("depth field checked", flush=True)
    for mode in (0, 1):
        sc.frame.view_depth_field = mode
        g = gpu.render_on_gpu(sc, aux=True, ctx=ctx)
        pix = sampled_pixels(g, 3000, 300, seed=mode)
        r = lazy.render(sc.frame, sc.width, sc.height, pixels=pix)
        compare_sample(g, r, pix, "large world mode %d" % mode)
        # the grazing camera drives walks far beyond the 384 cap of the reference
        assert int(g["steps"].max()) > 100
        # size-independent: rays = pixels + shadow rays in the masks
        pc = np.unpackbits(g["shadow_eval"].view(np.uint8)).sum()
        assert g["rays"] == g["steps"].size + int(pc)
        print("mode", mode, "ok, max primary steps", int(g["steps"].max()), flush=True)
    ctx.close()


def test_wide_world_takes_the_64bit_index_path(gpu, orc):
    dims = (4097, 4097, 8)  # X*Y > 2^24
    p = gpu.world_params(dims, 9, caves=40, boxes=300)
    world = gpu.generate_world(p)
    rng = np.random.default_rng(9)
    f = gpu.make_frame(p, dims, world, rng, 320, 200, 16, 384, area=(0.3, 0.7))
    f.cam_pos[2] = 4.5
    f.rotate_matrix[:] = gpu.camera_matrix(0.25, 1.5707964).tolist()  # looking along +x
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    wdf = ctx.download()
    xyz = np.stack([rng.integers(0, dims[0], 50000), rng.integers(0, 160, 50000), rng.integers(0, 8, 50000)],
                   1).astype(np.int32)
    idx = xyz[:, 0].astype(np.int64) + dims[0] * xyz[:, 1] + dims[0] * dims[1] * xyz[:, 2].astype(np.int64)
    assert np.array_equal(wdf[idx], orc.fix_depth_values(world, dims, xyz))
    tgt = gpu.Target(320, 200, depth=True, aux=True)
    ctx.stats(reset=True)
    ctx.draw(tgt, frame=f)
    g = tgt.download()
    g.update(ctx.stats(reset=True))
    r = orc.render(wdf, dims, f, 320, 200)
    assert_frames_match(g, r, "wide world")
    assert (r["hit"] >= dims[0] * dims[1]).any()  # hits beyond the first x-y slab
    ctx.close()
