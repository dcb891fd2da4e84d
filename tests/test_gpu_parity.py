"""GPU parity: libvrt's HIP kernels against the CPU oracle, bit for bit.

Bar: hit index, face, shadow-evaluated and shadow-occluded bits, primary step
counts, ray totals and depth-field words bit-exact; RGB8 within 1 LSB per
channel (tolerance written in assert_frames_match); DDA-step totals equal to
the oracle's with the shadow cutoff off, and never above them with it on."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RGB_TOL = 1  # LSB per channel


def same_floats(a, b):
    """Bitwise equal, or NaN on both sides: IEEE 754 leaves the sign and payload
    of a NaN made by an invalid operation open (x86 gives 0xFFC00000, gfx950
    0x7FC00000), and no comparison or int() conversion depends on them."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def assert_frames_match(g, r, label=""):
    for k in ("hit", "face", "shadow_eval", "shadow_occ", "steps"):
        if k not in g:  # a target without that map
            continue
        bad = np.argwhere(g[k] != r[k])
        assert bad.size == 0, "%s %s differs at %d pixels, first %s: gpu %s oracle %s" % (
            label, k, len(bad), bad[0].tolist(), g[k][tuple(bad[0])], r[k][tuple(bad[0])])
    if "depth" in g:
        assert same_floats(g["depth"], r["depth"]), label + " depth differs"
    d = np.abs(g["rgb"].astype(np.int16) - r["rgb"].astype(np.int16))
    assert int(d.max()) <= RGB_TOL, "%s rgb differs by %d" % (label, int(d.max()))
    assert g["rays"] == r["rays"], "%s rays %d vs %d" % (label, g["rays"], r["rays"])
    # the shadow cutoff stops shadow rays early, never late; without it the
    # iteration totals are the reference's
    if g.get("shadow_cutoff", True):
        assert g["dda_steps"] <= r["dda_steps"], "%s steps %d > %d" % (label, g["dda_steps"], r["dda_steps"])
    else:
        assert g["dda_steps"] == r["dda_steps"], "%s steps %d vs %d" % (label, g["dda_steps"], r["dda_steps"])


@pytest.mark.parametrize("name", ["tiny", "terrain128"])
def test_depth_field_build_matches_oracle(gpu, orc, name):
    sc = gpu.make_scene(name, seed=3)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    got = ctx.download()
    ref = orc.depth_field(sc.world, sc.dims)
    assert np.array_equal(got, ref), "depth field differs at %d voxels" % int((got != ref).sum())
    ctx.close()


@pytest.mark.parametrize("name,seed,mode", [("tiny", 1, 0), ("tiny", 2, 1), ("terrain128", 1, 0),
                                            ("terrain128", 4, 0), ("terrain128", 5, 1)])
def test_frame_matches_oracle(gpu, orc, name, seed, mode):
    print("frame case", name, seed, mode, flush=True)
    sc = gpu.make_scene(name, seed=seed, depth_view=mode)
    wdf = orc.depth_field(sc.world, sc.dims)
    g = gpu.render_on_gpu(sc, aux=True)
    r = orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height)
    assert_frames_match(g, r, "%s/%d/%d" % (name, seed, mode))
    assert (r["hit"] >= 0).mean() > 0.2  # the frame sees terrain


@pytest.mark.parametrize("name,seed", [("tiny", 1), ("terrain128", 4)])
def test_shadow_cutoff_changes_only_the_step_total(gpu, orc, name, seed):
    """Every output is the same with the shadow cutoff on and off; off, the
    DDA-step total is the oracle's; on, it is smaller (sun rays rise above
    the terrain and stop)."""
    sc = gpu.make_scene(name, seed=seed)
    wdf = orc.depth_field(sc.world, sc.dims)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    on = gpu.render_on_gpu(sc, aux=True, ctx=ctx, shadow_cutoff=True)
    off = gpu.render_on_gpu(sc, aux=True, ctx=ctx, shadow_cutoff=False)
    r = orc.render(wdf, sc.dims, sc.frame, sc.width, sc.height)
    assert_frames_match(off, r, "cutoff off")
    assert_frames_match(on, r, "cutoff on")
    for k in ("rgb", "depth", "hit", "face", "shadow_eval", "shadow_occ", "steps"):
        assert on[k].tobytes() == off[k].tobytes(), k
    assert on["dda_steps"] < off["dda_steps"] == r["dda_steps"]
    ctx.close()


@pytest.mark.parametrize("how", ["upload_region", "place_voxel", "device_pointer"])
def test_shadow_cutoff_follows_new_solid_voxels(gpu, orc, how):
    """A roof of solid voxels added 10 cells above the terrain's top must shade
    what lies under it: sun rays jump through the gap above the old top, so a
    solid-top bound that did not rise with the new voxels would end them there
    as misses.  One case per way of adding voxels."""
    sc = gpu.make_scene("tiny", seed=1)
    X, Y, Z = sc.dims
    world = sc.world.copy()
    old_top = int(np.nonzero((world.reshape(Z, Y, X) >= 0).any(axis=(0, 2)))[0].max())
    ry = min(Y - 2, old_top + 10)
    assert sc.frame.cam_pos[1] < ry
    ctx = gpu.Context(sc.dims)
    ctx.upload(world)
    roof = world.reshape(Z, Y, X).copy()
    # x in [0, X), z in [0, Z-1): what one vrt_upload_region call with an
    # in-grid end corner can cover (the reference's updatePartialGeometry rule)
    roof[:Z - 1, ry, :] = 0x808080
    roof = roof.reshape(-1)
    if how == "upload_region":
        ctx.upload_region(roof, (0, ry, 0), (X - 1, ry + 1, Z - 1))
    elif how == "place_voxel":
        for z in range(Z - 1):
            for x in range(X):
                ctx.place_voxel(x, ry, z, 0x808080)
    else:
        p = gpu.P()
        gpu.call("vrt_grid_device_ptr", ctx.h, gpu.C.byref(p))
        buf = gpu.DeviceBuffer(4 * roof.size)
        buf.upload(roof)
        gpu.hip_check(gpu.hip().hipMemcpy(p, buf.ptr, buf.nbytes, gpu.D2D), "hipMemcpy D2D")
        gpu.hip_check(gpu.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        ctx.grid_modified()
    ctx.build_depth_field()
    g = gpu.render_on_gpu(sc, aux=True, ctx=ctx)
    assert np.array_equal(ctx.download() >= 0, roof >= 0)
    r = orc.render(orc.depth_field(roof, sc.dims), sc.dims, sc.frame, sc.width, sc.height)
    assert_frames_match(g, r, "roof via " + how)
    # the roof shades many pixels that the open sky left lit
    r0 = orc.render(orc.depth_field(world, sc.dims), sc.dims, sc.frame, sc.width, sc.height)
    newly = (r["shadow_occ"] & 1) & ~(r0["shadow_occ"] & 1)
    assert int(newly.sum()) > 1000
    ctx.close()


def test_cast_rays_match_oracle(gpu, orc):
    sc = gpu.make_scene("tiny", seed=7)
    wdf = orc.depth_field(sc.world, sc.dims)
    rng = np.random.default_rng(11)
    X, Y, Z = sc.dims
    n = 20000
    start = np.stack([rng.uniform(0, X, n), rng.uniform(30, Y, n), rng.uniform(0, Z, n)], 1).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # edge cases: zero and signed-zero components, axis-aligned, integer starts,
    # starts on cell faces, starts outside the grid, zero budget
    special = []
    for axis in range(3):
        for s in (1.0, -1.0):
            v = np.zeros(3, np.float32)
            v[axis] = s
            special.append(([40.5, 45.5, 40.5], v))
            special.append(([40.0, 45.0, 40.0], v))
            w = v.copy()
            w[(axis + 1) % 3] = -0.0
            special.append(([40.25, 45.75, 40.5], w))
    special += [([40.5, 45.5, 40.5], [0.6, -0.8, 0.0]), ([48.0, 50.0, 48.0], [0.0, -0.6, 0.8]),
                ([10.0, 40.0, 10.0], [0.57735026, -0.57735026, 0.57735026]), ([-5.0, 40.0, 10.0], [1.0, 0.0, 0.0]),
                ([40.5, 63.999, 40.5], [1e-30, -1.0, 1e-30]), ([0.0, 0.0, 0.0], [1.0, 1.0, 1.0])]
    ss = np.array([s for s, _ in special], np.float32)
    dd = np.array([v for _, v in special], np.float32)
    start = np.concatenate([start, ss])
    d = np.concatenate([d, dd])
    dist = rng.integers(0, 400, start.shape[0]).astype(np.int32)
    dist[-len(special):] = 384
    dist[:50] = 0
    ctx = gpu.Context(sc.dims)
    ctx.upload(wdf)
    g = ctx.cast_rays(start, d, dist, 384)
    r = orc.cast_rays(wdf, sc.dims, start, d, dist, 384)
    for k in ("hit", "face", "steps"):
        bad = np.nonzero(g[k] != r[k])[0]
        assert bad.size == 0, "%s differs for %d rays, first %d" % (k, bad.size, bad[0])
    assert same_floats(g["hit_pos"], r["hit_pos"])
    ctx.close()


def test_remove_sphere_sequence_matches_oracle(gpu, orc):
    sc = gpu.make_scene("terrain128", seed=9)
    X, Y, Z = sc.dims
    ref = orc.depth_field(sc.world, sc.dims)
    ctx = gpu.Context(sc.dims)
    ctx.upload(ref)
    rng = np.random.default_rng(5)
    edits = [(int(rng.integers(0, X)), int(rng.integers(0, 60)), int(rng.integers(0, Z)), int(rng.integers(1, 9)))
             for _ in range(12)]
    edits += [(0, 0, 0, 7), (X - 1, Y - 1, Z - 1, 8), (X // 2, 30, Z // 2, 0), (-3, 20, 5, 6)]
    for (cx, cy, cz, r) in edits:
        orc.remove_sphere(ref, sc.dims, (cx, cy, cz), r)
        ctx.remove_sphere(cx, cy, cz, r)
    got = ctx.download()
    assert np.array_equal(got, ref), "world differs at %d voxels after edits" % int((got != ref).sum())
    # the edited world still renders identically
    g = gpu.render_on_gpu(sc, aux=True, ctx=ctx)
    r = orc.render(ref, sc.dims, sc.frame, sc.width, sc.height)
    assert_frames_match(g, r, "after edits")
    ctx.close()


def test_place_destroy_and_region_upload(gpu, orc):
    sc = gpu.make_scene("tiny", seed=4)
    X, Y, Z = sc.dims
    ref = orc.depth_field(sc.world, sc.dims)
    ctx = gpu.Context(sc.dims)
    ctx.upload(ref)
    for (x, y, z, v) in [(5, 5, 5, 0x123456), (X - 1, 0, Z - 1, 0x00FF00), (X, 0, 0, 7), (-1, 3, 3, 9)]:
        orc.place_voxel(ref, sc.dims, x, y, z, v)
        ctx.place_voxel(x, y, z, v)
    for (x, y, z) in [(6, 6, 6), (0, 0, 0), (X + 4, 1, 1)]:
        orc.destroy_voxel(ref, sc.dims, x, y, z)
        ctx.destroy_voxel(x, y, z)
    assert np.array_equal(ctx.download(), ref)
    # host edits then a partial upload of the box around them
    host = ref.copy()
    for (x, y, z) in [(20, 10, 20), (21, 11, 22), (30, 12, 25)]:
        host[x + X * y + X * Y * z] = -1
    ctx.upload_region(host, (15, 5, 15), (35, 20, 30))
    expect = ref.copy()
    hv = host.reshape(Z, Y, X)
    ev = expect.reshape(Z, Y, X)
    ev[15:30, 5:20, 15:36] = hv[15:30, 5:20, 15:36]
    assert np.array_equal(ctx.download(), expect)
    # start x outside the grid: getVoxelIndex(-1, y, z) is -1 for every row,
    # so updatePartialGeometry uploads nothing
    host2 = host.copy()
    host2[host2 >= 0] = 0x010203
    ctx.upload_region(host2, (-1, 0, 0), (5, 3, 3))
    assert np.array_equal(ctx.download(), expect)
    ctx.close()


def test_region_upload_with_float_corners(gpu):
    """updatePartialGeometry with float corners (src/render.cpp:204-223): the
    float loops' rows and the run length int(end.x - start.x) + 1, on cases
    whose rows follow from the rules by hand."""
    X, Y, Z = 24, 8, 6
    rng = np.random.default_rng(5)
    host = rng.integers(0, 1 << 24, X * Y * Z).astype(np.int32)
    f32 = np.float32

    def run(start, end):
        ctx = gpu.Context((X, Y, Z))  # every cell -1 until uploaded
        ctx.upload_region_f(host, start, end)
        got = ctx.download().reshape(Z, Y, X)
        ctx.close()
        return got

    def expect(cells):
        e = np.full((Z, Y, X), -1, np.int32)
        h = host.reshape(Z, Y, X)
        for (x, y, z) in cells:
            e[z, y, x] = h[z, y, x]
        return e

    # i = 0.5, 1.5, 2.5 (< 2.9) -> z 0..2; j = 0.5, 1.5 (< 2.2) -> y 0, 1; run int(3.2) + 1 = 4 from x 0
    got = run((0.5, 0.5, 0.5), (3.7, 2.2, 2.9))
    assert np.array_equal(got, expect([(x, y, z) for z in range(3) for y in range(2) for x in range(4)]))
    # j = 0.99999994 -> row 0; j + 1 = 1.99999994 rounds (to even) to 2.0 -> row 2; 3.0 -> row 3: row 1 skipped
    j0 = float(np.nextafter(f32(1.0), f32(0.0)))
    got = run((2.0, j0, 1.0), (4.0, 3.5, 2.0))
    assert np.array_equal(got, expect([(x, y, 1) for y in (0, 2, 3) for x in range(2, 5)]))
    # corners given high to low: swapped first (getVoxelIndex of start above end's)
    got = run((4.0, 3.5, 2.0), (2.0, 0.0, 1.0))
    assert np.array_equal(got, expect([(x, y, 1) for y in range(4) for x in range(2, 5)]))
    # int(start.x) = -1: every row's getVoxelIndex is -1, nothing is uploaded
    assert np.array_equal(run((-1.5, 0.0, 0.0), (5.0, 3.0, 3.0)), expect([]))
    # start and end both outside (no swap): rows y = 0, 1 of z = 0 (j = -1 skipped), each a run of
    # int(3.0) + 1 = 4 words from x = X - 2, which runs on into the next row
    got = run((X - 2.0, -1.0, 0.0), (X + 1.0, 2.0, 1.0))
    cells = [(X - 2, 0, 0), (X - 1, 0, 0), (0, 1, 0), (1, 1, 0), (X - 2, 1, 0), (X - 1, 1, 0), (0, 2, 0), (1, 2, 0)]
    assert np.array_equal(got, expect(cells))
    # whole rows of whole slabs (runs that meet end to end): z = 1, 2 entirely
    got = run((0.0, 0.0, 1.0), (X - 1.0, Y - 0.5, 3.0))
    assert np.array_equal(got, expect([(x, y, z) for z in (1, 2) for y in range(Y) for x in range(X)]))
    # both corners outside (no swap); runs of int(2X - (X - 2)) + 1 = X + 3 words from
    # x = X - 2 of rows 0 and 1 (j = -1 skipped), the second starting inside the first:
    # cells X - 2 .. 3X of slab 0 in row-major order
    got = run((X - 2.0, -1.0, 0.0), (2.0 * X, 2.0, 1.0))
    cells = [(k % X, k // X, 0) for k in range(X - 2, 3 * X + 1)]
    assert np.array_equal(got, expect(cells))
    # corners whose float loops would not end are refused
    ctx = gpu.Context((X, Y, Z))
    w = np.ascontiguousarray(host)
    s = (gpu.C.c_float * 3)(0.0, 0.0, 0.0)
    e = (gpu.C.c_float * 3)(3.0, 3.0, 3.0e7)
    assert gpu.lib().vrt_upload_region_f(ctx.h, w.ctypes.data_as(gpu.P), s, e) == gpu.VRT_ERR_INVALID
    ctx.close()


@pytest.fixture(scope="module")
def frame1080(gpu, orc):
    """The 1080p drop-in world on the GPU, and the oracle's own depth field of
    it (computeDepthField over every voxel on the CPU), which must equal the
    GPU-built one word for word; the frame tests render the oracle from the
    oracle's grid, so nothing the GPU built enters the expected frames."""
    sc = gpu.make_scene("frame1080", seed=1)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    print("1080p world on the GPU; oracle depth field of all %d voxels" % sc.world.size, flush=True)
    odf = orc.depth_field(sc.world, sc.dims)
    gdf = ctx.download()
    bad = np.nonzero(gdf != odf)[0]
    assert bad.size == 0, "GPU depth field differs from the oracle's at %d voxels, first %d" % (bad.size, bad[0])
    del gdf
    yield sc, ctx, odf
    ctx.close()


def test_frame1080_matches_oracle(gpu, orc, frame1080):
    """The 1080p drop-in frame, every pixel, against the oracle rendering its
    own depth field of the raw world."""
    sc, ctx, odf = frame1080
    g = gpu.render_on_gpu(sc, aux=True, ctx=ctx)
    r = orc.render(odf, sc.dims, sc.frame, sc.width, sc.height)
    assert_frames_match(g, r, "frame1080")


def test_frame1080_more_cameras_match_oracle(gpu, orc, frame1080):
    """Six more 1080p cameras over the C2 world (lights clustered around the
    camera so that many are in range, one depth-view frame), cutoff on."""
    sc, ctx, odf = frame1080
    rng = np.random.default_rng(31)
    for k in range(6):
        f = gpu.make_frame(sc.params, sc.dims, sc.world, rng, 1920, 1080, 16, 384, light_spread=48.0,
                           depth_view=1 if k == 5 else 0)
        t = gpu.Target(1920, 1080, depth=True, aux=True)
        ctx.stats(reset=True)
        ctx.draw(t, frame=f)
        g = t.download()
        g.update(ctx.stats(reset=True))
        r = orc.render(odf, sc.dims, f, 1920, 1080)
        assert_frames_match(g, r, "1080p camera %d" % k)


@pytest.mark.parametrize("case", ["frame1080_cameras", "terrain128"])
def test_primary_cutoff_keeps_every_output(gpu, orc, case, request):
    """Without a step map (and outside the depth view) primary rays heading
    up stop once above every solid voxel, like shadow rays: every map the
    target has still equals the oracle's, the iteration total drops below the
    oracle's, and with a step map (or the cutoff off) nothing stops early."""
    if case == "frame1080_cameras":
        sc, ctx, wdf = request.getfixturevalue("frame1080")  # the oracle's own depth field
        own = False
    else:
        sc = gpu.make_scene("terrain128", seed=4)
        ctx = gpu.Context(sc.dims)
        ctx.upload(sc.world)
        ctx.build_depth_field()
        wdf = orc.depth_field(sc.world, sc.dims)
        own = True
    rng = np.random.default_rng(77)
    frames = [sc.frame]
    for _ in range(2):
        # pitched up as well as down, so that many primary rays end in the sky
        f = gpu.make_frame(sc.params, sc.dims, sc.world, rng, sc.width, sc.height, 16, 384,
                           pitch_range=(-0.4, 0.3), light_spread=48.0)
        frames.append(f)
    for k, f in enumerate(frames):
        r = orc.render(wdf, sc.dims, f, sc.width, sc.height)
        res = {}
        for steps in (False, True):
            t = gpu.Target(sc.width, sc.height, depth=True, aux=True, steps=steps)
            ctx.stats(reset=True)
            ctx.draw(t, frame=f)
            g = t.download()
            g.update(ctx.stats(reset=True))
            assert_frames_match(g, r, "%s frame %d steps=%d" % (case, k, steps))
            res[steps] = g["dda_steps"]
        ctx.set_shadow_cutoff(False)
        t = gpu.Target(sc.width, sc.height, depth=True, aux=True, steps=False)
        ctx.stats(reset=True)
        ctx.draw(t, frame=f)
        g = t.download()
        g.update(ctx.stats(reset=True))
        g["shadow_cutoff"] = False
        assert_frames_match(g, r, "%s frame %d cutoff off" % (case, k))
        ctx.set_shadow_cutoff(True)
        assert res[False] <= res[True] <= r["dda_steps"]
        if k > 0:
            assert res[False] < res[True], "no primary ray stopped early in frame %d" % k
    if own:
        ctx.close()


def entering_rays(dims, n_per_face, rng):
    """Rays whose start cell is one cell outside the grid on one axis (-1 or
    the dimension) and whose direction points into the grid on that axis."""
    starts, dirs = [], []
    for axis in range(3):
        for low in (True, False):
            st = np.stack([rng.uniform(0, dims[k], n_per_face) for k in range(3)], 1)
            st[:, axis] = rng.uniform(-1.99, -1.0, n_per_face) if low else rng.uniform(dims[axis], dims[axis] + 0.99,
                                                                                      n_per_face)
            d = rng.normal(size=(n_per_face, 3)) * 0.6
            d[:, axis] = np.abs(rng.normal(1.0, 0.3, n_per_face)) * (1.0 if low else -1.0)
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            starts.append(st)
            dirs.append(d)
    return np.concatenate(starts).astype(np.float32), np.concatenate(dirs).astype(np.float32)


@pytest.mark.parametrize("dims", [(100, 37, 53), (4096, 600, 2), (4100, 2050, 2)],
                         ids=["byte_offset32", "index32", "index64"])
def test_cast_rays_entering_from_one_cell_outside(gpu, orc, dims):
    """A start cell at -1 (or at the dimension) on an axis, stepping into the
    grid: the voxel index of the entered cell must be exact although the start
    cell's coordinates are out of range.  X and X*Y are not multiples of 256,
    so a 24-bit unsigned multiply of -1 would give a wrong index (and a read
    outside the grid).  The sizes take the three index forms: 32-bit byte
    offsets, 32-bit indices (4*X*Y >= 2^23), 64-bit indices."""
    rng = np.random.default_rng(23)
    X, Y, Z = dims
    world = np.where(rng.random(X * Y * Z) < 0.08, rng.integers(0, 1 << 24, X * Y * Z), -1).astype(np.int32)
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    wdf = ctx.download()
    start, d = entering_rays(dims, 1500, rng)
    dist = np.full(start.shape[0], 384, np.int32)
    g = ctx.cast_rays(start, d, dist, 384)
    r = orc.cast_rays(wdf, dims, start, d, dist, 384)
    for k in ("hit", "face", "steps"):
        bad = np.nonzero(g[k] != r[k])[0]
        assert bad.size == 0, "%s differs for %d rays, first %d" % (k, bad.size, bad[0])
    assert same_floats(g["hit_pos"], r["hit_pos"])
    assert int((g["hit"] >= 0).sum()) > 1000
    ctx.close()


def test_cast_rays_near_and_past_the_grid_faces(gpu, orc):
    """Walks that start from 3 cells inside to 5 cells outside a face (either
    side, every axis), pointing in, out, or nearly along it; walks that jump
    out of the grid near a face; and unnormalized directions, whose
    depth-field jumps land up to hundreds of cells away.  Every exit must come
    at the reference's iteration (a cell off the grid ends the walk, one back
    inside continues it)."""
    rng = np.random.default_rng(31)
    dims = (45, 38, 41)
    X, Y, Z = dims
    world = np.where(rng.random(X * Y * Z) < 0.03, rng.integers(0, 1 << 24, X * Y * Z), -1).astype(np.int32)
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    wdf = ctx.download()
    starts, dirs = [], []
    n = 400
    for axis in range(3):
        for low in (True, False):
            for off in (-3.5, -0.5, 0.5, 1.5, 2.5, 5.5):
                st = np.stack([rng.uniform(-0.5, dims[k] + 0.5, n) for k in range(3)], 1)
                st[:, axis] = -off if low else dims[axis] - 1 + off
                d = rng.normal(size=(n, 3))
                # a third pointing in, a third out, a third along the face
                sgn = (1.0 if low else -1.0) * np.where(np.arange(n) % 3 == 0, 1.0,
                                                          np.where(np.arange(n) % 3 == 1, -1.0, 0.05))
                d[:, axis] = np.abs(d[:, axis]) * sgn
                d /= np.linalg.norm(d, axis=1, keepdims=True)
                d *= np.where(np.arange(n) % 5 == 0, 40.0, 1.0)[:, None]  # some unnormalized
                starts.append(st)
                dirs.append(d)
    start = np.concatenate(starts).astype(np.float32)
    d = np.concatenate(dirs).astype(np.float32)
    dist = rng.integers(1, 200, start.shape[0]).astype(np.int32)
    for budget in (384, 1 << 24):
        g = ctx.cast_rays(start, d, dist if budget == 384 else np.full_like(dist, budget), budget)
        r = orc.cast_rays(wdf, dims, start, d, dist if budget == 384 else np.full_like(dist, budget), budget)
        for k in ("hit", "face", "steps"):
            bad = np.nonzero(g[k] != r[k])[0]
            assert bad.size == 0, "%s differs for %d rays, first %d" % (k, bad.size, bad[0])
        assert same_floats(g["hit_pos"], r["hit_pos"])
        assert int((g["hit"] >= 0).sum()) > 1000
    ctx.close()


def test_cast_rays_extreme_values_match_oracle(gpu, orc):
    """int(float) at NaN, +-inf and beyond int32 range, zero and denormal
    direction components, huge budgets: the conversions saturate and NaN goes
    to 0 on both sides."""
    sc = gpu.make_scene("tiny", seed=8)
    wdf = orc.depth_field(sc.world, sc.dims)
    nan, inf = float("nan"), float("inf")
    starts = [[nan, 40.0, 40.0], [40.0, nan, 40.0], [inf, 40.0, 40.0], [-inf, 40.0, 40.0], [3e9, 40.0, 40.0],
              [-3e9, 40.0, 40.0], [2147483520.0, 40.0, 40.0], [40.5, 45.5, 40.5], [40.5, 45.5, 40.5],
              [40.5, 45.5, 40.5], [40.5, 45.5, 40.5], [1e-40, 50.0, 1e-40], [95.99999, 63.99999, 95.99999]]
    dirs = [[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [-1.0, 0.0, 0.0],
            [1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [nan, 0.5, 0.5], [inf, -1.0, 0.0], [1e-45, -1.0, 1e-45],
            [0.0, 0.0, 0.0], [0.7071, -0.7071, 0.0], [-0.57735, -0.57735, -0.57735]]
    n = len(starts)
    ctx = gpu.Context(sc.dims)
    ctx.upload(wdf)
    for budget in (384, 1 << 24):
        dist = np.full(n, budget, np.int32)
        g = ctx.cast_rays(starts, dirs, dist, budget)
        r = orc.cast_rays(wdf, sc.dims, starts, dirs, dist, budget)
        for k in ("hit", "face", "steps"):
            assert np.array_equal(g[k], r[k]), (budget, k, g[k], r[k])
        assert same_floats(g["hit_pos"], r["hit_pos"])
    ctx.close()


def test_cast_rays_fuzz_small_and_odd_grids(gpu, orc):
    """castRay on 24 grids of random size (1 to 37 cells per axis, so the
    guard cells and the clamp meet one- and two-cell extents), random solid
    density, depth field built; rays from inside, near and far outside, with
    zero, tiny, unnormalized and axis-aligned directions and random budgets."""
    rng = np.random.default_rng(4321)
    total_hits = 0
    for case in range(24):
        dims = tuple(int(v) for v in rng.integers(1, 38, 3))
        X, Y, Z = dims
        n = X * Y * Z
        world = np.where(rng.random(n) < rng.uniform(0.0, 0.4), rng.integers(0, 1 << 24, n), -1).astype(np.int32)
        ctx = gpu.Context(dims)
        ctx.upload(world)
        ctx.build_depth_field()
        wdf = ctx.download()
        m = 3000
        start = np.stack([rng.uniform(-6.0, dims[k] + 6.0, m) for k in range(3)], 1)
        d = rng.normal(size=(m, 3))
        kind = rng.integers(0, 6, m)
        d[kind == 1] *= 60.0                                       # unnormalized, long jumps
        d[kind == 2, rng.integers(0, 3)] = 0.0                     # a zero component
        d[kind == 3] = np.eye(3)[rng.integers(0, 3, int((kind == 3).sum()))] * rng.choice([-1, 1], (int((kind == 3).sum()), 1))
        d[kind == 4, rng.integers(0, 3)] = 1e-7                    # a tiny component
        start[kind == 5] = np.floor(start[kind == 5])              # integer starts (cell faces)
        start, d = start.astype(np.float32), d.astype(np.float32)
        dist = rng.integers(0, 300, m).astype(np.int32)
        for steps in (True, False):
            g = ctx.cast_rays(start, d, dist, 256, steps=steps)
            r = orc.cast_rays(wdf, dims, start, d, dist, 256)
            keys = ("hit", "face", "steps") if steps else ("hit", "face")
            for k in keys:
                bad = np.nonzero(g[k] != r[k])[0]
                assert bad.size == 0, "grid %s: %s differs for %d rays, first %d (start %s dir %s)" % (
                    dims, k, bad.size, bad[0], start[bad[0]].tolist(), d[bad[0]].tolist())
            assert same_floats(g["hit_pos"], r["hit_pos"]), "grid %s hit_pos" % (dims,)
        total_hits += int((r["hit"] >= 0).sum())
        ctx.close()
    assert total_hits > 5000, total_hits


def test_random_views_sweep(gpu, orc):
    """Random cameras over two worlds: low and high above the terrain (below
    and above the grid's solid top), pitched up and down, with and without a
    step map, in odd frame sizes.  Plain draws deal their tile rows in a
    camera-dependent order and walk the padded render copy; every map must
    equal the oracle's."""
    rng = np.random.default_rng(2024)
    hits = []
    for name, seed in (("tiny", 9), ("terrain128", 6)):
        sc = gpu.make_scene(name, seed=seed)
        X, Y, Z = sc.dims
        ctx = gpu.Context(sc.dims)
        ctx.upload(sc.world)
        ctx.build_depth_field()
        wdf = ctx.download()
        top = int(np.nonzero((sc.world.reshape(Z, Y, X) >= 0).any(axis=(0, 2)))[0].max())
        for k in range(8):
            w, h = int(rng.integers(40, 200)), int(rng.integers(30, 150))
            high = k % 2 == 1
            ch = (top - sc.params.height_min + 2, top - sc.params.height_min + 12) if high else (2, 10)
            # cameras above the top look down (positive pitch), the others either way
            f = gpu.make_frame(sc.params, sc.dims, sc.world, rng, w, h, int(rng.integers(0, 17)), 384,
                               pitch_range=(0.15, 0.6) if high else (-0.6, 0.6), cam_height=ch, light_spread=40.0)
            r = orc.render(wdf, sc.dims, f, w, h)
            for steps in (False, True):
                t = gpu.Target(w, h, depth=True, aux=True, steps=steps)
                ctx.stats(reset=True)
                ctx.draw(t, frame=f)
                g = t.download()
                g.update(ctx.stats(reset=True))
                assert_frames_match(g, r, "%s view %d (%dx%d, cam y %.1f, top %d) steps=%d" % (
                    name, k, w, h, f.cam_pos[1], top, steps))
            hits.append(int((r["hit"] >= 0).sum()))
        ctx.close()
    assert sum(1 for v in hits if v > 100) >= 12, hits


def test_every_writer_keeps_the_render_copy_current(gpu, orc):
    """The render kernels read a padded copy of the grid that each entry point
    writing the grid refreshes.  After every kind of write, a frame rendered
    on the GPU must equal the oracle's frame of the grid as downloaded."""
    sc = gpu.make_scene("tiny", seed=5)
    X, Y, Z = sc.dims
    ctx = gpu.Context(sc.dims)
    rng = np.random.default_rng(12)
    hit_px = []

    def check(label):
        w = ctx.download()
        g = gpu.render_on_gpu(sc, aux=True, ctx=ctx)
        r = orc.render(w, sc.dims, sc.frame, sc.width, sc.height)
        assert_frames_match(g, r, label)
        hit_px.append(int((r["hit"] >= 0).sum()))
        return r

    check("fresh context (all -1)")
    ctx.generate(sc.params)
    r = check("device generator")
    ctx.build_depth_field()
    check("depth field")
    # cells the camera sees: carve and place around a few hit voxels
    hits = r["hit"][r["hit"] >= 0]
    picks = rng.choice(hits, 6, replace=False)
    cells = [(int(i % X), int((i // X) % Y), int(i // (X * Y))) for i in picks]
    ctx.remove_sphere(*cells[0], 5)
    check("removeSphere")
    ctx.place_voxel(cells[1][0], cells[1][1] + 1, cells[1][2], 0x30A050)
    check("placeVoxel")
    ctx.destroy_voxel(*cells[2])
    check("destroyVoxel")
    lo = (max(cells[3][0] - 4, 0), max(cells[3][1] - 4, 0), max(cells[3][2] - 4, 0))
    ctx.fix_depth_field_box(lo, (lo[0] + 9, lo[1] + 9, lo[2] + 9))
    check("fixDepthField box")
    w = ctx.download().reshape(Z, Y, X)
    x, y, z = cells[4]
    w[max(z - 3, 0):z + 3, y + 1:y + 4, max(x - 3, 0):x + 3] = 0xC08040
    # end corner inside the grid, or the reference's corner rule would swap the corners
    ctx.upload_region(w.reshape(-1), (max(x - 3, 0), y + 1, max(z - 3, 0)),
                      (min(x + 2, X - 1), min(y + 4, Y - 1), min(z + 3, Z - 1)))
    check("region upload")
    w = ctx.download()
    w[rng.random(w.size) < 0.002] = 0x102030
    ctx.upload(w)
    check("full upload")
    other = gpu.DeviceBuffer(w.nbytes)
    w2 = w.copy()
    w2[rng.random(w.size) < 0.002] = -1
    other.upload(w2)
    gpu.call("vrt_upload_device", ctx.h, other.addr)
    check("device-to-device upload")
    p = gpu.P()
    gpu.call("vrt_grid_device_ptr", ctx.h, gpu.C.byref(p))
    w3 = w2.copy()
    w3[rng.random(w.size) < 0.002] = 0x405060
    buf = gpu.DeviceBuffer(w3.nbytes)
    buf.upload(w3)
    gpu.hip_check(gpu.hip().hipMemcpy(p, buf.ptr, buf.nbytes, gpu.D2D), "hipMemcpy D2D")
    gpu.hip_check(gpu.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
    ctx.grid_modified()
    check("device pointer + vrt_grid_modified")
    ctx.close()
    assert hit_px[0] == 0 and min(hit_px[1:]) > 1000, hit_px


def test_device_copy_and_generator(gpu):
    """vrt_upload_device copies a whole grid on the device; the device world
    generator writes the same words as the host generator."""
    for name, seed in (("tiny", 3), ("terrain128", 2)):
        cfg = gpu.CONFIGS[name]
        hmin, hmax = cfg.get("heights", (16, 96))
        p = gpu.world_params(cfg["dims"], seed, caves=cfg["caves"], boxes=cfg["boxes"] + 50, height_min=hmin,
                             height_max=hmax)
        host = gpu.generate_world(p)
        ctx = gpu.Context(cfg["dims"])
        ctx.generate(p)
        dev = ctx.download()
        assert np.array_equal(dev, host), "%s: device generator differs at %d voxels" % (name, int((dev != host).sum()))
        other = gpu.DeviceBuffer(host.nbytes)
        other.upload(host[::-1].copy())
        gpu.call("vrt_upload_device", ctx.h, other.addr)
        assert np.array_equal(ctx.download(), host[::-1])
        ctx.close()


def degenerate_cases(gpu):
    rng = np.random.default_rng(12)
    cases = []
    for name, dims, fill in [("empty", (40, 30, 50), "empty"), ("solid", (40, 30, 50), "solid"),
                             ("single", (40, 30, 50), "single"), ("one-cell", (1, 1, 1), "solid"),
                             ("column", (1, 64, 1), "random"), ("plate", (64, 1, 64), "random"),
                             ("sparse", (33, 17, 29), "random")]:
        n = dims[0] * dims[1] * dims[2]
        if fill == "empty":
            w = np.full(n, -1, np.int32)
        elif fill == "solid":
            w = rng.integers(0, 1 << 24, n).astype(np.int32)
        elif fill == "single":
            w = np.full(n, -1, np.int32)
            w[20 + 40 * 10 + 40 * 30 * 25] = 0x00FF00
        else:
            w = np.where(rng.random(n) < 0.1, rng.integers(0, 1 << 24, n), -1).astype(np.int32)
        cases.append((name, dims, w))
    return cases


@pytest.mark.parametrize("idx", range(7))
def test_degenerate_worlds_and_ragged_frames(gpu, orc, idx):
    name, dims, w = degenerate_cases(gpu)[idx]
    ctx = gpu.Context(dims)
    ctx.upload(w)
    ctx.build_depth_field()
    wdf = ctx.download()
    assert np.array_equal(wdf, orc.depth_field(w, dims)), name + " depth field"
    rng = np.random.default_rng(idx)
    for (W, H, rd) in [(13, 7, 384), (1, 1, 384), (64, 40, 0), (50, 50, 5)]:
        f = gpu.frame_defaults(dims[0], W, H)
        f.cam_pos[:] = [dims[0] * 0.5 + 0.25, dims[1] * 0.5 + 0.1, dims[2] * 0.5 - 0.3]
        f.rotate_matrix[:] = gpu.camera_matrix(rng.uniform(-1, 1), rng.uniform(0, 6.28)).tolist()
        f.render_dist = rd
        f.lights[0].pos[:] = [dims[0] * 0.5, dims[1] * 0.5, dims[2] * 0.5]
        f.lights[0].rgb[:] = [1.0, 0.5, 2.0]
        t = gpu.Target(W, H, depth=True, aux=True)
        ctx.stats(reset=True)
        ctx.draw(t, frame=f)
        g = t.download()
        g.update(ctx.stats(reset=True))
        r = orc.render(wdf, dims, f, W, H)
        assert_frames_match(g, r, "%s %dx%d rd=%d" % (name, W, H, rd))
    ctx.close()


def test_cast_rays_cutoff_without_steps(gpu, orc):
    """vrt_cast_rays without a steps array lets rays heading up stop above
    every solid voxel: hit, position and face equal the run with steps (which
    walks them to the end) and the oracle's."""
    sc = gpu.make_scene("tiny", seed=3)
    ctx = gpu.Context(sc.dims)
    ctx.upload(sc.world)
    ctx.build_depth_field()
    wdf = ctx.download()
    X, Y, Z = sc.dims
    rng = np.random.default_rng(5)
    n = 20000
    start = np.stack([rng.uniform(0, X, n), rng.uniform(Y * 0.3, Y, n), rng.uniform(0, Z, n)], 1).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 1] = np.abs(d[:, 1]) * np.where(rng.random(n) < 0.7, 1.0, -1.0)  # mostly upward
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dist = np.full(n, 384, np.int32)
    r = orc.cast_rays(wdf, sc.dims, start, d, dist, 384)
    full = ctx.cast_rays(start, d, dist, 384)
    cut = ctx.cast_rays(start, d, dist, 384, steps=False)
    for k in ("hit", "face"):
        assert np.array_equal(full[k], r[k]) and np.array_equal(cut[k], r[k]), k
    assert same_floats(cut["hit_pos"], r["hit_pos"]) and np.array_equal(full["steps"], r["steps"])
    assert int((r["hit"] == -1).sum()) > n // 4
    ctx.close()


def test_pin_rays_match_oracle(gpu, orc):
    """The closed-form rays that pin the oracle (start-cell truncation, exactly
    zero direction components, the 1e-6 step bias) through the kernels."""
    import test_oracle_pins as pins
    for make in (pins.truncation_rays, pins.zero_component_rays, pins.step_bias_rays):
        dims, w, start, d = make()
        start = np.asarray(start, np.float32)
        d = np.asarray(d, np.float32)
        dist = np.full(start.shape[0], 384, np.int32)
        ctx = gpu.Context(dims)
        ctx.upload(w)
        g = ctx.cast_rays(start, d, dist, 384)
        r = orc.cast_rays(w, dims, start, d, dist, 384)
        for k in ("hit", "face", "steps"):
            assert np.array_equal(g[k], r[k]), "%s: %s gpu %s oracle %s" % (make.__name__, k, g[k], r[k])
        assert same_floats(g["hit_pos"], r["hit_pos"]), make.__name__
        ctx.close()


@pytest.mark.parametrize("case", ["range", "axis_range", "overbright", "falloff27", "falloff28", "roof", "clamp0",
                                  "clamp15", "clamp14_15"])
def test_pin_frames_match_oracle(gpu, orc, case):
    """The exact local-light scenes that pin the oracle (light at exactly its
    radius, overbright reached at exactly 1.25, linear falloff, shadow rays
    that stop at the light, the clamp after the first, the last and the
    next-to-last light slot) through the kernels, cutoff off and on."""
    import test_oracle_pins as pins
    if case == "roof":
        dims, w = pins.roof_world()
        f = pins.roof_frame(gpu, dims)
    else:
        dims, w = pins.shading_world()
        f = {"range": lambda: pins.light_range_frame(gpu, dims),
             "axis_range": lambda: pins.axis_range_frame(gpu, dims),
             "overbright": lambda: pins.overbright_frame(gpu, dims),
             "falloff27": lambda: pins.falloff_frame(gpu, dims, float(2 ** 27)),
             "falloff28": lambda: pins.falloff_frame(gpu, dims, float(2 ** 28)),
             "clamp0": lambda: pins.clamp_frame(gpu, dims, (0,)),
             "clamp15": lambda: pins.clamp_frame(gpu, dims, (15,)),
             "clamp14_15": lambda: pins.clamp_frame(gpu, dims, (14, 15))}[case]()
    ctx = gpu.Context(dims)
    ctx.upload(w)
    ctx.build_depth_field()
    wdf = ctx.download()
    r = orc.render(wdf, dims, f, 8, 8)
    for cutoff in (False, True):
        ctx.set_shadow_cutoff(cutoff)
        t = gpu.Target(8, 8, depth=True, aux=True)
        ctx.stats(reset=True)
        ctx.draw(t, frame=f)
        g = t.download()
        g.update(ctx.stats(reset=True))
        g["shadow_cutoff"] = cutoff
        assert_frames_match(g, r, "%s cutoff=%d" % (case, cutoff))
        assert np.array_equal(g["rgb"], r["rgb"]), case  # exact lighting terms: no rounding slack needed
    ctx.close()
