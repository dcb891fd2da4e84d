"""The oracle's on-demand world (oracle_ctypes.LazyWorld: depth-field words
computed from the raw world and the edit log only where a walk reads them)
against the oracle's eager path it stands for: computeDepthField over the
whole grid, then removeSphere for each edit in order.  Every word of the grid
at every version, and every output of whole frames, must be identical.  The
large-world GPU tests rely on this mode, so it is pinned here on grids the
eager path finishes in seconds."""
import numpy as np
import pytest


def random_world(dims, density, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = dims
    w = np.full(X * Y * Z, -1, np.int32)
    # a floor with bumps, then scattered solids: near, mid and far cells all occur
    h = rng.integers(2, max(3, Y // 3), size=(Z, X))
    ys = np.arange(Y)[None, :, None]
    solid = ys < h[:, None, :]
    solid |= rng.random((Z, Y, X)) < density
    col = rng.integers(0, 1 << 24, size=solid.shape, dtype=np.int64).astype(np.int32)
    w = np.where(solid, col, -1).astype(np.int32).reshape(-1)
    return w


def frame_for(vrt, dims, seed, w=64, h=48, lights=6):
    rng = np.random.default_rng(seed)
    X, Y, Z = dims
    f = vrt.frame_defaults(X, w, h)
    f.cam_pos[:] = [rng.uniform(0.2, 0.8) * X, rng.uniform(0.55, 0.9) * Y, rng.uniform(0.2, 0.8) * Z]
    f.rotate_matrix[:] = vrt.camera_matrix(rng.uniform(0.2, 0.7), rng.uniform(-3.1, 3.1)).tolist()
    f.light_pos[:] = vrt.light_orbit(rng.uniform(-60, -30), (X / 2.0, 3.0 * X, Z / 2.0)).tolist()
    f.focal = vrt.FOV70_FOCAL
    for i in range(lights):
        L = f.lights[i]
        L.pos[:] = [rng.uniform(0, X - 1), rng.uniform(0.4, 0.95) * Y, rng.uniform(0, Z - 1)]
        L.rgb[:] = [rng.uniform(0.5, 2.0) for _ in range(3)]
        L.radius = rng.uniform(8.0, 40.0)
    return f


def same(a, b):
    for k in ("rgb", "hit", "face", "shadow_eval", "shadow_occ", "steps"):
        if not np.array_equal(a[k], b[k]):
            return k
    if not np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32)):
        return "depth"
    if (a["rays"], a["dda_steps"]) != (b["rays"], b["dda_steps"]):
        return "totals"
    return None


EDITS = [(10, 5, 12, 4), (0, 0, 0, 6), (36, 28, 40, 8), (18, 3, 20, 1), (18, 4, 21, 7), (50, 10, 10, 5),
         (20, 6, 19, 0), (22, 6, 22, -2), (-3, 6, 15, 5), (12, 2, 30, 8), (12, 3, 31, 3)]


@pytest.mark.parametrize("seed,density", [(1, 0.01), (2, 0.05), (3, 0.0)])
def test_lazy_world_equals_eager_build_and_edits(vrt, orc, seed, density):
    dims = (37, 29, 41)
    raw = random_world(dims, density, seed)
    lazy = orc.LazyWorld(raw, dims, EDITS)
    eager = orc.depth_field(raw, dims)
    idx = np.arange(raw.size)
    f = frame_for(vrt, dims, seed)
    for t in range(len(EDITS) + 1):
        if t > 0:
            orc.remove_sphere(eager, dims, EDITS[t - 1][:3], EDITS[t - 1][3])
        words = lazy.words(idx, version=t)
        bad = np.nonzero(words != eager)[0]
        assert bad.size == 0, "version %d: %d words differ, first cell %d: %x vs %x" % (
            t, bad.size, bad[0], words[bad[0]], eager[bad[0]])
        if t in (0, 3, len(EDITS)):
            r_e = orc.render(eager, dims, f, 64, 48)
            r_l = lazy.render(f, 64, 48, version=t)
            assert same(r_l, r_e) is None, "version %d frame differs in %s" % (t, same(r_l, r_e))


def test_lazy_world_on_a_generated_world_and_pixel_lists(vrt, orc):
    """A generated terrain (caves, boxes) with the depth view and a shaded
    frame; the pixel-list form returns the same outputs as the whole frame."""
    sc = vrt.make_scene("tiny", seed=7)
    eager = orc.depth_field(sc.world, sc.dims)
    lazy = orc.LazyWorld(sc.world, sc.dims)
    for mode in (0, 1):
        sc.frame.view_depth_field = mode
        r_e = orc.render(eager, sc.dims, sc.frame, sc.width, sc.height)
        r_l = lazy.render(sc.frame, sc.width, sc.height)
        assert same(r_l, r_e) is None
        pix = np.random.default_rng(mode).integers(0, sc.width * sc.height, 500)
        r_p = lazy.render(sc.frame, sc.width, sc.height, pixels=pix)
        for k in ("rgb", "hit", "face", "shadow_eval", "shadow_occ", "steps"):
            assert np.array_equal(r_p[k], r_e[k].reshape((sc.width * sc.height,) + r_e[k].shape[2:])[pix]), k
    assert (lazy.memo != 0).any()  # the memo filled in on the way
