"""Pins of the CPU oracle (CPU only).

The reference ships no tests or golden vectors, and its GLSL cannot run here,
so each oracle function is pinned by something that does not restate it:
closed forms, an exact-arithmetic voxel traversal (Amanatides & Woo 1987) on
tiny grids, the reference's own vendored GLM headers (oracle/_ref/glm_probe,
built where the reference is present), lattice-point counts, cube-distance
geometry and order independence."""
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def bits_f32(u):
    return struct.unpack("<f", struct.pack("<I", u & 0xFFFFFFFF))[0]


# ------------------------------------------------------------ depth table
def test_depth_table_counts(orc):
    dist, off = orc.depth_table()
    # lattice points with x^2+y^2+z^2 <= 49 (OEIS A000605: 1, 7, 33, 123, 257, 515, 925, 1419)
    assert len(dist) == 1419
    # computeDepthIndices walks z, then y, then x from -7: first entry (0,0,-7)
    assert off[0].tolist() == [0, 0, -7] and off[-1].tolist() == [0, 0, 7]
    # distance 0 exactly for the 27 offsets whose cube touches the centre cube
    assert int((dist == 0).sum()) == 27
    assert set(map(tuple, off[dist == 0].tolist())) == {(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1)
                                                         for c in (-1, 0, 1)}
    # the farthest entries are the six axis points at 7, cube distance 6
    assert int((dist == -6.0).sum()) == 6
    assert dist.min() == np.float32(-6.0) and dist.max() == 0.0


def test_depth_table_is_cube_distance(orc):
    """Each entry's distance is minus the distance between the unit cube at the
    origin and the unit cube at the offset, rounded to float."""
    dist, off = orc.depth_table()
    for d, (x, y, z) in zip(dist, off):
        gap = [max(abs(c) - 1, 0) for c in (x, y, z)]
        assert d == np.float32(-math.sqrt(sum(g * g for g in gap)))


def test_product_table_equals_oracle(orc, vrt):
    d1, o1 = orc.depth_table()
    d2, o2 = vrt.depth_table()
    assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32)) and np.array_equal(o1, o2)


# ---------------------------------------------------------------- GLM probe
def _probe_available(orc):
    return os.path.exists(orc.GLM_PROBE)


def _hexs(vals):
    return " ".join("%08x" % f32_bits(float(v)) for v in vals)


@pytest.fixture(scope="module")
def probe(orc):
    if not _probe_available(orc):
        pytest.skip("oracle/_ref/glm_probe is built only where the reference sources are present")
    return orc.glm_probe


def _rand_vecs(rng, n):
    v = rng.normal(size=(n, 3)).astype(np.float32) * rng.choice([1e-3, 1.0, 50.0, 1e4], size=(n, 1)).astype(np.float32)
    return v


def test_oracle_vector_maths_equals_glm(orc, probe):
    rng = np.random.default_rng(1)
    vs = _rand_vecs(rng, 300)
    ws = _rand_vecs(rng, 300)
    lines = []
    for v, w in zip(vs, ws):
        lines += ["normalize " + _hexs(v), "length " + _hexs(v), "dot " + _hexs(v) + " " + _hexs(w)]
    out = probe(lines)
    for i, (v, w) in enumerate(zip(vs, ws)):
        assert [f32_bits(x) for x in orc.normalize(v)] == out[3 * i]
        assert f32_bits(orc.length(v)) == out[3 * i + 1][0]
        assert f32_bits(orc.dot(v, w)) == out[3 * i + 2][0]


def test_oracle_primary_ray_equals_glm(orc, vrt, probe):
    rng = np.random.default_rng(2)
    lines, cases = [], []
    for _ in range(300):
        f = vrt.frame_defaults(512, 1920, 1080)
        f.aspect_ratio = float(rng.choice([1.0, 16 / 9, 4 / 3]))
        f.focal = float(rng.choice([1.0, vrt.FOV70_FOCAL]))
        m = vrt.camera_matrix(rng.uniform(-1.5, 1.5), rng.uniform(-6.3, 6.3))
        f.rotate_matrix[:] = m.tolist()
        vx, vy = np.float32(rng.uniform(-1, 1)), np.float32(rng.uniform(-1, 1))
        cases.append((f, vx, vy))
        lines.append("primary " + _hexs([f.aspect_ratio, f.focal, vx, vy]) + " " + _hexs(m))
    out = probe(lines)
    for (f, vx, vy), o in zip(cases, out):
        assert [f32_bits(x) for x in orc.primary_dir(f, float(vx), float(vy))] == o


def test_host_matrix_helpers_equal_glm(vrt, probe):
    rng = np.random.default_rng(3)
    lines, expect = [], []
    for _ in range(200):
        m = rng.normal(size=16).astype(np.float32)
        a = np.float32(rng.uniform(-7, 7))
        ax = rng.normal(size=3).astype(np.float32)
        lines.append("rotate " + _hexs(m) + " " + _hexs([a]) + " " + _hexs(ax))
        expect.append(vrt.mat4_rotate(m, float(a), ax))
        b = rng.normal(size=16).astype(np.float32)
        lines.append("matmul " + _hexs(m) + " " + _hexs(b))
        expect.append(vrt.mat4_mul(m, b))
        p, y = np.float32(rng.uniform(-1.6, 1.6)), np.float32(rng.uniform(-6.3, 6.3))
        lines.append("camera " + _hexs([p, y]))
        expect.append(vrt.camera_matrix(float(p), float(y)))
        deg = np.float32(rng.uniform(-400, 400))
        start = rng.uniform(0, 3000, 3).astype(np.float32)
        lines.append("orbit " + _hexs([deg]) + " " + _hexs(start))
        expect.append(vrt.light_orbit(float(deg), start))
    out = probe(lines)
    for e, o in zip(expect, out):
        assert [f32_bits(float(x)) for x in e] == o


# -------------------------------------------------------- traversal closed forms
def empty_grid(dims):
    return np.full(dims[0] * dims[1] * dims[2], -1, np.int32)


def put(w, dims, x, y, z, v):
    w[x + dims[0] * y + dims[0] * dims[1] * z] = v


def idx(dims, x, y, z):
    return x + dims[0] * y + dims[0] * dims[1] * z


def test_empty_grid_walks_to_the_edge(orc):
    dims = (16, 8, 12)
    w = empty_grid(dims)
    e = 1e-4
    d = np.array([1.0, e, e], np.float32) / np.float32(np.linalg.norm([1, e, e]))
    r = orc.cast_rays(w, dims, [[0.5, 0.5, 0.5]], [d], [384], 384)
    assert r["hit"][0] == -1 and r["steps"][0] == 16  # cells x=1..15 in the grid, the 16th step leaves
    r = orc.cast_rays(w, dims, [[0.5, 0.5, 0.5]], [d], [5], 384)
    assert r["steps"][0] == 5  # the budget stops it first
    r = orc.cast_rays(w, dims, [[0.5, 0.5, 0.5]], [d], [384], 3)
    assert r["steps"][0] == 3  # so does the step cap


@pytest.mark.parametrize("axis,sign", [(a, s) for a in range(3) for s in (1, -1)])
def test_single_voxel_hit_from_each_face(orc, axis, sign):
    dims = (12, 12, 12)
    w = empty_grid(dims)
    target = [6, 6, 6]
    put(w, dims, *target, 0x102030)
    start = [6.5, 6.5, 6.5]
    start[axis] = 6.5 - sign * 4.0  # four cells before the target
    e = [2e-4, 3e-4, 1e-4]
    d = np.array([e[0], e[1], e[2]], np.float64)
    d[axis] = sign
    d = (d / np.linalg.norm(d)).astype(np.float32)
    r = orc.cast_rays(w, dims, [start], [d], [384], 384)
    assert r["hit"][0] == idx(dims, *target)
    assert r["steps"][0] == 4
    assert r["face"][0] == axis * 3 + (sign + 1)  # step +1 on the axis, normal -1
    # the hit point lies on the entered face (within the 1e-6 bias of the step sizes)
    face_coord = 6.0 if sign > 0 else 7.0
    assert abs(r["hit_pos"][0][axis] - face_coord) < 1e-4


def test_tie_break_goes_to_z(orc):
    """intersect.x == intersect.y < intersect.z takes the z branch (fshader.glsl:87-104)."""
    dims = (8, 8, 8)
    w = empty_grid(dims)
    put(w, dims, 0, 0, 1, 0x0000FF)  # the z neighbour
    d = np.array([1.0, 1.0, 0.25], np.float32)
    d = d / np.float32(np.sqrt(np.float32(d @ d)))
    r = orc.cast_rays(w, dims, [[0.5, 0.5, 0.5]], [d], [384], 384)
    assert r["hit"][0] == idx(dims, 0, 0, 1) and r["face"][0] == 2 * 3 + 2 and r["steps"][0] == 1


def test_start_cell_is_never_tested(orc):
    dims = (8, 8, 8)
    w = empty_grid(dims)
    put(w, dims, 3, 3, 3, 0x111111)  # the start cell is solid
    put(w, dims, 5, 3, 3, 0x222222)
    d = np.array([1.0, 1e-4, 2e-4], np.float32)
    r = orc.cast_rays(w, dims, [[3.5, 3.5, 3.5]], [d], [384], 384)
    assert r["hit"][0] == idx(dims, 5, 3, 3) and r["steps"][0] == 2


def exact_first_solid(w, dims, start, d, max_steps):
    """Amanatides & Woo in exact rational arithmetic: the sequence of cells a
    ray from `start` along `d` enters; returns (index, step count) of the
    first solid one, or (-1, n) after leaving the grid / max_steps.  Returns
    None when two axis crossings are too close to order robustly (the
    shader's 1e-6 step bias may swap them)."""
    s = [Fraction(float(v)) for v in start]
    dd = [Fraction(float(v)) for v in d]
    cell = [int(math.floor(v)) for v in s]
    step = [1 if v > 0 else -1 for v in dd]
    nxt = [(Fraction(cell[a] + (1 if step[a] > 0 else 0)) - s[a]) / dd[a] for a in range(3)]
    inc = [Fraction(1) / abs(dd[a]) for a in range(3)]
    for n in range(1, max_steps + 1):
        order = sorted(range(3), key=lambda a: nxt[a])
        if nxt[order[1]] - nxt[order[0]] < Fraction(1, 1000):
            return None
        a = order[0]
        cell[a] += step[a]
        nxt[a] += inc[a]
        if not all(0 <= cell[k] < dims[k] for k in range(3)):
            return -1, n
        v = w[idx(dims, *cell)]
        if v >= 0:
            return idx(dims, *cell), n
    return -1, max_steps


def test_traversal_equals_exact_dda_on_tiny_grids(orc):
    rng = np.random.default_rng(7)
    checked = 0
    for trial in range(60):
        dims = tuple(int(v) for v in rng.integers(4, 12, 3))
        w = np.where(rng.random(dims[0] * dims[1] * dims[2]) < 0.08, rng.integers(0, 1 << 24), -1).astype(np.int32)
        n = 200
        start = rng.uniform(0.05, 0.95, (n, 3)) * np.array(dims)
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        start = start.astype(np.float32)
        d = d.astype(np.float32)
        r = orc.cast_rays(w, dims, start, d, np.full(n, 384, np.int32), 384)
        for i in range(n):
            if min(abs(d[i])) < 1e-3:
                continue
            e = exact_first_solid(w, dims, start[i], d[i], 384)
            if e is None:
                continue
            assert (r["hit"][i], r["steps"][i]) == e, (trial, i)
            checked += 1
    assert checked > 8000


def test_depth_field_jump_skips_empty_space(orc):
    """In an empty grid with its depth field every cell holds -6: each loop
    iteration moves one cell and jumps six, so a ray along +x crosses a
    64-wide grid in about 64/7 iterations instead of 64."""
    dims = (64, 4, 4)
    w = orc.depth_field(empty_grid(dims), dims)
    assert np.all(w == np.int32(f32_bits(-6.0) - (1 << 32)))
    d = np.array([1.0, 1e-5, 1e-5], np.float32)
    r = orc.cast_rays(w, dims, [[0.5, 1.5, 1.5]], [d], [384], 384)
    assert r["hit"][0] == -1 and 9 <= r["steps"][0] <= 10


def truncation_rays():
    """Rays that start outside the grid, in (-1, 0) and (-2, -1) on x.  GLSL's
    int(float) drops the fraction (GLSL 4.60 5.4.1), so x = -0.5 lies in cell
    0 and x = -1.5 in cell -1; the start cell is never tested."""
    dims = (8, 4, 4)
    w = empty_grid(dims)
    put(w, dims, 0, 1, 1, 0x111111)
    put(w, dims, 1, 1, 1, 0x222222)
    d = np.array([1.0, 1e-4, 2e-4], np.float32)
    return dims, w, [[-0.5, 1.5, 1.5], [-1.5, 1.5, 1.5]], [d, d]


def test_start_cell_truncates_toward_zero(orc):
    dims, w, start, d = truncation_rays()
    r = orc.cast_rays(w, dims, start, d, [384, 384], 384)
    assert r["hit"][0] == idx(dims, 1, 1, 1) and r["steps"][0] == 1  # cell 0 is the (untested) start
    assert r["hit"][1] == idx(dims, 0, 1, 1) and r["steps"][1] == 1  # cell -1 -> 0


def zero_component_rays():
    """Rays whose direction has an exactly zero component.  sign(0) = 0
    (GLSL 4.60 8.3), so that axis never moves and its first boundary
    distance is (cell - start) / dir with no forward step:
      * dir.y = +0 and a fractional start: -inf, the smallest, so y is the
        axis taken on every iteration and the ray never leaves its cell;
      * dir.y = -0: +inf, so y is never taken and the ray walks in x and z;
      * dir.y = +0 and start.y integral: 0/0 = NaN, every comparison with it
        is false, so z (the last branch) is taken on every iteration."""
    dims = (8, 4, 8)
    w = empty_grid(dims)
    for x in range(dims[0]):
        for y in range(dims[1]):
            put(w, dims, x, y, 6, 0x336699)  # a wall across z = 6
    d_pos = np.array([0.6, 0.0, 0.8], np.float32)
    d_neg = np.array([0.6, -0.0, 0.8], np.float32)
    start = [[0.5, 1.5, 0.5], [0.5, 1.5, 0.5], [0.5, 1.0, 0.5]]
    return dims, w, start, [d_pos, d_neg, d_pos]


def test_zero_direction_component(orc):
    dims, w, start, d = zero_component_rays()
    assert math.copysign(1.0, d[1][1]) < 0
    r = orc.cast_rays(w, dims, start, d, [384, 384, 384], 384)
    assert r["hit"][0] == -1 and r["steps"][0] == 384  # stuck in its cell until the budget ends
    # -0: an ordinary walk in the xz plane from (0.5, 0.5) along (0.6, 0.8)
    e = exact_first_solid(w, dims, [0.5, 1.5, 0.5], [0.6, 1e-9, 0.8], 384)
    assert (r["hit"][1], r["steps"][1]) == e and r["hit"][1] >= 0
    assert r["hit"][2] == idx(dims, 0, 1, 6) and r["steps"][2] == 6  # straight along z


def step_bias_rays():
    """fshader.glsl:79-81 adds 1e-6 to each direction component before |.|
    in the step sizes, so a step along +0.5 is 1/0.500001 = 1.999996 and one
    along -0.5 is 1/0.499999 = 2.000004 (2 without the bias).  Each ray is
    built so that its first y boundary lies between the biased and the
    unbiased second x boundary, 1 + step:
      A: dir (+0.5, 0.25), first y boundary at 4 * (1 - 0.2500005) = 2.999998:
         1 + 1.999996 is smaller, so x is taken twice and the ray enters
         (2, 0, 0); unbiased it would take y and enter (1, 1, 0);
      B: dir (-0.5, 0.25), first y boundary at 4 * (1 - 0.2499995) = 3.000002:
         1 + 2.000004 is larger, so y comes second and the ray enters
         (6, 1, 0); unbiased it would take x again and enter (5, 0, 0)."""
    dims = (8, 4, 2)
    w = empty_grid(dims)
    for c in [(2, 0, 0), (1, 1, 0), (6, 1, 0), (5, 0, 0)]:
        put(w, dims, *c, 0x445566)
    start = [[0.5, 0.2500005, 0.5], [7.5, 0.2499995, 0.5]]
    d = [np.array([0.5, 0.25, 1e-4], np.float32), np.array([-0.5, 0.25, 1e-4], np.float32)]
    return dims, w, start, d


def test_step_bias_shortens_positive_steps(orc):
    dims, w, start, d = step_bias_rays()
    f = np.float32
    one = f(1)
    step_pos = one / abs(f(0.5) + f(1e-6))
    step_neg = one / abs(f(-0.5) + f(1e-6))
    y_a = (one - f(start[0][1])) / f(0.25)
    y_b = (one - f(start[1][1])) / f(0.25)
    assert one + step_pos < y_a < one + one + one < y_b < one + step_neg  # the premise above
    r = orc.cast_rays(w, dims, start, d, [384, 384], 384)
    assert r["hit"][0] == idx(dims, 2, 0, 0) and r["steps"][0] == 2
    assert r["hit"][1] == idx(dims, 6, 1, 0) and r["steps"][1] == 2


# ----------------------------------------------------------- shading closed forms
def shading_world():
    dims = (32, 32, 32)
    w = empty_grid(dims)
    for x in range(32):
        for z in range(32):
            put(w, dims, x, 4, z, 0xC86432)  # floor at y = 4, colour (200, 100, 50)
    return dims, w


def down_frame(vrt, dims, light_pos):
    f = vrt.frame_defaults(dims[0], 8, 8)
    f.cam_pos[:] = [16.3, 20.7, 16.2]
    f.rotate_matrix[:] = vrt.camera_matrix(math.pi / 2 - 0.05, 0.3).tolist()  # looking down
    f.light_pos[:] = light_pos
    f.focal = 4.0
    return f


def unorm8(c):
    c = np.float32(c)
    k = c if 0 < c < 1 else (np.float32(1) if c >= 1 else np.float32(0))
    return int(np.float32(k) * np.float32(255.0) + np.float32(0.5))


def test_ambient_only_when_sun_is_below(orc, vrt):
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    f = down_frame(vrt, dims, [16.0, -500.0, 16.0])  # sun under the floor: lambert term 0
    r = orc.render(wdf, dims, f, 8, 8)
    assert np.all(r["hit"] >= 0)
    m = np.float32(0.4)
    expect = [unorm8(np.float32(np.float32(c) / np.float32(255.0)) * m) for c in (200, 100, 50)]
    assert np.all(r["rgb"].reshape(-1, 3) == expect)


def test_fully_lit_floor(orc, vrt):
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    f = down_frame(vrt, dims, [16.0, 1e6, 16.0])  # sun straight above: lambert ~1
    r = orc.render(wdf, dims, f, 8, 8)
    assert np.all(r["shadow_eval"] == 1) and np.all(r["shadow_occ"] == 0)
    assert np.all(r["face"] == 1 * 3 + 0)  # entered through the top face (step -1 on y)
    # 0.4 + 0.8 * ~1 = 1.2: red saturates, green 100/255*1.2, blue 50/255*1.2
    rgb = r["rgb"].reshape(-1, 3)
    assert np.all(rgb[:, 0] == 240) and np.all(np.abs(rgb[:, 1].astype(int) - 120) <= 1)


def test_overbright_caps_at_1_25(orc, vrt):
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    f = down_frame(vrt, dims, [16.0, 1e6, 16.0])
    for i in range(4):
        L = f.lights[i]
        L.pos[:] = [16.0 + i, 8.0, 16.0]
        L.rgb[:] = [5.0, 5.0, 5.0]
        L.radius = 64.0
    r = orc.render(wdf, dims, f, 8, 8)
    # sun 1.2, first light pushes past 1.25, second iteration clamps to 1.25 and stops
    assert np.all(r["shadow_eval"] == 0b11)
    expect = [unorm8(np.float32(np.float32(c) / np.float32(255.0)) * np.float32(1.25)) for c in (200, 100, 50)]
    assert np.all(r["rgb"].reshape(-1, 3) == expect)


# Lights 2^26 cells straight above the floor's top face (y = 5) make the
# lighting terms exact: from a floor point the offset to the light is
# (lx, 2^26, lz) with |lx|, |lz| < 16 once 2^26 + 8 - 5 is rounded to float
# (spacing 8 there), lx^2 + 2^52 + lz^2 rounds to 2^52, so the distance is
# 2^26, the direction is (lx / 2^26, 1, lz / 2^26) and N.L = 1 exactly.
HIGH = float(2 ** 26 + 8)


def high_light(L, x, radius, rgb, z=16.0):
    L.pos[:] = [x, HIGH, z]
    L.radius = radius
    L.rgb[:] = [rgb, rgb, rgb]


def lit(c, m):
    return unorm8(np.float32(np.float32(c) / np.float32(255.0)) * np.float32(m))


def light_range_frame(vrt, dims):
    """Sun overhead (multiplier 0.4 + 0.8 = 1.2); lights whose distance 2^26
    equals, exceeds and falls short of their radius."""
    f = down_frame(vrt, dims, [16.0, HIGH, 16.0])
    high_light(f.lights[0], 16.0, float(2 ** 26), 0.0)      # d == r: in range, falloff 0
    high_light(f.lights[1], 16.0, float(2 ** 26 - 4), 0.0)  # d > r: skipped
    high_light(f.lights[2], 0.0, float(2 ** 60), 0.0)       # x == 0 is a placed light (unused slots hold -1)
    return f


def test_light_at_exactly_its_radius_is_evaluated(orc, vrt):
    """fshader.glsl:167 tests localLightDist <= radius and :162 pos >= 0."""
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, light_range_frame(vrt, dims), 8, 8)
    assert np.all(r["hit"] >= 0)
    assert np.all(r["shadow_eval"] == 0b1011) and np.all(r["shadow_occ"] == 0)
    m = np.float32(0.4) + np.float32(0.8)
    assert np.all(r["rgb"].reshape(-1, 3) == [lit(c, m) for c in (200, 100, 50)])


def axis_range_frame(vrt, dims):
    """Sun overhead; two lights 2^26 cells along +x from the hit points (x =
    2^26 + 16: every hit x near 16 rounds the x difference to exactly 2^26,
    and the y and z differences vanish beside 2^52 in the squared length), of
    radius 2^26 (in range) and 2^26 - 8 (out): the range decision rests on
    the x difference alone, at its boundary."""
    f = down_frame(vrt, dims, [16.0, HIGH, 16.0])
    for i, radius in ((0, float(2 ** 26)), (1, float(2 ** 26 - 8))):
        L = f.lights[i]
        L.pos[:] = [float(2 ** 26 + 16), 6.5, 16.0]
        L.radius = radius
        L.rgb[:] = [1.0, 1.0, 1.0]
    return f


def test_light_range_decided_exactly_along_x(orc, vrt):
    """fshader.glsl:171: length <= radius with the length carried by one axis."""
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, axis_range_frame(vrt, dims), 8, 8)
    assert np.all(r["hit"] >= 0)
    assert np.all((r["shadow_eval"] & 0b110) == 0b010)  # light 0 evaluated, light 1 not
    m = np.float32(0.4) + np.float32(0.8)  # light 0 adds N.L * 0 (falloff 0 at d == r)
    assert np.all(r["rgb"].reshape(-1, 3) == [lit(c, m) for c in (200, 100, 50)])


def overbright_frame(vrt, dims):
    """Sun overhead gives 0.4 + 0.8 = 1.2 and light 0 (falloff 1, since
    2^60 - 2^26 rounds to 2^60) adds 0.05: float32 rounds 1.2 + 0.05 to
    exactly 1.25, so the >= 1.25 test at the top of the next iteration
    (fshader.glsl:158) stops the loop before light 1."""
    f = down_frame(vrt, dims, [16.0, HIGH, 16.0])
    for i in range(3):
        high_light(f.lights[i], 16.0, float(2 ** 60), 0.05)
    return f


def test_overbright_stops_at_exactly_1_25(orc, vrt):
    f32 = np.float32
    assert (f32(0.4) + f32(0.8)) + f32(0.05) == f32(1.25)  # the premise above
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, overbright_frame(vrt, dims), 8, 8)
    assert np.all(r["shadow_eval"] == 0b11) and np.all(r["shadow_occ"] == 0)
    assert np.all(r["rgb"].reshape(-1, 3) == [lit(c, 1.25) for c in (200, 100, 50)])


def clamp_frame(vrt, dims, slots):
    """Sun overhead (1.2) and lights of colour 0.5 with falloff 1 and N.L = 1
    in the given slots only (the others unused, -1): the first one lifts the
    multiplier to 1.7."""
    f = down_frame(vrt, dims, [16.0, HIGH, 16.0])
    for i in slots:
        high_light(f.lights[i], 16.0, float(2 ** 60), 0.5)
    return f


@pytest.mark.parametrize("slots,m,evaluated", [((0,), 1.25, 0b11), ((15,), 1.7, 0b1 | 2 << 15),
                                               ((14, 15), 1.25, 0b1 | 2 << 14)])
def test_overbright_clamp_needs_a_following_iteration(orc, vrt, slots, m, evaluated):
    """The clamp to 1.25 is the first statement of the light loop
    (fshader.glsl:158-164), so it happens only if an iteration follows the
    light that crossed 1.25, placed or not: after slot 0 it does (slot 1's
    iteration clamps), after slot 15 it does not (no clamp after the last
    light); with slots 14 and 15 placed, slot 15's iteration clamps and
    stops before evaluating it."""
    f32 = np.float32
    m_after = (f32(0.4) + f32(0.8)) + f32(0.5)
    assert m_after > f32(1.25)
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, clamp_frame(vrt, dims, slots), 8, 8)
    assert np.all(r["shadow_eval"] == evaluated) and np.all(r["shadow_occ"] == 0)
    mm = f32(1.25) if m == 1.25 else m_after
    assert np.all(r["rgb"].reshape(-1, 3) == [lit(c, mm) for c in (200, 100, 50)])


def falloff_frame(vrt, dims, radius):
    """Sun under the floor (lambert 0, multiplier stays 0.4) and one light of
    colour 0.5 at distance 2^26 with radius `radius`."""
    f = down_frame(vrt, dims, [16.0, -500.0, 16.0])
    high_light(f.lights[0], 16.0, radius, 0.5)
    return f


@pytest.mark.parametrize("log_r,fall", [(27, 0.5), (28, 0.75), (30, 0.9375)])
def test_light_falloff_is_linear_in_distance(orc, vrt, log_r, fall):
    """Colour * N.L * (radius - d) / radius (fshader.glsl:170-177) with
    N.L = 1 and d = 2^26: exactly 1 - 2^(26 - log_r)."""
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, falloff_frame(vrt, dims, float(2 ** log_r)), 8, 8)
    assert np.all(r["shadow_eval"] == 0b11) and np.all(r["shadow_occ"] == 0b00)
    m = np.float32(0.4) + np.float32(0.5 * fall)
    assert np.all(r["rgb"].reshape(-1, 3) == [lit(c, m) for c in (200, 100, 50)])


def roof_world():
    dims, w = shading_world()
    for x in range(32):
        for z in range(32):
            put(w, dims, x, 28, z, 0x808080)  # a roof at y = 28 over the floor at y = 4
    return dims, w


def roof_frame(vrt, dims):
    """Camera (y = 20.7) between floor and roof.  The sun overhead is behind
    the roof.  Light 0 (y = 12) is under the roof: its shadow ray gets
    int(d + 1) iterations (fshader.glsl:175), d < 14 here, and reaching the
    roof takes at least 23 (one per y boundary from 5 to 28), so the roof
    cannot shadow it.  Light 1, 2^26 up, is behind the roof."""
    f = down_frame(vrt, dims, [16.0, 1e6, 16.0])
    L = f.lights[0]
    L.pos[:] = [16.0, 12.0, 16.0]
    L.radius = 64.0
    L.rgb[:] = [0.1, 0.1, 0.1]
    high_light(f.lights[1], 16.0, float(2 ** 60), 0.1)
    return f


def test_local_shadow_rays_stop_at_the_light(orc, vrt):
    dims, w = roof_world()
    wdf = orc.depth_field(w, dims)
    r = orc.render(wdf, dims, roof_frame(vrt, dims), 8, 8)
    assert np.all(r["hit"] >= 0) and np.all(r["face"] == 1 * 3 + 0)  # all on the floor
    assert np.all(r["shadow_eval"] == 0b111)
    assert np.all(r["shadow_occ"] == 0b101)  # sun and light 1 behind the roof, light 0 not


def test_depth_view_is_step_count(orc, vrt):
    dims, w = shading_world()
    wdf = orc.depth_field(w, dims)
    f = down_frame(vrt, dims, [16.0, 1e6, 16.0])
    f.view_depth_field = 1
    r = orc.render(wdf, dims, f, 8, 8)
    grey = [unorm8(np.float32(s) / np.float32(100.0)) for s in r["steps"].ravel()]
    assert np.all(r["rgb"].reshape(-1, 3)[:, 0] == grey) and r["rays"] == 64


# ------------------------------------------------------------ depth field / edits
def small_world(seed, dims=(24, 16, 20), p=0.12):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    return np.where(rng.random(n) < p, rng.integers(0, 1 << 24, n), -1).astype(np.int32)


def test_depth_field_order_independent(orc):
    dims = (24, 16, 20)
    w = small_world(1, dims)
    full = orc.depth_field(w, dims)
    seq = w.copy()
    rng = np.random.default_rng(2)
    cells = [(x, y, z) for z in range(dims[2]) for y in range(dims[1]) for x in range(dims[0])]
    for i in rng.permutation(len(cells)):
        orc.lib().orc_fix_depth_field(seq.ctypes.data_as(orc.P), *dims, *cells[i])
    assert np.array_equal(full, seq)


def test_depth_field_values_are_cube_distances(orc):
    """Every jump word equals minus the smallest cube distance to a solid voxel
    among the table's offsets (searched here by brute force over the 15^3 box),
    written only when that distance is at least 2."""
    dims = (20, 14, 18)
    w = small_world(3, dims, p=0.03)
    df = orc.depth_field(w, dims)
    X, Y, Z = dims
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                v = w[idx(dims, x, y, z)]
                got = df[idx(dims, x, y, z)]
                if v >= 0:
                    assert got == v
                    continue
                best = 6.0
                for oz in range(-7, 8):
                    for oy in range(-7, 8):
                        for ox in range(-7, 8):
                            if ox * ox + oy * oy + oz * oz > 49:
                                continue
                            nx, ny, nz = x + ox, y + oy, z + oz
                            if 0 <= nx < X and 0 <= ny < Y and 0 <= nz < Z and w[idx(dims, nx, ny, nz)] >= 0:
                                gap = math.sqrt(sum(max(abs(c) - 1, 0) ** 2 for c in (ox, oy, oz)))
                                best = min(best, gap)
                if best < 2.0:
                    assert got == -1
                else:
                    assert bits_f32(int(got)) == np.float32(-best)


@pytest.mark.parametrize("p", [0.25, 0.02])
def test_remove_sphere_equals_rebuild_inside_repair_radius(orc, p):
    """removeSphere (level.cpp:30-56) carves radius r and repairs the depth
    field within r + 3; there it must equal a full rebuild.  The sparse world
    has cells beyond r whose nearest solid was carved."""
    dims = (40, 24, 36)
    w = small_world(4, dims, p=p)
    df = orc.depth_field(w, dims)
    before = df.copy()
    c, r = (18, 10, 17), 6
    orc.remove_sphere(df, dims, c, r)
    X, Y, Z = dims
    # the same carve on the raw world, then a full build
    raw = w.copy()
    for z in range(-r, r):
        for y in range(-r, r):
            for x in range(-r, r):
                if x * x + y * y + z * z < r * r:
                    raw[idx(dims, c[0] + x, c[1] + y, c[2] + z)] = -1
    rebuilt = orc.depth_field(raw, dims)
    R = r + 3
    for z in range(-R, R):
        for y in range(-R, R):
            for x in range(-R, R):
                if x * x + y * y + z * z < R * R:
                    k = idx(dims, c[0] + x, c[1] + y, c[2] + z)
                    assert df[k] == rebuilt[k]
    # removal never shortens a jump: |jump| after >= |jump| before on cells empty before
    def jump(v):
        return 0.0 if v == -1 else -bits_f32(int(v))
    for k in np.nonzero(before < 0)[0]:
        assert jump(df[k]) >= jump(before[k])
    assert np.all((df >= 0) <= (before >= 0))  # solids only disappear
    if p < 0.1:
        shell = [idx(dims, c[0] + x, c[1] + y, c[2] + z) for z in range(-R, R) for y in range(-R, R)
                 for x in range(-R, R) if r * r <= x * x + y * y + z * z < R * R]
        assert np.count_nonzero(df[shell] != before[shell]) > 50  # the repair reached past the carve


# ------------------------------------------------------------ reference world
@pytest.mark.parametrize("dims", [(512, 96, 512), (200, 60, 130), (25, 40, 31)])
def test_reference_world_product_equals_oracle(orc, vrt, dims):
    a = orc.world_reference(dims)
    b = vrt.reference_world(dims)
    assert np.array_equal(a, b), "differs at %d voxels" % int((a != b).sum())


def test_reference_world_closed_forms(orc):
    dims = (512, 96, 512)
    w = orc.world_reference(dims).reshape(512, 96, 512)  # [z][y][x]
    # layers (src/level.cpp:93-125) away from trees and the origin quirk
    assert w[300, 10, 301] == ((90 + 5 * ((301 + 10 + 300) % 3)) * 0x010101)
    assert w[300, 30, 301] == (((120 + 5 * ((301 + 30 + 300) % 3)) << 16) | ((100 + 5 * ((301 + 30 + 300) % 3)) << 8))
    assert w[300, 36, 301] == ((10 << 16) | ((130 + 5 * ((301 + 36 + 300) % 3)) << 8) | 10)
    assert np.all(w[:, 37:, :][w[:, 37:, :] >= 0] > 0)  # only trees above the grass
    # 16 x 19 trees: x in 30..480 step 30, z in 25..500 step 25
    assert ((w[:, 40, :] >= 0).sum()) > 16 * 19 * 9
    # destroyVoxel with local offsets clears cells at the origin (quirk kept)
    assert w[0, 0, 0] == -1 and w[0, 5, 0] == -1 and w[0, 0, 5] == -1 and w[5, 0, 0] == -1
    assert w[0, 0, 6] >= 0 and w[0, 6, 0] >= 0
