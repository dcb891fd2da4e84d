"""The headless C++ driver (voxel-rt_amd/app/vrt_headless.cpp): the
reference's frame loop on the reference's own world, through the C-ABI.
Recorded frames are re-rendered by the oracle after replaying the same
edit log."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from test_gpu_parity import assert_frames_match

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "voxel-rt_amd", "vrt_headless")


def test_headless_driver_frames_match_oracle(gpu, orc, tmp_path):
    W, H, N = 160, 120, 26
    r = subprocess.run([APP, "--frames", str(N), "--width", str(W), "--height", str(H), "--out", str(tmp_path),
                        "--record"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    edits = [tuple(int(v) for v in ln.split()) for ln in open(tmp_path / "edits.txt").read().split("\n") if ln]
    assert len(edits) == N // 8
    dims = (512, 96, 512)
    world = orc.depth_field(orc.world_reference(dims), dims)
    fsize = C.sizeof(gpu.vrt_frame)
    applied = 0
    checked = 0
    for k in range(N):
        while applied < len(edits) and edits[applied][0] < k:  # edits of earlier iterations
            _, x, y, z, rad = edits[applied]
            orc.remove_sphere(world, dims, (x, y, z), rad)
            applied += 1
        if k not in (0, 5, 7, 8, 9, 16, 17, 21, 24, 25):
            continue
        raw = open(tmp_path / ("frame_%04d.bin" % k), "rb").read()
        frame = gpu.vrt_frame.from_buffer_copy(raw[:fsize])
        rgb = np.frombuffer(raw[fsize:fsize + 3 * W * H], np.uint8).reshape(H, W, 3)
        hit = np.frombuffer(raw[fsize + 3 * W * H:], np.int64).reshape(H, W)
        ref = orc.render(world, dims, frame, W, H)
        assert np.array_equal(hit, ref["hit"]), "frame %d hit map differs at %d pixels" % (k, int((hit != ref["hit"]).sum()))
        assert int(np.abs(rgb.astype(np.int16) - ref["rgb"].astype(np.int16)).max()) <= 1, "frame %d rgb" % k
        checked += 1
    assert checked == 10
    # the scripted loop did what it says: a light was placed, the depth view toggled
    last = gpu.vrt_frame.from_buffer_copy(open(tmp_path / ("frame_%04d.bin" % 21), "rb").read()[:fsize])
    assert last.view_depth_field == 1 and last.lights[0].pos[0] >= 0


def test_headless_entity_frames_match_oracle(gpu, orc, tmp_path):
    """--entity: the reference's test entity moves its voxel on the host array
    and re-uploads its box every frame (updatePartialGeometry), next to the
    on-device sphere carves.  The oracle replays both edit streams."""
    W, H, N = 96, 72, 40
    r = subprocess.run([APP, "--frames", str(N), "--width", str(W), "--height", str(H), "--out", str(tmp_path),
                        "--record", "--entity"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in open(tmp_path / "edits.txt").read().split("\n") if ln]
    spheres = [tuple(int(v) for v in ln) for ln in lines if ln[0] != "e"]
    moves = [tuple(int(v) for v in ln[1:]) for ln in lines if ln[0] == "e"]
    assert len(spheres) == N // 8 and len(moves) == N
    # the centre starts at (200, 55, 170) and crosses into z = 171 after 30 frames of 2/60
    assert moves[0][1:4] == (200, 55, 170) and moves[-1][4:7] == (200, 55, 171)
    dims = (512, 96, 512)
    world = orc.depth_field(orc.world_reference(dims), dims)
    fsize = C.sizeof(gpu.vrt_frame)
    si = 0
    checked = 0
    for k in range(N):
        if k > 0:  # the edits of iteration k-1: its carve (if any), then the entity move
            if si < len(spheres) and spheres[si][0] == k - 1:
                _, x, y, z, rad = spheres[si]
                orc.remove_sphere(world, dims, (x, y, z), rad)
                si += 1
            _, fx, fy, fz, tx, ty, tz = moves[k - 1]
            orc.destroy_voxel(world, dims, fx, fy, fz)
            orc.place_voxel(world, dims, tx, ty, tz, 0x00AA0000)
        if k % 3 and k != N - 1:
            continue
        raw = open(tmp_path / ("frame_%04d.bin" % k), "rb").read()
        frame = gpu.vrt_frame.from_buffer_copy(raw[:fsize])
        rgb = np.frombuffer(raw[fsize:fsize + 3 * W * H], np.uint8).reshape(H, W, 3)
        hit = np.frombuffer(raw[fsize + 3 * W * H:], np.int64).reshape(H, W)
        ref = orc.render(world, dims, frame, W, H)
        assert np.array_equal(hit, ref["hit"]), "frame %d hit map differs at %d pixels" % (k, int((hit != ref["hit"]).sum()))
        assert int(np.abs(rgb.astype(np.int16) - ref["rgb"].astype(np.int16)).max()) <= 1, "frame %d rgb" % k
        checked += 1
    assert checked == 14
    # the entity voxel is on screen: some frame's hit map holds its index
    ent = [200 + 512 * 55 + 512 * 96 * z for z in (170, 171)]
    seen = 0
    for k in range(0, N, 3):
        raw = open(tmp_path / ("frame_%04d.bin" % k), "rb").read()
        hit = np.frombuffer(raw[fsize + 3 * W * H:], np.int64)
        seen += int(np.isin(hit, ent).sum())
    assert seen > 0


@pytest.mark.parametrize("lights", [0, 4])
def test_reference_world_startup_view_matches_oracle(gpu, orc, lights):
    """The reference's world and start-up frame (src/main.cpp:26-39, 800x600,
    identity rotateMatrix, lightPos = startLightPos), optionally with local
    lights placed the way the T key does (src/controls.cpp:46-49)."""
    dims = (512, 96, 512)
    world = gpu.reference_world(dims)
    ctx = gpu.Context(dims)
    ctx.upload(world)
    ctx.build_depth_field()
    wdf = ctx.download()
    rng = np.random.default_rng(lights)
    n = 100000
    xyz = np.stack([rng.integers(0, 512, n), rng.integers(0, 96, n), rng.integers(0, 512, n)], 1).astype(np.int32)
    idx = xyz[:, 0].astype(np.int64) + 512 * xyz[:, 1] + 512 * 96 * xyz[:, 2].astype(np.int64)
    assert np.array_equal(wdf[idx], orc.fix_depth_values(world, dims, xyz))
    f = gpu.frame_defaults(512, 800, 600)
    for i in range(lights):
        assert gpu.lib().vrt_place_local_light(C.byref(f), 195.0 + 6 * i, 55.0 - 1.5, 165.0 + 4 * i, 0.5) == i
    tgt = gpu.Target(800, 600, depth=True, aux=True)
    ctx.stats(reset=True)
    ctx.draw(tgt, frame=f)
    g = tgt.download()
    g.update(ctx.stats(reset=True))
    r = orc.LazyWorld(world, dims).render(f, 800, 600)  # from the raw world, not the GPU's grid
    assert_frames_match(g, r, "reference start-up view, %d lights" % lights)
    assert (r["hit"] >= 0).mean() > 0.3
    if lights:
        assert (r["shadow_eval"] > 1).any()
    ctx.close()
