"""The C-ABI boundary on the CPU: libvrt.so loads, exports every function
include/vrt.h declares, and its host-only entry points (frame helpers,
placeLocalLight slot rule, world generator, error paths) behave; no GPU work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vrt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(vrt_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function(vrt):
    names = declared_functions()
    assert len(names) >= 45
    lib = C.CDLL(vrt.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python bindings cover the same set
    assert sorted(vrt.SIGNATURES) == names


def test_no_device_is_reported_cleanly(vrt):
    if vrt.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    rc = vrt.lib().vrt_create(C.byref(h), 8, 8, 8, 0)
    assert rc == vrt.VRT_ERR_NODEVICE and not h.value
    assert b"device" in vrt.lib().vrt_error_string()


def test_invalid_arguments(vrt):
    h = C.c_void_p()
    assert vrt.lib().vrt_create(C.byref(h), 0, 8, 8, 0) == vrt.VRT_ERR_INVALID
    assert vrt.lib().vrt_create(C.byref(h), 1 << 31, 8, 8, 0) == vrt.VRT_ERR_INVALID
    assert vrt.lib().vrt_draw(None, None) == vrt.VRT_ERR_INVALID
    assert vrt.lib().vrt_remove_sphere(None, 0, 0, 0, 3) == vrt.VRT_ERR_INVALID
    assert vrt.lib().vrt_frame_defaults(None, 512, 8, 8) == vrt.VRT_ERR_INVALID


def test_frame_defaults_are_the_reference_start_up(vrt):
    f = vrt.frame_defaults(512, 800, 600)
    assert list(f.cam_pos) == [195.0, 55.0, 155.0]  # src/main.cpp:26
    assert list(f.light_pos) == [256.0, 1536.0, 256.0]  # startLightPos, src/main.cpp:33
    assert list(f.rotate_matrix) == [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    assert f.aspect_ratio == np.float32(800 / 600) and f.focal == 1.0
    assert f.render_dist == 384 and f.view_depth_field == 0
    for L in f.lights:  # initLocalLights, src/render.cpp:304-311
        assert list(L.pos) == [-1.0, -1.0, -1.0] and list(L.rgb) == [0.0, 0.0, 0.0] and L.radius == 64.0


def test_place_local_light_takes_first_free_slot(vrt):
    f = vrt.frame_defaults(512, 8, 8)
    slots = [vrt.lib().vrt_place_local_light(C.byref(f), 10.0 + i, 20.0, 30.0, 0.5) for i in range(17)]
    assert slots == list(range(16)) + [-1]  # limit 16 (README "T to place local light (limit 16)")
    f.lights[5].pos[1] = -1.0  # a slot becomes free when any component is negative
    assert vrt.lib().vrt_place_local_light(C.byref(f), 1.0, 2.0, 3.0, 0.25) == 5
    assert list(f.lights[5].rgb) == [0.25, 0.25, 0.25]


def test_world_generator_deterministic_and_in_range(vrt):
    p = vrt.world_params((64, 48, 40), 7, caves=6, boxes=20, height_min=8, height_max=30)
    a = vrt.generate_world(p, threads=1)
    b = vrt.generate_world(p, threads=5)
    assert np.array_equal(a, b)
    assert np.all((a == -1) | ((a >= 0) & (a < (1 << 24))))
    w = a.reshape(40, 48, 64)
    for (x, z) in [(0, 0), (63, 39), (17, 23)]:
        h = vrt.surface(p, x, z)
        assert 8 <= h <= 30
    # terrain is solid at the bottom (caves aside) and empty at the top
    assert (w[:, 0, :] >= 0).mean() > 0.8 and np.all(w[:, 47, :] == -1)
    # a different seed gives a different world
    q = vrt.world_params((64, 48, 40), 8, caves=6, boxes=20, height_min=8, height_max=30)
    assert not np.array_equal(a, vrt.generate_world(q))


def test_world_params_rejected(vrt):
    p = vrt.world_params((64, 48, 40), 7)
    p.height_max = 2  # below height_min
    out = np.empty(64 * 48 * 40, np.int32)
    assert vrt.lib().vrt_world_generate(C.byref(p), out.ctypes.data_as(C.c_void_p), 1) == vrt.VRT_ERR_INVALID


C_CLIENT = r"""
#include <stdio.h>
#include "vrt.h"
int main(void) {
    vrt_frame f;
    float m[16];
    if (vrt_frame_defaults(&f, 512, 800, 600) != VRT_OK) return 1;
    if (vrt_place_local_light(&f, 10.0f, 20.0f, 30.0f, 0.5f) != 0) return 2; /* slot 0 */
    if (vrt_camera_matrix(0.25f, 1.5f, m) != VRT_OK) return 3;
    printf("%zu %zu %.9g %.9g %.9g %.9g\n", sizeof(vrt_frame), sizeof(vrt_target), f.cam_pos[0],
           f.lights[0].pos[1], f.lights[0].rgb[0], m[5]);
    return 0;
}
"""


def test_header_is_plain_c(vrt, tmp_path):
    """include/vrt.h compiles as C99 with warnings as errors, a C program links
    libvrt.so and calls its host-side helpers, and the struct sizes equal the
    ctypes mirror's."""
    src = tmp_path / "client.c"
    src.write_text(C_CLIENT)
    exe = tmp_path / "client"
    libdir = os.path.dirname(vrt.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", libdir, "-lvrt", "-Wl,-rpath," + libdir,
                    "-Wl,--allow-shlib-undefined"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(vrt.vrt_frame) and int(out[1]) == C.sizeof(vrt.vrt_target)
    f = vrt.frame_defaults(512, 800, 600)
    assert vrt.lib().vrt_place_local_light(C.byref(f), 10.0, 20.0, 30.0, 0.5) == 0
    assert float(out[2]) == f.cam_pos[0] and float(out[3]) == f.lights[0].pos[1] and float(out[4]) == f.lights[0].rgb[0]
    assert np.float32(float(out[5])) == np.float32(vrt.camera_matrix(0.25, 1.5)[5])
