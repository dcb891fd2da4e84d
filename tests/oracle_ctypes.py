"""ctypes access to the CPU oracle (oracle/liboracle.so) for tests and the
benchmark's cpu_baseline leg.  The oracle is the checker, never the product."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# VRT_ORACLE_LIB names another build of the oracle; only the mutant test sets it,
# to run the pins against deliberately broken builds.
LIB_PATH = os.environ.get("VRT_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")
GLM_PROBE = os.path.join(ORACLE_DIR, "_ref", "glm_probe")

P = C.c_void_p
_lib = None


class orc_outputs(C.Structure):
    _fields_ = [("rgb", P), ("depth", P), ("hit", P), ("face", P), ("eval", P), ("occ", P), ("steps", P)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH) and not os.environ.get("VRT_ORACLE_LIB"):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)
        L = C.CDLL(LIB_PATH)
        i64, i32 = C.c_int64, C.c_int32
        L.orc_depth_table.argtypes = [P, P, i32]
        L.orc_depth_table.restype = C.c_int
        L.orc_fix_depth_field.argtypes = [P, i64, i64, i64, i32, i32, i32]
        L.orc_fix_depth_values.argtypes = [P, i64, i64, i64, P, i64, P, i32]
        L.orc_build_depth_field.argtypes = [P, i64, i64, i64, i32]
        L.orc_place_voxel.argtypes = [P, i64, i64, i64, i32, i32, i32, i32]
        L.orc_destroy_voxel.argtypes = [P, i64, i64, i64, i32, i32, i32]
        L.orc_remove_sphere.argtypes = [P, i64, i64, i64, i32, i32, i32, i32]
        L.orc_cast_rays.argtypes = [P, i64, i64, i64, i64, P, P, P, i32, P, P, P, P]
        L.orc_render.argtypes = [P, i64, i64, i64, P, i32, i32, P, i64, C.POINTER(orc_outputs), P, i32]
        L.orc_render_lazy.argtypes = [P, i64, i64, i64, P, i32, P, P, i32, i32, P, i64, C.POINTER(orc_outputs), P,
                                      i32]
        L.orc_lazy_words.argtypes = [P, i64, i64, i64, P, i32, P, P, i64, P, i32]
        L.orc_max_threads.restype = C.c_int
        L.orc_normalize.argtypes = [P, P]
        L.orc_length.argtypes = [P]
        L.orc_length.restype = C.c_float
        L.orc_dot.argtypes = [P, P]
        L.orc_dot.restype = C.c_float
        L.orc_primary_dir.argtypes = [P, C.c_float, C.c_float, P]
        L.orc_world_reference.argtypes = [P, i64, i64, i64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


def threads():
    return int(os.environ.get("ORACLE_THREADS", "0")) or lib().orc_max_threads()


def depth_table():
    dist = np.zeros(4096, np.float32)
    off = np.zeros(3 * 4096, np.int32)
    n = lib().orc_depth_table(_p(dist), _p(off), 4096)
    return dist[:n].copy(), off[:3 * n].reshape(n, 3).copy()


def depth_field(world, dims, nthreads=None):
    """computeDepthField over a copy of `world`."""
    w = np.array(world, dtype=np.int32, copy=True)
    lib().orc_build_depth_field(_p(w), dims[0], dims[1], dims[2], nthreads or threads())
    return w


def fix_depth_values(world, dims, xyz, nthreads=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
    out = np.empty(xyz.shape[0], np.int32)
    lib().orc_fix_depth_values(_p(world), dims[0], dims[1], dims[2], _p(xyz), xyz.shape[0], _p(out),
                               nthreads or threads())
    return out


def remove_sphere(world, dims, c, r):
    """removeSphere in place."""
    lib().orc_remove_sphere(_p(world), dims[0], dims[1], dims[2], int(c[0]), int(c[1]), int(c[2]), int(r))


def place_voxel(world, dims, x, y, z, v):
    lib().orc_place_voxel(_p(world), dims[0], dims[1], dims[2], int(x), int(y), int(z), int(v))


def destroy_voxel(world, dims, x, y, z):
    lib().orc_destroy_voxel(_p(world), dims[0], dims[1], dims[2], int(x), int(y), int(z))


def cast_rays(world, dims, start, direction, dist, render_dist):
    start = np.ascontiguousarray(start, np.float32).reshape(-1, 3)
    direction = np.ascontiguousarray(direction, np.float32).reshape(-1, 3)
    dist = np.ascontiguousarray(dist, np.int32).reshape(-1)
    n = start.shape[0]
    hit = np.empty(n, np.int64)
    pos = np.empty((n, 3), np.float32)
    face = np.empty(n, np.uint8)
    steps = np.empty(n, np.uint32)
    lib().orc_cast_rays(_p(world), dims[0], dims[1], dims[2], n, _p(start), _p(direction), _p(dist),
                        int(render_dist), _p(hit), _p(pos), _p(face), _p(steps))
    return {"hit": hit, "hit_pos": pos, "face": face, "steps": steps}


def render(world_df, dims, frame, width, height, pixels=None, nthreads=None):
    """One frame of fshader.glsl main() over `world_df` (depth field built).
    pixels: optional array of pixel ids (row*W + px); outputs then follow it."""
    if pixels is None:
        shape, n, pix = (height, width), width * height, None
    else:
        pix = np.ascontiguousarray(pixels, np.int64)
        shape, n = (pix.shape[0],), pix.shape[0]
    out = {"rgb": np.empty(shape + (3,), np.uint8), "depth": np.empty(shape, np.float32),
           "hit": np.empty(shape, np.int64), "face": np.empty(shape, np.uint8),
           "shadow_eval": np.empty(shape, np.uint32), "shadow_occ": np.empty(shape, np.uint32),
           "steps": np.empty(shape, np.uint32)}
    o = orc_outputs(_p(out["rgb"]), _p(out["depth"]), _p(out["hit"]), _p(out["face"]), _p(out["shadow_eval"]),
                    _p(out["shadow_occ"]), _p(out["steps"]))
    totals = np.zeros(2, np.uint64)
    lib().orc_render(_p(world_df), dims[0], dims[1], dims[2], C.byref(frame), int(width), int(height), _p(pix),
                     n if pix is not None else 0, C.byref(o), _p(totals), nthreads or threads())
    out["rays"] = int(totals[0])
    out["dda_steps"] = int(totals[1])
    return out


class LazyWorld:
    """The world after its depth-field build and the first `version` edits of
    an edit log, evaluated by the oracle on demand from the raw world (solid
    colours and -1): for worlds too large for orc.depth_field on the CPU.
    edits: (cx, cy, cz, r) removeSphere calls in order; the memo of the
    version-0 words is kept across calls."""

    def __init__(self, raw, dims, edits=()):
        self.raw = np.ascontiguousarray(raw, np.int32)
        self.dims = tuple(int(d) for d in dims)
        assert self.raw.size == self.dims[0] * self.dims[1] * self.dims[2]
        self.memo = np.zeros(self.raw.size, np.int32)
        self.set_edits(edits)

    def set_edits(self, edits):
        self.edits = np.ascontiguousarray(np.asarray(list(edits), np.int32).reshape(-1, 4))

    def render(self, frame, width, height, version=None, pixels=None, nthreads=None):
        T = len(self.edits) if version is None else int(version)
        assert 0 <= T <= len(self.edits)
        if pixels is None:
            shape, n, pix = (height, width), width * height, None
        else:
            pix = np.ascontiguousarray(pixels, np.int64)
            shape, n = (pix.shape[0],), pix.shape[0]
        out = {"rgb": np.empty(shape + (3,), np.uint8), "depth": np.empty(shape, np.float32),
               "hit": np.empty(shape, np.int64), "face": np.empty(shape, np.uint8),
               "shadow_eval": np.empty(shape, np.uint32), "shadow_occ": np.empty(shape, np.uint32),
               "steps": np.empty(shape, np.uint32)}
        o = orc_outputs(_p(out["rgb"]), _p(out["depth"]), _p(out["hit"]), _p(out["face"]), _p(out["shadow_eval"]),
                        _p(out["shadow_occ"]), _p(out["steps"]))
        totals = np.zeros(2, np.uint64)
        X, Y, Z = self.dims
        lib().orc_render_lazy(_p(self.raw), X, Y, Z, _p(self.edits), T, _p(self.memo), C.byref(frame), int(width),
                              int(height), _p(pix), n if pix is not None else 0, C.byref(o), _p(totals),
                              nthreads or threads())
        out["rays"] = int(totals[0])
        out["dda_steps"] = int(totals[1])
        return out

    def words(self, idx, version=None, nthreads=None):
        """The words of the listed linear cell indices at `version`."""
        T = len(self.edits) if version is None else int(version)
        idx = np.ascontiguousarray(idx, np.int64).reshape(-1)
        out = np.empty(idx.size, np.int32)
        X, Y, Z = self.dims
        lib().orc_lazy_words(_p(self.raw), X, Y, Z, _p(self.edits), T, _p(self.memo), _p(idx), idx.size, _p(out),
                             nthreads or threads())
        return out


def world_reference(dims=(512, 96, 512)):
    """initVoxels' world (src/level.cpp:82-138)."""
    w = np.empty(dims[0] * dims[1] * dims[2], np.int32)
    lib().orc_world_reference(_p(w), dims[0], dims[1], dims[2])
    return w


def normalize(v):
    v = np.ascontiguousarray(v, np.float32)
    o = np.empty(3, np.float32)
    lib().orc_normalize(_p(v), _p(o))
    return o


def length(v):
    v = np.ascontiguousarray(v, np.float32)
    return np.float32(lib().orc_length(_p(v)))


def dot(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().orc_dot(_p(a), _p(b)))


def primary_dir(frame, vx, vy):
    o = np.empty(3, np.float32)
    lib().orc_primary_dir(C.byref(frame), C.c_float(vx), C.c_float(vy), _p(o))
    return o


def glm_probe(lines):
    """Runs oracle/_ref/glm_probe (built only where the reference is present)."""
    r = subprocess.run([GLM_PROBE], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return [[int(w, 16) for w in ln.split()] for ln in r.stdout.strip().split("\n")]
