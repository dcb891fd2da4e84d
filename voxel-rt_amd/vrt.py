"""Python host side of libvrt: ctypes bindings of include/vrt.h, device
buffers through the HIP runtime, and the synthetic scenes the benchmarks and
tests render.

The reference's host (src/render.cpp, src/main.cpp) is C++; this module is
for tests, benchmarks and scripting.  Every render goes through libvrt.so;
there is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvrt.so")
HIP_LIB = "/opt/rocm/lib/libamdhip64.so"

VRT_OK = 0
VRT_ERR_INVALID = -1
VRT_ERR_NOMEM = -2
VRT_ERR_HIP = -3
VRT_ERR_BUDGET = -4
VRT_ERR_NOTFOUND = -5
VRT_ERR_NODEVICE = -6
MAX_LOCAL_LIGHTS = 16
FACE_MISS = 255
FOV70_FOCAL = 1.0 / math.tan(math.radians(35.0))


class vrt_local_light(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("rgb", C.c_float * 3), ("radius", C.c_float), ("reserved", C.c_float)]


class vrt_frame(C.Structure):
    _fields_ = [("cam_pos", C.c_float * 3), ("light_pos", C.c_float * 3), ("rotate_matrix", C.c_float * 16),
                ("aspect_ratio", C.c_float), ("focal", C.c_float), ("view_depth_field", C.c_int32),
                ("render_dist", C.c_int32), ("lights", vrt_local_light * MAX_LOCAL_LIGHTS)]


class vrt_target(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.c_void_p), ("depth", C.c_void_p),
                ("hit", C.c_void_p), ("face", C.c_void_p), ("shadow_eval", C.c_void_p),
                ("shadow_occ", C.c_void_p), ("steps", C.c_void_p)]


class vrt_stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("dda_steps", C.c_uint64), ("pixels", C.c_uint64), ("draws", C.c_uint64)]


class vrt_view(C.Structure):
    _fields_ = [("frame", vrt_frame), ("target", vrt_target)]


class vrt_batch_config(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("chunk_tiles", C.c_int32),
                ("waves_per_cu", C.c_int32)]


class vrt_tune_entry(C.Structure):
    _fields_ = [("config", vrt_batch_config), ("ms", C.c_double)]


class vrt_server_config(C.Structure):
    _fields_ = [("framebuffer_budget", C.c_uint64), ("max_batch_views", C.c_int32), ("streams", C.c_int32),
                ("want_depth", C.c_int32), ("want_aux", C.c_int32), ("batch", vrt_batch_config),
                ("max_batch_pixels", C.c_int64)]


class vrt_result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("world_version", C.c_uint64), ("width", C.c_int32), ("height", C.c_int32),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_exec", C.c_double),
                ("t_gpu_done", C.c_double), ("t_done", C.c_double),
                ("target", vrt_target)]


class vrt_world_params(C.Structure):
    _fields_ = [("size", C.c_int64 * 3), ("seed", C.c_uint64), ("height_min", C.c_int32),
                ("height_max", C.c_int32), ("octaves", C.c_int32), ("base_cell", C.c_int32),
                ("caves", C.c_int32), ("cave_r_min", C.c_int32), ("cave_r_max", C.c_int32),
                ("boxes", C.c_int32), ("box_min", C.c_int32), ("box_max", C.c_int32), ("jitter", C.c_int32)]


P = C.c_void_p
I32, I64, U64, F32 = C.c_int32, C.c_int64, C.c_uint64, C.c_float
FP = C.POINTER(C.c_float)

# name -> (restype, argtypes); every function declared in include/vrt.h
SIGNATURES = {
    "vrt_error_string": (C.c_char_p, []),
    "vrt_version": (C.c_char_p, []),
    "vrt_device_count": (C.c_int, [C.POINTER(I32)]),
    "vrt_create": (C.c_int, [C.POINTER(P), I64, I64, I64, I32]),
    "vrt_destroy": (C.c_int, [P]),
    "vrt_set_stream": (C.c_int, [P, P]),
    "vrt_get_stream": (P, [P]),
    "vrt_sync": (C.c_int, [P]),
    "vrt_grid_device_ptr": (C.c_int, [P, C.POINTER(P)]),
    "vrt_grid_dims": (C.c_int, [P, C.POINTER(I64)]),
    "vrt_grid_modified": (C.c_int, [P]),
    "vrt_set_shadow_cutoff": (C.c_int, [P, I32]),
    "vrt_upload": (C.c_int, [P, P]),
    "vrt_upload_device": (C.c_int, [P, P]),
    "vrt_upload_region": (C.c_int, [P, P, I32, I32, I32, I32, I32, I32]),
    "vrt_upload_region_f": (C.c_int, [P, P, FP, FP]),
    "vrt_download": (C.c_int, [P, P]),
    "vrt_depth_table": (C.c_int, [FP, C.POINTER(I32), I32, C.POINTER(I32)]),
    "vrt_build_depth_field": (C.c_int, [P]),
    "vrt_fix_depth_field_box": (C.c_int, [P, I32, I32, I32, I32, I32, I32]),
    "vrt_remove_sphere": (C.c_int, [P, I32, I32, I32, I32]),
    "vrt_place_voxel": (C.c_int, [P, I32, I32, I32, I32]),
    "vrt_destroy_voxel": (C.c_int, [P, I32, I32, I32]),
    "vrt_frame_defaults": (C.c_int, [C.POINTER(vrt_frame), I64, I32, I32]),
    "vrt_init_local_lights": (C.c_int, [C.POINTER(vrt_frame)]),
    "vrt_place_local_light": (C.c_int, [C.POINTER(vrt_frame), F32, F32, F32, F32]),
    "vrt_mat4_rotate": (C.c_int, [FP, F32, FP, FP]),
    "vrt_mat4_mul": (C.c_int, [FP, FP, FP]),
    "vrt_camera_matrix": (C.c_int, [F32, F32, FP]),
    "vrt_light_orbit": (C.c_int, [F32, FP, FP]),
    "vrt_set_frame": (C.c_int, [P, C.POINTER(vrt_frame)]),
    "vrt_draw": (C.c_int, [P, C.POINTER(vrt_target)]),
    "vrt_draw_frame": (C.c_int, [P, C.POINTER(vrt_frame), C.POINTER(vrt_target)]),
    "vrt_read_stats": (C.c_int, [P, C.POINTER(vrt_stats), I32]),
    "vrt_cast_rays": (C.c_int, [P, I64, P, P, P, I32, P, P, P, P]),
    "vrt_batch_create": (C.c_int, [P, C.POINTER(vrt_view), I32, C.POINTER(vrt_batch_config), C.POINTER(P)]),
    "vrt_batch_launch": (C.c_int, [P]),
    "vrt_batch_destroy": (C.c_int, [P]),
    "vrt_draw_views_plain": (C.c_int, [P, C.POINTER(vrt_view), I32]),
    "vrt_draw_views_noalloc": (C.c_int, [P, C.POINTER(vrt_view), I32, C.POINTER(vrt_batch_config)]),
    "vrt_autotune_batch": (C.c_int, [P, C.POINTER(vrt_view), I32, I32, C.POINTER(vrt_batch_config),
                                     C.POINTER(vrt_tune_entry), I32, C.POINTER(I32)]),
    "vrt_server_create": (C.c_int, [P, C.POINTER(vrt_server_config), C.POINTER(P)]),
    "vrt_server_destroy": (C.c_int, [P]),
    "vrt_server_submit": (C.c_int, [P, C.POINTER(vrt_frame), I32, I32, C.c_double, C.POINTER(U64)]),
    "vrt_server_submit_edit": (C.c_int, [P, I32, I32, I32, I32, C.POINTER(U64)]),
    "vrt_server_pump": (C.c_int, [P, I32]),
    "vrt_server_poll": (C.c_int, [P, C.POINTER(vrt_result), I32, C.POINTER(I32)]),
    "vrt_server_release": (C.c_int, [P, U64]),
    "vrt_server_cancel": (C.c_int, [P, U64]),
    "vrt_server_status": (C.c_int, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32), C.POINTER(U64)]),
    "vrt_server_clock": (C.c_double, [P]),
    "vrt_selftest_division": (C.c_int, [P, C.c_uint32, C.c_uint32, C.POINTER(U64)]),
    "vrt_selftest_reciprocal": (C.c_int, [P, C.c_uint64, C.c_uint64, C.POINTER(U64)]),
    "vrt_world_generate": (C.c_int, [C.POINTER(vrt_world_params), P, I32]),
    "vrt_world_generate_device": (C.c_int, [P, C.POINTER(vrt_world_params)]),
    "vrt_world_surface": (C.c_int, [C.POINTER(vrt_world_params), I64, I64, C.POINTER(I32)]),
    "vrt_world_reference": (C.c_int, [I64, I64, I64, P]),
}

_lib = None
_hip = None


class VrtError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__("%s failed (%d): %s" % (fn, code, msg))
        self.code = code


def load_library(path=LIB_PATH):
    """Loads libvrt.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libvrt.so not built (%s): run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return load_library()


def check(rc, fn="vrt call"):
    if rc != VRT_OK:
        raise VrtError(rc, fn, lib().vrt_error_string().decode())
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)


def device_count():
    n = I32(0)
    lib().vrt_device_count(C.byref(n))
    return n.value


# ---------------------------------------------------------------- HIP memory
def hip():
    global _hip
    if _hip is None:
        h = C.CDLL(HIP_LIB)
        h.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
        h.hipFree.argtypes = [P]
        h.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [P, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipEventCreate.argtypes = [C.POINTER(P)]
        h.hipEventRecord.argtypes = [P, P]
        h.hipEventSynchronize.argtypes = [P]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), P, P]
        h.hipEventDestroy.argtypes = [P]
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _hip = h
    return _hip


def hip_check(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(rc).decode()))


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """A hipMalloc allocation."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = P()
        hip_check(hip().hipMalloc(C.byref(self.ptr), max(1, self.nbytes)), "hipMalloc(%d)" % self.nbytes)

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        hip_check(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(P), arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, shape):
        """Copy to the host after all queued device work: hipMemcpy on the null
        stream does not wait for a context's non-blocking stream."""
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        hip_check(hip().hipMemcpy(out.ctypes.data_as(P), self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def fill(self, byte):
        hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        hip_check(hip().hipMemset(self.ptr, byte, self.nbytes), "hipMemset")
        hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ev = P()
        hip_check(hip().hipEventCreate(C.byref(self.ev)), "hipEventCreate")

    def record(self, stream):
        hip_check(hip().hipEventRecord(self.ev, stream), "hipEventRecord")

    def sync(self):
        hip_check(hip().hipEventSynchronize(self.ev), "hipEventSynchronize")

    def ms_since(self, start):
        t = C.c_float(0)
        hip_check(hip().hipEventElapsedTime(C.byref(t), start.ev, self.ev), "hipEventElapsedTime")
        return t.value

    def __del__(self):
        try:
            if self.ev:
                hip().hipEventDestroy(self.ev)
        except Exception:
            pass


# --------------------------------------------------------------- frame buffers
class Target:
    """Device frame buffers for one view: rgb (+ depth, + aux maps; steps=False
    leaves out the primary step map, which lets primary rays use the cutoff)."""

    def __init__(self, width, height, depth=True, aux=False, steps=True):
        self.width, self.height = int(width), int(height)
        n = self.width * self.height
        self.rgb = DeviceBuffer(3 * n)
        self.depth = DeviceBuffer(4 * n) if depth else None
        self.aux = aux
        if aux:
            self.hit = DeviceBuffer(8 * n)
            self.face = DeviceBuffer(n)
            self.eval = DeviceBuffer(4 * n)
            self.occ = DeviceBuffer(4 * n)
            self.steps = DeviceBuffer(4 * n) if steps else None

    def struct(self):
        t = vrt_target()
        t.width, t.height = self.width, self.height
        t.rgb = self.rgb.addr
        t.depth = self.depth.addr if self.depth else None
        if self.aux:
            t.hit, t.face = self.hit.addr, self.face.addr
            t.shadow_eval, t.shadow_occ = self.eval.addr, self.occ.addr
            t.steps = self.steps.addr if self.steps else None
        return t

    def download(self):
        h, w = self.height, self.width
        out = {"rgb": self.rgb.download(np.uint8, (h, w, 3))}
        if self.depth:
            out["depth"] = self.depth.download(np.float32, (h, w))
        if self.aux:
            out["hit"] = self.hit.download(np.int64, (h, w))
            out["face"] = self.face.download(np.uint8, (h, w))
            out["shadow_eval"] = self.eval.download(np.uint32, (h, w))
            out["shadow_occ"] = self.occ.download(np.uint32, (h, w))
            if self.steps:
                out["steps"] = self.steps.download(np.uint32, (h, w))
        return out


def download_target(t, aux=True, depth=True):
    """Downloads a vrt_target struct's device buffers (e.g. a server result)."""
    h, w = t.height, t.width
    n = w * h

    def get(ptr, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        hip_check(hip().hipMemcpy(out.ctypes.data_as(P), P(ptr), out.nbytes, D2H), "hipMemcpy D2H")
        return out

    out = {"rgb": get(t.rgb, np.uint8, (h, w, 3))}
    if depth and t.depth:
        out["depth"] = get(t.depth, np.float32, (h, w))
    if aux and t.hit:
        out["hit"] = get(t.hit, np.int64, (h, w))
        out["face"] = get(t.face, np.uint8, (h, w))
        out["shadow_eval"] = get(t.shadow_eval, np.uint32, (h, w))
        out["shadow_occ"] = get(t.shadow_occ, np.uint32, (h, w))
        out["steps"] = get(t.steps, np.uint32, (h, w))
    assert n >= 0
    return out


# ------------------------------------------------------------------- context
class Context:
    """A libvrt context: the grid resident on one device."""

    def __init__(self, dims, device=0):
        self.dims = tuple(int(d) for d in dims)
        self.h = P()
        call("vrt_create", C.byref(self.h), self.dims[0], self.dims[1], self.dims[2], device)
        self.shadow_cutoff = True

    def set_shadow_cutoff(self, enable):
        """Shadow rays above every solid voxel stop early (outputs unchanged;
        with it off the DDA-step totals equal the reference's)."""
        call("vrt_set_shadow_cutoff", self.h, 1 if enable else 0)
        self.shadow_cutoff = bool(enable)

    def grid_modified(self):
        call("vrt_grid_modified", self.h)

    def close(self):
        if self.h:
            lib().vrt_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().vrt_get_stream(self.h)

    def sync(self):
        call("vrt_sync", self.h)

    def upload(self, world):
        w = np.ascontiguousarray(world, dtype=np.int32)
        assert w.size == self.dims[0] * self.dims[1] * self.dims[2]
        call("vrt_upload", self.h, w.ctypes.data_as(P))

    def upload_region(self, world, lo, hi):
        w = np.ascontiguousarray(world, dtype=np.int32)
        call("vrt_upload_region", self.h, w.ctypes.data_as(P), *[int(v) for v in lo], *[int(v) for v in hi])

    def upload_region_f(self, world, start, end):
        """updatePartialGeometry with the reference's float corners."""
        w = np.ascontiguousarray(world, dtype=np.int32)
        s = (C.c_float * 3)(*[float(v) for v in start])
        e = (C.c_float * 3)(*[float(v) for v in end])
        call("vrt_upload_region_f", self.h, w.ctypes.data_as(P), s, e)

    def download(self):
        out = np.empty(self.dims[0] * self.dims[1] * self.dims[2], dtype=np.int32)
        call("vrt_download", self.h, out.ctypes.data_as(P))
        return out

    def generate(self, params):
        call("vrt_world_generate_device", self.h, C.byref(params))

    def build_depth_field(self):
        call("vrt_build_depth_field", self.h)

    def fix_depth_field_box(self, lo, hi):
        call("vrt_fix_depth_field_box", self.h, *[int(v) for v in lo], *[int(v) for v in hi])

    def remove_sphere(self, cx, cy, cz, r):
        call("vrt_remove_sphere", self.h, int(cx), int(cy), int(cz), int(r))

    def place_voxel(self, x, y, z, v):
        call("vrt_place_voxel", self.h, int(x), int(y), int(z), int(v))

    def destroy_voxel(self, x, y, z):
        call("vrt_destroy_voxel", self.h, int(x), int(y), int(z))

    def set_frame(self, frame):
        call("vrt_set_frame", self.h, C.byref(frame))

    def draw(self, target, frame=None):
        t = target.struct() if isinstance(target, Target) else target
        if frame is None:
            call("vrt_draw", self.h, C.byref(t))
        else:
            call("vrt_draw_frame", self.h, C.byref(frame), C.byref(t))

    def stats(self, reset=False):
        s = vrt_stats()
        call("vrt_read_stats", self.h, C.byref(s), 1 if reset else 0)
        return {"rays": s.rays, "dda_steps": s.dda_steps, "pixels": s.pixels, "draws": s.draws}

    def selftest_division(self, d_begin, d_count):
        bad = U64(0)
        call("vrt_selftest_division", self.h, int(d_begin), int(d_count), C.byref(bad))
        return bad.value

    def selftest_reciprocal(self, first, count):
        bad = U64(0)
        call("vrt_selftest_reciprocal", self.h, int(first), int(count), C.byref(bad))
        return bad.value

    def cast_rays(self, start, direction, dist, render_dist, steps=True):
        """castRay for each row; steps=False leaves out the iteration counts
        (the rays may then take the cutoff; the other outputs are the same)."""
        start = np.ascontiguousarray(start, dtype=np.float32).reshape(-1, 3)
        direction = np.ascontiguousarray(direction, dtype=np.float32).reshape(-1, 3)
        dist = np.ascontiguousarray(dist, dtype=np.int32).reshape(-1)
        n = start.shape[0]
        bs, bd, bl = DeviceBuffer(start.nbytes), DeviceBuffer(direction.nbytes), DeviceBuffer(dist.nbytes)
        bs.upload(start)
        bd.upload(direction)
        bl.upload(dist)
        bh, bp, bf = DeviceBuffer(8 * n), DeviceBuffer(12 * n), DeviceBuffer(n)
        bt = DeviceBuffer(4 * n) if steps else None
        call("vrt_cast_rays", self.h, n, bs.addr, bd.addr, bl.addr, int(render_dist), bh.addr, bp.addr,
             bf.addr, bt.addr if bt else None)
        self.sync()
        out = {"hit": bh.download(np.int64, (n,)), "hit_pos": bp.download(np.float32, (n, 3)),
               "face": bf.download(np.uint8, (n,))}
        if bt:
            out["steps"] = bt.download(np.uint32, (n,))
        return out


# ------------------------------------------------------------- frame helpers
def frame_defaults(grid_width, width, height):
    f = vrt_frame()
    call("vrt_frame_defaults", C.byref(f), int(grid_width), int(width), int(height))
    return f


def camera_matrix(pitch, yaw):
    out = (C.c_float * 16)()
    call("vrt_camera_matrix", F32(pitch), F32(yaw), out)
    return np.frombuffer(out, dtype=np.float32).copy()


def light_orbit(deg, start):
    s = (C.c_float * 3)(*start)
    out = (C.c_float * 3)()
    call("vrt_light_orbit", F32(deg), s, out)
    return np.frombuffer(out, dtype=np.float32).copy()


def mat4_rotate(m, angle, axis):
    mm = (C.c_float * 16)(*np.asarray(m, dtype=np.float32).ravel())
    ax = (C.c_float * 3)(*axis)
    out = (C.c_float * 16)()
    call("vrt_mat4_rotate", mm, F32(angle), ax, out)
    return np.frombuffer(out, dtype=np.float32).copy()


def mat4_mul(a, b):
    aa = (C.c_float * 16)(*np.asarray(a, dtype=np.float32).ravel())
    bb = (C.c_float * 16)(*np.asarray(b, dtype=np.float32).ravel())
    out = (C.c_float * 16)()
    call("vrt_mat4_mul", aa, bb, out)
    return np.frombuffer(out, dtype=np.float32).copy()


def depth_table():
    n = I32(0)
    dist = (C.c_float * 2048)()
    off = (C.c_int32 * (3 * 2048))()
    call("vrt_depth_table", dist, off, 2048, C.byref(n))
    k = n.value
    return np.frombuffer(dist, dtype=np.float32)[:k].copy(), np.frombuffer(off, dtype=np.int32)[:3 * k].reshape(k, 3).copy()


# --------------------------------------------------------- synthetic scenes
def world_params(dims, seed, caves=0, boxes=0, height_min=16, height_max=96, octaves=4, base_cell=64,
                 cave_r=(3, 10), box_size=(1, 12), jitter=8):
    p = vrt_world_params()
    p.size[0], p.size[1], p.size[2] = [int(d) for d in dims]
    p.seed = int(seed)
    p.height_min, p.height_max = height_min, height_max
    p.octaves, p.base_cell = octaves, base_cell
    p.caves, p.cave_r_min, p.cave_r_max = caves, cave_r[0], cave_r[1]
    p.boxes, p.box_min, p.box_max = boxes, box_size[0], box_size[1]
    p.jitter = jitter
    return p


def generate_world(params, threads=0):
    n = params.size[0] * params.size[1] * params.size[2]
    out = np.empty(n, dtype=np.int32)
    call("vrt_world_generate", C.byref(params), out.ctypes.data_as(P), int(threads))
    return out


def reference_world(dims=(512, 96, 512)):
    """initVoxels' world (src/level.cpp:82-138), host array."""
    out = np.empty(dims[0] * dims[1] * dims[2], dtype=np.int32)
    call("vrt_world_reference", int(dims[0]), int(dims[1]), int(dims[2]), out.ctypes.data_as(P))
    return out


def surface(params, x, z):
    h = I32(0)
    call("vrt_world_surface", C.byref(params), int(x), int(z), C.byref(h))
    return h.value


# World parameters of the benchmark configurations: value-noise heights
# 16-96, 4 octaves, colour jitter +-8, caves of radius 3-10 (32 per 128x128
# of area, or the count giving ~35% empty cells below the surface for the
# large world), plus 2000 boxes of 1-12 voxels for the 1080p drop-in world.
CONFIGS = {
    "terrain128": dict(dims=(128, 128, 128), caves=32, boxes=0, size=(256, 256), lights=0, render_dist=384),
    "frame1080": dict(dims=(512, 256, 512), caves=512, boxes=2000, size=(1920, 1080), lights=16, render_dist=384),
    "large4k": dict(dims=(2048, 512, 2048), caves=70000, boxes=0, size=(3840, 2160), lights=16, render_dist=4096),
    "views64": dict(dims=(1024, 256, 1024), caves=2048, boxes=0, size=None, lights=None, render_dist=384),
    "served2g": dict(dims=(1024, 512, 1024), caves=2048, boxes=0, size=None, lights=None, render_dist=384),
    # small scenes for smoke and quick parity runs
    "tiny": dict(dims=(96, 64, 96), caves=8, boxes=40, size=(160, 120), lights=4, render_dist=384, heights=(8, 40)),
}
MIXED_RESOLUTIONS = [(320, 180), (640, 360), (1280, 720), (1920, 1080), (3840, 2160)]


class Scene:
    def __init__(self, name, dims, params, world, frame, width, height):
        self.name, self.dims, self.params = name, dims, params
        self.world, self.frame, self.width, self.height = world, frame, width, height


def is_open(world, dims, x, y, z):
    X, Y, Z = dims
    xi, yi, zi = int(x), int(y), int(z)
    if not (0 <= xi < X and 0 <= yi < Y and 0 <= zi < Z):
        return False
    return world[xi + X * yi + X * Y * zi] < 0


def make_frame(params, dims, world, rng, width, height, n_lights, render_dist, pitch_range=(0.15, 0.5),
               area=(0.25, 0.75), cam_height=(8, 40), light_spread=None, depth_view=0):
    """A seeded camera above the terrain, the sun on the reference's orbit and
    n_lights point lights in open air above the terrain."""
    X, Y, Z = dims
    f = frame_defaults(X, width, height)
    for _ in range(1000):
        cx = rng.uniform(area[0] * X, area[1] * X)
        cz = rng.uniform(area[0] * Z, area[1] * Z)
        cy = min(Y - 1.5, surface(params, int(cx), int(cz)) + rng.uniform(*cam_height))
        if is_open(world, dims, cx, cy, cz):
            break
    f.cam_pos[:] = [cx, cy, cz]
    # yaw toward the middle of the world (+-90 degrees): the camera looks along
    # rotY(yaw) * (0, 0, 1) = (sin yaw, 0, cos yaw)
    yaw = math.atan2(X / 2.0 - cx, Z / 2.0 - cz) + rng.uniform(-0.5 * math.pi, 0.5 * math.pi)
    m = camera_matrix(rng.uniform(*pitch_range), yaw)
    f.rotate_matrix[:] = m.tolist()
    f.light_pos[:] = light_orbit(rng.uniform(-60.0, -30.0), (X / 2.0, 3.0 * X, Z / 2.0)).tolist()
    f.aspect_ratio = width / height
    f.focal = FOV70_FOCAL
    f.render_dist = render_dist
    f.view_depth_field = depth_view
    for i in range(n_lights):
        for _ in range(1000):
            if light_spread is None:
                lx, lz = rng.uniform(0, X - 1), rng.uniform(0, Z - 1)
            else:
                lx = float(np.clip(cx + rng.uniform(-light_spread, light_spread), 0, X - 1))
                lz = float(np.clip(cz + rng.uniform(-light_spread, light_spread), 0, Z - 1))
            ly = min(Y - 1.5, surface(params, int(lx), int(lz)) + rng.uniform(2.0, 20.0))
            if is_open(world, dims, lx, ly, lz):
                break
        L = f.lights[i]
        L.pos[:] = [lx, ly, lz]
        L.rgb[:] = [rng.uniform(0.5, 2.0) for _ in range(3)]
        L.radius = rng.uniform(16.0, 64.0)
    return f


def make_scene(name, seed=1, world=None, depth_view=0):
    """A benchmark configuration's world (host, without depth field) and one frame."""
    cfg = CONFIGS[name]
    dims = cfg["dims"]
    hmin, hmax = cfg.get("heights", (16, 96))
    params = world_params(dims, seed, caves=cfg["caves"], boxes=cfg["boxes"], height_min=hmin, height_max=hmax)
    if world is None:
        world = generate_world(params)
    rng = np.random.default_rng(seed)
    if cfg["size"] is None:
        w, h = MIXED_RESOLUTIONS[-1]
        n_lights = 16
    else:
        (w, h), n_lights = cfg["size"], cfg["lights"]
    pitch = (0.035, 0.175) if name == "large4k" else (0.15, 0.5)
    frame = make_frame(params, dims, world, rng, w, h, n_lights, cfg["render_dist"], pitch_range=pitch,
                       depth_view=depth_view)
    return Scene(name, dims, params, world, frame, w, h)


def make_views(params, dims, world, seed, n_views, resolutions=MIXED_RESOLUTIONS, render_dist=384,
               max_lights=16):
    """n_views seeded views: resolution uniform from `resolutions`, pose random
    over the world, light count uniform in [0, max_lights] around the camera."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_views):
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        nl = int(rng.integers(0, max_lights + 1))
        f = make_frame(params, dims, world, rng, w, h, nl, render_dist, pitch_range=(0.1, 0.6), area=(0.05, 0.95),
                       light_spread=96.0)
        views.append((f, w, h))
    return views


def render_on_gpu(scene, aux=True, depth=True, ctx=None, shadow_cutoff=True):
    """Uploads the scene's world, builds its depth field on the device and draws
    one frame; returns the host copies of the outputs plus ray/step counts
    (and whether the shadow cutoff was on)."""
    own = ctx is None
    if own:
        ctx = Context(scene.dims)
        ctx.upload(scene.world)
        ctx.build_depth_field()
    ctx.set_shadow_cutoff(shadow_cutoff)
    tgt = Target(scene.width, scene.height, depth=depth, aux=aux)
    ctx.stats(reset=True)
    ctx.draw(tgt, frame=scene.frame)
    st = ctx.stats(reset=True)
    out = tgt.download()
    out.update(st)
    out["shadow_cutoff"] = bool(shadow_cutoff)
    if own:
        ctx.close()
    return out
