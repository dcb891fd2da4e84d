// vrt_api.hip — C-ABI entry points of libvrt: contexts, uploads, depth field,
// edits, frame constants and the plain draw.  Each function cites the
// reference interface it stands in for (src/render.hpp:30-41).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

#include "vrt_internal.h"

namespace vrt {

namespace {
thread_local std::string g_error;
}

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    // consume the runtime's last-error state so that the launch checks of
    // later calls do not report this failure again
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory) return fail(VRT_ERR_NOMEM, std::string(what) + ": " + hipGetErrorString(e));
    return fail(VRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// computeDepthIndices, src/render.cpp:66-100: offsets with x^2+y^2+z^2 <= 49
// in z, y, x order; dist = -sqrt(shrunk^2) where each component moves one step
// toward zero; sqrt of the int is the double overload, stored as float.
namespace {
struct TableEntry {
    float dist;
    int32_t x, y, z;
};
std::vector<TableEntry> reference_table() {
    const int R = VRT_DEPTH_FIELD_RADIUS;
    std::vector<TableEntry> t;
    for (int z = -R; z <= R; z++)
        for (int y = -R; y <= R; y++)
            for (int x = -R; x <= R; x++) {
                if (x * x + y * y + z * z > R * R) continue;
                const int sx = x - (x > 0) + (x < 0), sy = y - (y > 0) + (y < 0), sz = z - (z > 0) + (z < 0);
                const float d = (float)(-std::sqrt((double)(sx * sx + sy * sy + sz * sz)));
                if (d <= (float)R) t.push_back({d, x, y, z});
            }
    return t;
}
}  // namespace

std::vector<float> depth_table_dist() {
    std::vector<float> out;
    for (auto& e : reference_table()) out.push_back(e.dist);
    return out;
}

std::vector<int32_t> depth_table_offsets() {
    std::vector<int32_t> out;
    for (auto& e : reference_table()) out.insert(out.end(), {e.x, e.y, e.z});
    return out;
}

// Largest float t with sqrtf(t) <= r (IEEE sqrt, correctly rounded like the
// device's); NaN for NaN r, -inf when no t qualifies (r < 0).
float sqrt_le_limit(float r) {
    if (r != r) return r;
    if (r < 0.0f) return -INFINITY;
    if (r == INFINITY) return INFINITY;
    float t = r * r;
    while (t > 0.0f && !(std::sqrt(t) <= r)) t = std::nextafter(t, -INFINITY);
    for (;;) {
        const float up = std::nextafter(t, INFINITY);
        if (!(std::sqrt(up) <= r)) break;
        t = up;
    }
    return t;
}

// Largest float a >= 0 with a * a <= t (one IEEE product, as on the device);
// -inf for NaN or negative t, +inf for t = +inf.
float square_le_limit(float t) {
    if (!(t >= 0.0f)) return -INFINITY;
    if (t == INFINITY) return INFINITY;
    float a = std::sqrt(t);
    while (a > 0.0f && !(a * a <= t)) a = std::nextafter(a, -INFINITY);
    for (;;) {
        const float up = std::nextafter(a, INFINITY);
        if (!(up * up <= t)) break;
        a = up;
    }
    return a;
}

FrameDesc make_frame_desc(const vrt_frame& f, const vrt_target& target) {
    const int32_t width = target.width, height = target.height;
    FrameDesc d;
    d.f = f;
    d.render_lim = (float)f.render_dist;
    d.light_valid = 0;
    d.inv_w = width > 0 ? 1.0f / (float)width : 0.0f;
    d.inv_h = height > 0 ? 1.0f / (float)height : 0.0f;
    d.primary_cutoff = (f.view_depth_field != 1 && target.steps == nullptr) ? 1 : 0;
    d.row_rot = 0;
    for (int i = 0; i < VRT_MAX_LOCAL_LIGHTS; i++) {
        const vrt_local_light& L = f.lights[i];
        d.light_dd_max[i] = sqrt_le_limit(L.radius);
        d.light_axis_max[i] = square_le_limit(d.light_dd_max[i]);
        d.light_rinv[i] = 1.0f / L.radius;
        if (L.pos[0] >= 0.0f && L.pos[1] >= 0.0f && L.pos[2] >= 0.0f) d.light_valid |= 1u << i;
    }
    return d;
}

// First tile row of a plain draw.  A workgroup's time is set by its longest
// walk, and the hardware deals workgroups in order, so the frame ends when
// the last-dealt rows' walks end: rows of moderate cost belong last.  With
// primary rays stopping above the grid's solid top and the camera below that
// top, a row whose rays all rise (at the left edge, centre and right edge of
// the frame) steeply enough to clear the top within kCheapWalk cells is such
// a row; those at the top of the frame (the sky of a camera without roll) are
// dealt after all the others (-4% to -6% on a 1080p frame whose camera is 5
// cells under the top).  A camera above the top keeps row-major order: its
// sky rays end after one jump, and dealing those near-empty workgroups last
// made a 4K frame 30% slower.  Only the order changes, never a result.
int32_t plain_row_rotation(const vrt_frame& f, const vrt_target& t, int32_t top, bool primary_cutoff) {
    if (!primary_cutoff || top < 0 || t.height <= 0) return 0;
    constexpr double kCheapWalk = 96.0;
    const int32_t tiles_y = (t.height + 7) / 8;
    const double rise = (double)top + 1.0 - (double)f.cam_pos[1];
    if (!(rise > 0.0)) return 0;
    const float* M = f.rotate_matrix;
    int32_t r = 0;
    for (; r < tiles_y; r++) {
        const int32_t py = std::min(8 * r + 4, t.height - 1);
        const double vy = (2.0 * (t.height - 1 - py) + 1.0) / t.height - 1.0;
        double min_dy = INFINITY;
        for (double vx : {-1.0, 0.0, 1.0}) {
            const double ux = vx * f.aspect_ratio, uy = vy, uz = f.focal;
            const double n = std::sqrt(ux * ux + uy * uy + uz * uz);
            min_dy = std::min(min_dy, (M[1] * ux + M[5] * uy + M[9] * uz) / n);
        }
        const bool cheap = min_dy > 0.0 && rise < kCheapWalk * min_dy;
        if (!cheap) break;
    }
    return r == tiles_y ? 0 : r;
}

namespace {
// True when work on `s` is being captured into a graph (or the state cannot
// be read: then the caller takes the path that is right under capture).
bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

// Adds the plain draws' pending partial counts into the counters.
int fold_parts(vrt_ctx* c) {
    if (c->parts_used <= 0) return VRT_OK;
    if (int rc = launch_fold_parts(c->parts, c->parts_used, c->counters, c->stream)) return rc;
    c->parts_used = 0;
    return VRT_OK;
}
}  // namespace

int check_frame(const vrt_frame& f) {
    if (f.render_dist < 0 || f.render_dist > (1 << 24))
        return fail(VRT_ERR_INVALID, "render_dist must be in [0, 2^24]");
    return VRT_OK;
}

int check_target(const vrt_target& t) {
    if (t.width <= 0 || t.height <= 0) return fail(VRT_ERR_INVALID, "target width and height must be positive");
    if ((int64_t)t.width * t.height > ((int64_t)1 << 31)) return fail(VRT_ERR_INVALID, "target too large");
    if (!t.rgb) return fail(VRT_ERR_INVALID, "target rgb buffer is required");
    return VRT_OK;
}

namespace {

int check_ctx(vrt_ctx* c) {
    if (!c || !c->grid.vox) return fail(VRT_ERR_INVALID, "invalid context");
    VRT_HIP(hipSetDevice(c->device));
    return VRT_OK;
}

// Depth table on the device, sorted by decreasing distance, entries at -6
// (never counted: the test at src/render.cpp:240 is strict) dropped.
int upload_table(vrt_ctx* c) {
    auto t = reference_table();
    std::vector<int> order(t.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t[a].dist > t[b].dist; });
    std::vector<int32_t> packed;
    std::vector<float> dist;
    int32_t n_near = 0;
    for (int i : order) {
        if (!(t[i].dist > -6.0f)) continue;
        packed.push_back((t[i].x + 8) | ((t[i].y + 8) << 5) | ((t[i].z + 8) << 10));
        dist.push_back(t[i].dist);
        if (t[i].dist > -2.0f) n_near++;
    }
    VRT_HIP(hipMalloc(&c->table_packed, packed.size() * sizeof(int32_t)));
    VRT_HIP(hipMalloc(&c->table_dist, dist.size() * sizeof(float)));
    VRT_HIP(hipMemcpy(c->table_packed, packed.data(), packed.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    VRT_HIP(hipMemcpy(c->table_dist, dist.data(), dist.size() * sizeof(float), hipMemcpyHostToDevice));
    c->table.packed = c->table_packed;
    c->table.dist = c->table_dist;
    c->table.n = (int32_t)packed.size();
    c->table.n_near = n_near;
    // rows: the entries of one distance and one (dy, dz), their dx as a mask
    std::vector<int32_t> rows;
    std::vector<float> rows_dist;
    int32_t n_rows_near = 0;
    for (size_t a = 0; a < packed.size();) {
        size_t e = a;
        while (e < packed.size() && dist[e] == dist[a]) e++;  // one distance level
        for (size_t i = a; i < e; i++) {
            const int32_t dy = (packed[i] >> 5) & 31, dz = (packed[i] >> 10) & 31;
            bool seen = false;
            for (size_t j = a; j < i; j++) seen |= ((packed[j] >> 5) & 31) == dy && ((packed[j] >> 10) & 31) == dz;
            if (seen) continue;
            int32_t mask = 0;
            for (size_t j = i; j < e; j++)
                if (((packed[j] >> 5) & 31) == dy && ((packed[j] >> 10) & 31) == dz) mask |= 1 << ((packed[j] & 31) - 1);
            rows.push_back(mask | (dy << 16) | (dz << 21));
            rows_dist.push_back(dist[a]);
            if (dist[a] > -2.0f) n_rows_near++;
        }
        a = e;
    }
    VRT_HIP(hipMalloc(&c->table_rows, rows.size() * sizeof(int32_t)));
    VRT_HIP(hipMalloc(&c->table_rows_dist, rows_dist.size() * sizeof(float)));
    VRT_HIP(hipMemcpy(c->table_rows, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    VRT_HIP(hipMemcpy(c->table_rows_dist, rows_dist.data(), rows_dist.size() * sizeof(float), hipMemcpyHostToDevice));
    c->table.rows = c->table_rows;
    c->table.rows_dist = c->table_rows_dist;
    c->table.n_rows = (int32_t)rows.size();
    c->table.n_rows_near = n_rows_near;
    return VRT_OK;
}

void release(vrt_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    if (c->grid.vox) (void)hipFree(c->grid.vox);
    if (c->rgrid.vox) (void)hipFree(c->rgrid.vox);
    if (c->grid.top) (void)hipFree(c->grid.top);
    if (c->counters) (void)hipFree(c->counters);
    if (c->parts) (void)hipFree(c->parts);
    if (c->table_packed) (void)hipFree(c->table_packed);
    if (c->table_dist) (void)hipFree(c->table_dist);
    if (c->table_rows) (void)hipFree(c->table_rows);
    if (c->table_rows_dist) (void)hipFree(c->table_rows_dist);
    if (c->bricks) (void)hipFree(c->bricks);
    if (c->solid_bits) (void)hipFree(c->solid_bits);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

// glm's 4x4 helpers on column-major float[16] (src/glm/ext/matrix_transform.inl:18-46,
// src/glm/detail/type_mat4x4.inl:560-571,630-648, func_geometric.inl:88).
void mat4_rotate(const float* m, float angle, const float* v, float* out) {
    const float a = angle;
    const float c = std::cos(a);
    const float s = std::sin(a);
    const float dd = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(dd);
    const float axis[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    const float temp[3] = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    float R[3][3];
    R[0][0] = c + temp[0] * axis[0];
    R[0][1] = temp[0] * axis[1] + s * axis[2];
    R[0][2] = temp[0] * axis[2] - s * axis[1];
    R[1][0] = temp[1] * axis[0] - s * axis[2];
    R[1][1] = c + temp[1] * axis[1];
    R[1][2] = temp[1] * axis[2] + s * axis[0];
    R[2][0] = temp[2] * axis[0] + s * axis[1];
    R[2][1] = temp[2] * axis[1] - s * axis[0];
    R[2][2] = c + temp[2] * axis[2];
    float res[16];
    for (int j = 0; j < 3; j++)
        for (int r = 0; r < 4; r++) res[4 * j + r] = (m[r] * R[j][0] + m[4 + r] * R[j][1]) + m[8 + r] * R[j][2];
    for (int r = 0; r < 4; r++) res[12 + r] = m[12 + r];
    std::memcpy(out, res, sizeof(res));
}

void mat4_mul(const float* a, const float* b, float* out) {
    float res[16];
    for (int j = 0; j < 4; j++)
        for (int r = 0; r < 4; r++)
            res[4 * j + r] = ((a[r] * b[4 * j] + a[4 + r] * b[4 * j + 1]) + a[8 + r] * b[4 * j + 2]) +
                             a[12 + r] * b[4 * j + 3];
    std::memcpy(out, res, sizeof(res));
}

const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

}  // namespace

// Reads *grid.top into top_host (waits for the context's stream).
int read_top(vrt_ctx* c) {
    VRT_HIP(hipMemcpyAsync(&c->top_host, c->grid.top, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    VRT_HIP(hipStreamSynchronize(c->stream));
    return VRT_OK;
}

int solid_top_full(vrt_ctx* c) {
    return launch_solid_top(c->grid, 0, 0, 0, c->grid.X, c->grid.Y, c->grid.Z, true, c->stream);
}

int refresh_render(vrt_ctx* c, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1, int64_t z1) {
    return refresh_render_on(c, c->stream, x0, y0, z0, x1, y1, z1);
}

int refresh_render_on(vrt_ctx* c, hipStream_t s, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1,
                      int64_t z1) {
    return refresh_copy_on(c, c->rgrid, s, x0, y0, z0, x1, y1, z1);
}

int refresh_copy_on(vrt_ctx* c, const GridDesc& dst, hipStream_t s, int64_t x0, int64_t y0, int64_t z0, int64_t x1,
                    int64_t y1, int64_t z1) {
    x0 = std::max<int64_t>(x0, 0);
    y0 = std::max<int64_t>(y0, 0);
    z0 = std::max<int64_t>(z0, 0);
    x1 = std::min<int64_t>(x1, c->grid.X);
    y1 = std::min<int64_t>(y1, c->grid.Y);
    z1 = std::min<int64_t>(z1, c->grid.Z);
    if (x1 <= x0 || y1 <= y0 || z1 <= z0) return VRT_OK;
    return launch_copy_box(c->grid, dst, x0, y0, z0, x1 - x0, y1 - y0, z1 - z0, s);
}

int remove_sphere_grid(vrt_ctx* c, int32_t cx, int32_t cy, int32_t cz, int32_t radius, hipStream_t s) {
    if (int rc = launch_carve_sphere(c->grid, cx, cy, cz, radius, s)) return rc;
    const int32_t R = radius + (VRT_DEPTH_FIELD_RADIUS >> 1);
    if (R > 0) {
        const int64_t x0 = (int64_t)cx - R, x1 = (int64_t)cx + R, y0 = (int64_t)cy - R, y1 = (int64_t)cy + R,
                      z0 = (int64_t)cz - R, z1 = (int64_t)cz + R;
        auto c32 = [](int64_t v) { return (int32_t)std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, v)); };
        if (int rc = launch_fix_depth_box(c->grid, c->table, c32(x0), c32(y0), c32(z0), c32(x1), c32(y1), c32(z1),
                                          cx, cy, cz, R, s))
            return rc;
    }
    return VRT_OK;
}

int remove_sphere_refresh(vrt_ctx* c, int32_t cx, int32_t cy, int32_t cz, int32_t radius, hipStream_t s) {
    return remove_sphere_refresh_into(c, c->rgrid, cx, cy, cz, radius, s);
}

int remove_sphere_refresh_into(vrt_ctx* c, const GridDesc& dst, int32_t cx, int32_t cy, int32_t cz, int32_t radius,
                               hipStream_t s) {
    // the carve covers [c - |r|, c + |r|) at most, the repair [c - R, c + R)
    const int64_t R = (int64_t)radius + (VRT_DEPTH_FIELD_RADIUS >> 1);
    const int64_t M = std::max<int64_t>(radius < 0 ? -(int64_t)radius : (int64_t)radius, R);
    return refresh_copy_on(c, dst, s, (int64_t)cx - M, (int64_t)cy - M, (int64_t)cz - M, (int64_t)cx + M,
                           (int64_t)cy + M, (int64_t)cz + M);
}

int refresh_render_all(vrt_ctx* c) { return refresh_render(c, 0, 0, 0, c->grid.X, c->grid.Y, c->grid.Z); }

}  // namespace vrt

using namespace vrt;

extern "C" {

const char* vrt_error_string(void) { return g_error.c_str(); }

const char* vrt_version(void) { return "vrt 0.1 (gfx950)"; }

int vrt_device_count(int32_t* count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return n > 0 ? VRT_OK : fail(VRT_ERR_NODEVICE, "no HIP device visible");
}

// initRender, src/render.cpp:313-372: grid allocated and filled with -1
// (:349-351).
int vrt_create(vrt_ctx** out, int64_t X, int64_t Y, int64_t Z, int32_t device) {
    if (!out) return fail(VRT_ERR_INVALID, "null out");
    *out = nullptr;
    if (X <= 0 || Y <= 0 || Z <= 0 || X > 0x7fffffff || Y > 0x7fffffff || Z > 0x7fffffff)
        return fail(VRT_ERR_INVALID, "grid dimensions must be in [1, 2^31)");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(VRT_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(VRT_ERR_INVALID, "device out of range");
    vrt_ctx* c = new vrt_ctx();
    c->device = device;
    const uint64_t count = (uint64_t)X * (uint64_t)Y * (uint64_t)Z;
    auto bail = [&](int rc) {
        release(c);
        return rc;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return bail(hip_fail(e, "hipSetDevice"));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return bail(hip_fail(e, "hipGetDeviceProperties"));
    c->num_cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    c->stream = c->own_stream;
    if (hipMalloc(&c->grid.vox, count * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        c->grid.vox = nullptr;
        return bail(fail(VRT_ERR_NOMEM, "cannot allocate the voxel grid"));
    }
    c->grid.X = (uint32_t)X;
    c->grid.Y = (uint32_t)Y;
    c->grid.Z = (uint32_t)Z;
    c->grid.XY = (uint64_t)X * (uint64_t)Y;
    c->grid.count = count;
    c->grid.SX = (uint32_t)X;
    c->grid.SXY = c->grid.XY;
    c->grid.origin = 0;
    {
        const uint64_t PX = (uint64_t)X + 2 * kGuard, PY = (uint64_t)Y + 2 * kGuard, PZ = (uint64_t)Z + 2 * kGuard;
        GridDesc& r = c->rgrid;
        r = c->grid;
        r.SX = (uint32_t)PX;
        r.SXY = PX * PY;
        r.count = PX * PY * PZ;
        r.origin = (uint64_t)kGuard * (1 + PX + PX * PY);
        r.vox = nullptr;
        if (hipMalloc(&r.vox, r.count * sizeof(int32_t)) != hipSuccess) {
            (void)hipGetLastError();
            r.vox = nullptr;
            return bail(fail(VRT_ERR_NOMEM, "cannot allocate the render copy of the voxel grid"));
        }
    }
    if ((e = hipMalloc(&c->grid.top, sizeof(int32_t))) != hipSuccess) return bail(hip_fail(e, "hipMalloc top"));
    const size_t counter_bytes = (size_t)kCounterSlots * kCounterStride * sizeof(unsigned long long);
    if ((e = hipMalloc(&c->counters, counter_bytes)) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc counters"));
    if ((e = hipMalloc(&c->parts, (size_t)kPartsCap * sizeof(unsigned long long))) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc partial counts"));
    const uint64_t nbricks = (uint64_t)((X + 7) / 8) * ((Y + 7) / 8) * ((Z + 7) / 8);
    if ((e = hipMalloc(&c->bricks, nbricks)) != hipSuccess) return bail(hip_fail(e, "hipMalloc bricks"));
    int rc = upload_table(c);
    if (rc) return bail(rc);
    if ((e = hipMemsetAsync(c->counters, 0, counter_bytes, c->stream)) != hipSuccess)
        return bail(hip_fail(e, "hipMemsetAsync"));
    rc = launch_fill(c->grid.vox, count, -1, c->stream);
    if (!rc) rc = launch_fill(c->rgrid.vox, c->rgrid.count, kGuardValue, c->stream);
    if (!rc) rc = refresh_render_all(c);
    if (rc) return bail(rc);
    rc = launch_solid_top(c->grid, 0, 0, 0, 0, 0, 0, true, c->stream);  // no solid voxel yet: -1
    if (rc) return bail(rc);
    c->top_host = -1;
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(hip_fail(e, "hipStreamSynchronize"));
    vrt_frame_defaults(&c->frame, X, 800, 600);
    *out = c;
    return VRT_OK;
}

int vrt_destroy(vrt_ctx* ctx) {
    if (!ctx) return fail(VRT_ERR_INVALID, "invalid context");
    release(ctx);
    return VRT_OK;
}

int vrt_set_stream(vrt_ctx* ctx, void* stream) {
    if (int rc = check_ctx(ctx)) return rc;
    const hipStream_t next = stream ? (hipStream_t)stream : ctx->own_stream;
    if (next == ctx->stream) return VRT_OK;
    // work queued on the old stream (uploads, edits, draws whose counts are
    // not folded yet) stays ahead of everything queued on the new one
    hipEvent_t e = nullptr;
    VRT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipError_t err = hipEventRecord(e, ctx->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(next, e, 0);
    (void)hipEventDestroy(e);
    if (err != hipSuccess) return hip_fail(err, "ordering the new stream after the old one");
    ctx->stream = next;
    return VRT_OK;
}

void* vrt_get_stream(vrt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int vrt_sync(vrt_ctx* ctx) {
    if (int rc = check_ctx(ctx)) return rc;
    VRT_HIP(hipStreamSynchronize(ctx->stream));
    return VRT_OK;
}

int vrt_grid_device_ptr(vrt_ctx* ctx, int32_t** out) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!out) return fail(VRT_ERR_INVALID, "null out");
    *out = ctx->grid.vox;
    return VRT_OK;
}

int vrt_grid_dims(vrt_ctx* ctx, int64_t dims[3]) {
    if (int rc = check_ctx(ctx)) return rc;
    dims[0] = ctx->grid.X;
    dims[1] = ctx->grid.Y;
    dims[2] = ctx->grid.Z;
    return VRT_OK;
}

// updateGeometry, src/render.cpp:199-202.
int vrt_upload(vrt_ctx* ctx, const int32_t* host_voxels) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!host_voxels) return fail(VRT_ERR_INVALID, "null voxels");
    VRT_HIP(hipMemcpyAsync(ctx->grid.vox, host_voxels, ctx->grid.count * sizeof(int32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    if (int rc = solid_top_full(ctx)) return rc;
    if (int rc = refresh_render_all(ctx)) return rc;
    return read_top(ctx);
}

int vrt_upload_device(vrt_ctx* ctx, const int32_t* device_voxels) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!device_voxels) return fail(VRT_ERR_INVALID, "null voxels");
    if (int rc = launch_copy(ctx->grid.vox, device_voxels, ctx->grid.count, ctx->stream)) return rc;
    if (int rc = solid_top_full(ctx)) return rc;
    if (int rc = refresh_render_all(ctx)) return rc;
    return read_top(ctx);
}

int vrt_grid_modified(vrt_ctx* ctx) {
    if (int rc = check_ctx(ctx)) return rc;
    if (int rc = solid_top_full(ctx)) return rc;
    if (int rc = refresh_render_all(ctx)) return rc;
    return read_top(ctx);
}

int vrt_set_shadow_cutoff(vrt_ctx* ctx, int32_t enable) {
    if (int rc = check_ctx(ctx)) return rc;
    ctx->shadow_cutoff = enable != 0;
    return VRT_OK;
}

// updatePartialGeometry, src/render.cpp:204-223: corners swapped when the
// start voxel's index exceeds the end's, then one x-row of x1-x0+1 words per
// (y, z) with y in [y0, y1), z in [z0, z1); rows clipped to the grid.
int vrt_upload_region(vrt_ctx* ctx, const int32_t* host_voxels, int32_t x0, int32_t y0, int32_t z0, int32_t x1,
                      int32_t y1, int32_t z1) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!host_voxels) return fail(VRT_ERR_INVALID, "null voxels");
    const GridDesc& g = ctx->grid;
    auto index = [&](int64_t x, int64_t y, int64_t z) -> int64_t {
        if (x >= 0 && y >= 0 && z >= 0 && x < g.X && y < g.Y && z < g.Z) return x + (int64_t)g.X * y + (int64_t)g.XY * z;
        return -1;
    };
    if (index(x0, y0, z0) > index(x1, y1, z1)) {
        std::swap(x0, x1);
        std::swap(y0, y1);
        std::swap(z0, z1);
    }
    // every row starts at getVoxelIndex(x0, y, z), which is -1 (the row is
    // skipped) for x0 outside the grid
    if (x0 < 0 || (int64_t)x0 >= (int64_t)g.X) return VRT_OK;
    const int64_t xs = x0;
    const int64_t xe = std::min<int64_t>((int64_t)g.X - 1, x1);  // inclusive end of the row
    if (xe < xs) return VRT_OK;
    const size_t row_bytes = (size_t)(xe - xs + 1) * sizeof(int32_t);
    const int64_t ys = std::max<int64_t>(0, y0), ye = std::min<int64_t>(g.Y, y1);
    const int64_t zs = std::max<int64_t>(0, z0), ze = std::min<int64_t>(g.Z, z1);
    if (ye <= ys || ze <= zs) return VRT_OK;
    // One strided 3D copy instead of (ye-ys)*(ze-zs) row copies.
    hipMemcpy3DParms p = {};
    p.srcPtr = make_hipPitchedPtr((void*)host_voxels, g.X * sizeof(int32_t), g.X, g.Y);
    p.dstPtr = make_hipPitchedPtr(ctx->grid.vox, g.X * sizeof(int32_t), g.X, g.Y);
    p.srcPos = make_hipPos(xs * sizeof(int32_t), ys, zs);
    p.dstPos = make_hipPos(xs * sizeof(int32_t), ys, zs);
    p.extent = make_hipExtent(row_bytes, ye - ys, ze - zs);
    p.kind = hipMemcpyHostToDevice;
    VRT_HIP(hipMemcpy3DAsync(&p, ctx->stream));
    // the region can only have gained solid voxels above the old bound
    if (int rc = launch_solid_top(g, xs, ys, zs, xe + 1, ye, ze, false, ctx->stream)) return rc;
    if (int rc = refresh_render(ctx, xs, ys, zs, xe + 1, ye, ze)) return rc;
    return read_top(ctx);
}

// updatePartialGeometry, src/render.cpp:204-223, with its float corners and
// float loops as written: corners swapped when getVoxelIndex of the truncated
// start exceeds the end's; xLength = int(end.x - start.x) + 1; for float i
// from start.z while i < end.z (i++), for float j from start.y while j <
// end.y (j++): the xLength words from getVoxelIndex(int(start.x), int(j),
// int(i)), when that is a cell, copied as one run (running on into the next
// row when start.x + xLength passes the row's end, as glBufferSubData does;
// a run past the end of the grid is not copied: GL rejects it).
int vrt_upload_region_f(vrt_ctx* ctx, const int32_t* host_voxels, const float start[3], const float end[3]) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!host_voxels || !start || !end) return fail(VRT_ERR_INVALID, "null argument");
    float s0[3] = {start[0], start[1], start[2]}, e0[3] = {end[0], end[1], end[2]};
    for (int k = 0; k < 3; k++)
        if (!(std::fabs(s0[k]) < 16777216.0f) || !(std::fabs(e0[k]) < 16777216.0f))
            return fail(VRT_ERR_INVALID, "corners must be finite and below 2^24 in magnitude (the float loops "
                                         "of updatePartialGeometry would not end)");
    const GridDesc& g = ctx->grid;
    auto index = [&](float x, float y, float z) -> int64_t {
        const int64_t ix = (int32_t)x, iy = (int32_t)y, iz = (int32_t)z;
        if (ix >= 0 && iy >= 0 && iz >= 0 && ix < g.X && iy < g.Y && iz < g.Z)
            return ix + (int64_t)g.X * iy + (int64_t)g.XY * iz;
        return -1;
    };
    if (index(s0[0], s0[1], s0[2]) > index(e0[0], e0[1], e0[2]))
        for (int k = 0; k < 3; k++) std::swap(s0[k], e0[k]);
    const int32_t x_len = (int32_t)(e0[0] - s0[0]) + 1;
    if (x_len <= 0) return VRT_OK;
    // the int rows the two float loops visit (|values| < 2^24: each i++ moves on)
    auto rows = [](float from, float to, int64_t n) {
        std::vector<int64_t> out;
        for (float v = from; v < to; v++) {
            const int64_t r = (int32_t)v;
            if (r >= 0 && r < n) out.push_back(r);
        }
        return out;
    };
    const std::vector<int64_t> zs = rows(s0[2], e0[2], g.Z), ys = rows(s0[1], e0[1], g.Y);
    const int64_t xs = (int32_t)s0[0];
    if (xs < 0 || xs >= (int64_t)g.X || zs.empty() || ys.empty()) return VRT_OK;
    // Runs in increasing offset order (z outer, y inner, both ascending);
    // runs that touch or overlap are copied as one range: they carry the same
    // host words to the same cells.
    int64_t lo = INT64_MAX, hi = -1, run_lo = -1, run_hi = -1;
    auto flush = [&]() -> int {
        if (run_hi < 0) return VRT_OK;
        VRT_HIP(hipMemcpyAsync(g.vox + run_lo, host_voxels + run_lo, (size_t)(run_hi - run_lo) * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->stream));
        return VRT_OK;
    };
    for (int64_t z : zs)
        for (int64_t y : ys) {
            const int64_t off = xs + (int64_t)g.X * y + (int64_t)g.XY * z;
            if ((uint64_t)(off + x_len) > g.count) continue;
            if (run_hi >= 0 && off <= run_hi) {
                run_hi = std::max(run_hi, off + x_len);
            } else {
                if (int rc = flush()) return rc;
                run_lo = off;
                run_hi = off + x_len;
            }
            lo = std::min(lo, off);
            hi = std::max(hi, off + x_len);
        }
    if (int rc = flush()) return rc;
    if (hi < 0) return VRT_OK;
    // the cells written lie in whole x-rows of the z-slabs lo..hi-1 touch
    const int64_t z_lo = lo / (int64_t)g.XY, z_hi = (hi - 1) / (int64_t)g.XY + 1;
    int64_t y_lo = 0, y_hi = g.Y;
    if (z_hi - z_lo == 1) {
        y_lo = (lo % (int64_t)g.XY) / g.X;
        y_hi = ((hi - 1) % (int64_t)g.XY) / g.X + 1;
    }
    if (int rc = launch_solid_top(g, 0, y_lo, z_lo, g.X, y_hi, z_hi, false, ctx->stream)) return rc;
    if (int rc = refresh_render(ctx, 0, y_lo, z_lo, g.X, y_hi, z_hi)) return rc;
    return read_top(ctx);
}

int vrt_download(vrt_ctx* ctx, int32_t* host_voxels) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!host_voxels) return fail(VRT_ERR_INVALID, "null voxels");
    VRT_HIP(hipMemcpyAsync(host_voxels, ctx->grid.vox, ctx->grid.count * sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    VRT_HIP(hipStreamSynchronize(ctx->stream));
    return VRT_OK;
}

int vrt_depth_table(float* dist, int32_t* off, int32_t cap, int32_t* count) {
    auto t = reference_table();
    if (count) *count = (int32_t)t.size();
    for (int32_t i = 0; i < cap && i < (int32_t)t.size(); i++) {
        if (dist) dist[i] = t[i].dist;
        if (off) {
            off[3 * i] = t[i].x;
            off[3 * i + 1] = t[i].y;
            off[3 * i + 2] = t[i].z;
        }
    }
    return VRT_OK;
}

// computeDepthField threads, src/render.cpp:273-286,354-363.
int vrt_build_depth_field(vrt_ctx* ctx) {
    if (int rc = check_ctx(ctx)) return rc;
    const GridDesc& g = ctx->grid;
    if (!ctx->solid_bits) {
        const size_t bytes = (size_t)g.Y * g.Z * solid_bits_words(g) * sizeof(uint32_t);
        if (hipMalloc(&ctx->solid_bits, bytes) != hipSuccess) {
            (void)hipGetLastError();
            ctx->solid_bits = nullptr;
            return fail(VRT_ERR_NOMEM, "cannot allocate the depth-field build's solidity bitmap");
        }
    }
    if (int rc = launch_build_depth_field(g, ctx->table, ctx->bricks, ctx->solid_bits, ctx->stream)) return rc;
    return refresh_render_all(ctx);
}

// fixDepthField, src/render.cpp:226-253, over a box.
int vrt_fix_depth_field_box(vrt_ctx* ctx, int32_t x0, int32_t y0, int32_t z0, int32_t x1, int32_t y1, int32_t z1) {
    if (int rc = check_ctx(ctx)) return rc;
    if (int rc = launch_fix_depth_box(ctx->grid, ctx->table, x0, y0, z0, x1, y1, z1, 0, 0, 0, -1, ctx->stream))
        return rc;
    return refresh_render(ctx, x0, y0, z0, x1, y1, z1);
}

// removeSphere, src/level.cpp:30-56.
int vrt_remove_sphere(vrt_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t radius) {
    if (int rc = check_ctx(ctx)) return rc;
    if (radius < -3 || radius > 4096) return fail(VRT_ERR_INVALID, "radius must be in [-3, 4096]");
    if (int rc = remove_sphere_grid(ctx, cx, cy, cz, radius, ctx->stream)) return rc;
    return remove_sphere_refresh(ctx, cx, cy, cz, radius, ctx->stream);
}

// placeVoxel, src/render.cpp:256-262.
int vrt_place_voxel(vrt_ctx* ctx, int32_t x, int32_t y, int32_t z, int32_t voxel) {
    if (int rc = check_ctx(ctx)) return rc;
    const GridDesc& g = ctx->grid;
    if (x < 0 || y < 0 || z < 0 || (uint32_t)x >= g.X || (uint32_t)y >= g.Y || (uint32_t)z >= g.Z) return VRT_OK;
    if (int rc = launch_set_voxel(g, (int64_t)x + (int64_t)g.X * y + (int64_t)g.XY * z, y, voxel, ctx->stream))
        return rc;
    if (voxel >= 0) ctx->top_host = std::max(ctx->top_host, y);  // as k_set_voxel raises *top
    return refresh_render(ctx, x, y, z, (int64_t)x + 1, (int64_t)y + 1, (int64_t)z + 1);
}

// destroyVoxel, src/render.cpp:265-271.
int vrt_destroy_voxel(vrt_ctx* ctx, int32_t x, int32_t y, int32_t z) { return vrt_place_voxel(ctx, x, y, z, -1); }

// Start-up globals, src/main.cpp:26-39, src/render.cpp:304-311.
int vrt_frame_defaults(vrt_frame* f, int64_t grid_width, int32_t width, int32_t height) {
    if (!f || width <= 0 || height <= 0) return fail(VRT_ERR_INVALID, "invalid frame arguments");
    std::memset(f, 0, sizeof(*f));
    f->cam_pos[0] = 195.0f;
    f->cam_pos[1] = 55.0f;
    f->cam_pos[2] = 155.0f;
    f->light_pos[0] = (float)grid_width / 2.0f;
    f->light_pos[1] = (float)grid_width * 3.0f;
    f->light_pos[2] = (float)grid_width / 2.0f;
    std::memcpy(f->rotate_matrix, kIdentity, sizeof(kIdentity));
    f->aspect_ratio = (float)width / (float)height;
    f->focal = 1.0f;
    f->view_depth_field = 0;
    f->render_dist = 384;
    return vrt_init_local_lights(f);
}

// initLocalLights, src/render.cpp:304-311.
int vrt_init_local_lights(vrt_frame* f) {
    if (!f) return fail(VRT_ERR_INVALID, "null frame");
    for (int i = 0; i < VRT_MAX_LOCAL_LIGHTS; i++) {
        vrt_local_light& L = f->lights[i];
        L.pos[0] = L.pos[1] = L.pos[2] = -1.0f;
        L.rgb[0] = L.rgb[1] = L.rgb[2] = 0.0f;
        L.radius = 64.0f;
        L.reserved = 0.0f;
    }
    return VRT_OK;
}

// placeLocalLight, src/render.cpp:375-385.
int vrt_place_local_light(vrt_frame* f, float x, float y, float z, float diffuse) {
    if (!f) return fail(VRT_ERR_INVALID, "null frame");
    for (int i = 0; i < VRT_MAX_LOCAL_LIGHTS; i++) {
        vrt_local_light& L = f->lights[i];
        if (L.pos[0] < 0 || L.pos[1] < 0 || L.pos[2] < 0) {
            L.pos[0] = x;
            L.pos[1] = y;
            L.pos[2] = z;
            L.rgb[0] = L.rgb[1] = L.rgb[2] = diffuse;
            L.radius = 64.0f;
            return i;
        }
    }
    return -1;
}

int vrt_mat4_rotate(const float m[16], float angle_rad, const float axis[3], float out[16]) {
    if (!m || !axis || !out) return fail(VRT_ERR_INVALID, "null argument");
    mat4_rotate(m, angle_rad, axis, out);
    return VRT_OK;
}

int vrt_mat4_mul(const float a[16], const float b[16], float out[16]) {
    if (!a || !b || !out) return fail(VRT_ERR_INVALID, "null argument");
    mat4_mul(a, b, out);
    return VRT_OK;
}

// doMouseLook, src/controls.cpp:136-142.
int vrt_camera_matrix(float pitch_rad, float yaw_rad, float out[16]) {
    if (!out) return fail(VRT_ERR_INVALID, "null argument");
    const float ax[3] = {1, 0, 0}, ay[3] = {0, 1, 0};
    float rx[16], ry[16];
    mat4_rotate(kIdentity, pitch_rad, ax, rx);
    mat4_rotate(kIdentity, yaw_rad, ay, ry);
    mat4_mul(ry, rx, out);
    return VRT_OK;
}

// lightUpdate, src/render.cpp:400-401: rotate about z by radians(deg), then
// mat4 * vec4(start, 1) in GLM's pairwise order.
int vrt_light_orbit(float rotation_deg, const float start[3], float out[3]) {
    if (!start || !out) return fail(VRT_ERR_INVALID, "null argument");
    const float az[3] = {0, 0, 1};
    const float rad = rotation_deg * static_cast<float>(0.01745329251994329576923690768489);
    float r[16];
    mat4_rotate(kIdentity, rad, az, r);
    const float v[4] = {start[0], start[1], start[2], 1.0f};
    for (int k = 0; k < 3; k++) out[k] = (r[k] * v[0] + r[4 + k] * v[1]) + (r[8 + k] * v[2] + r[12 + k] * v[3]);
    return VRT_OK;
}

// updateUniforms, src/render.cpp:289-302.
int vrt_set_frame(vrt_ctx* ctx, const vrt_frame* f) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!f) return fail(VRT_ERR_INVALID, "null frame");
    if (int rc = check_frame(*f)) return rc;
    ctx->frame = *f;
    ctx->frame_set = true;
    return VRT_OK;
}

int vrt_draw_frame(vrt_ctx* ctx, const vrt_frame* f, const vrt_target* t) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!f || !t) return fail(VRT_ERR_INVALID, "null frame or target");
    if (int rc = check_frame(*f)) return rc;
    if (int rc = check_target(*t)) return rc;
    ctx->draws++;
    FrameDesc F = make_frame_desc(*f, *t);
    if (ctx->shadow_cutoff) F.row_rot = plain_row_rotation(*f, *t, ctx->top_host, F.primary_cutoff != 0);
    // the tile totals go into the partial-count array (kPartsCap) unless the
    // draw is being captured, the frame has more tiles than the array or
    // its 32-bit tile sums could overflow
    const int64_t tiles = (int64_t)((t->width + 7) / 8) * ((t->height + 7) / 8);
    unsigned long long* parts = nullptr;
    if (F.render_lim < kTileSum32Lim && tiles <= kPartsCap && !stream_capturing(ctx->stream)) {
        if (ctx->parts_used + tiles > kPartsCap)
            if (int rc = fold_parts(ctx)) return rc;
        parts = ctx->parts + ctx->parts_used;
    }
    if (int rc = launch_render_plain(render_grid(ctx), F, *t, ctx->counters, parts, ctx->stream)) return rc;
    if (parts) ctx->parts_used += tiles;
    return VRT_OK;
}

// glDrawArrays over the fullscreen quad, src/main.cpp:59.
int vrt_draw(vrt_ctx* ctx, const vrt_target* t) {
    if (int rc = check_ctx(ctx)) return rc;
    return vrt_draw_frame(ctx, &ctx->frame, t);
}

int vrt_read_stats(vrt_ctx* ctx, vrt_stats* out, int32_t reset) {
    if (int rc = check_ctx(ctx)) return rc;
    if (int rc = fold_parts(ctx)) return rc;
    std::vector<unsigned long long> h((size_t)kCounterSlots * kCounterStride);
    const size_t bytes = h.size() * sizeof(unsigned long long);
    VRT_HIP(hipMemcpyAsync(h.data(), ctx->counters, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VRT_HIP(hipStreamSynchronize(ctx->stream));
    if (out) {
        out->rays = out->dda_steps = out->pixels = 0;
        for (int i = 0; i < kCounterSlots; i++) {
            out->rays += h[(size_t)i * kCounterStride + 0];
            out->dda_steps += h[(size_t)i * kCounterStride + 1];
            out->pixels += h[(size_t)i * kCounterStride + 2];
        }
        out->draws = ctx->draws;
    }
    if (reset) {
        VRT_HIP(hipMemsetAsync(ctx->counters, 0, bytes, ctx->stream));
        VRT_HIP(hipStreamSynchronize(ctx->stream));
        ctx->draws = 0;
    }
    return VRT_OK;
}

// castRay, src/fshader.glsl:59-129, for a list of rays.
int vrt_cast_rays(vrt_ctx* ctx, int64_t n, const float* start, const float* dir, const int32_t* dist,
                  int32_t render_dist, int64_t* hit, float* hit_pos, uint8_t* face, uint32_t* steps) {
    if (int rc = check_ctx(ctx)) return rc;
    if (n < 0 || (n > 0 && (!start || !dir || !dist))) return fail(VRT_ERR_INVALID, "invalid ray arrays");
    if (render_dist < 0 || render_dist > (1 << 24)) return fail(VRT_ERR_INVALID, "render_dist must be in [0, 2^24]");
    return launch_cast_rays(render_grid(ctx), n, start, dir, dist, render_dist, hit, hit_pos, face, steps,
                            ctx->stream);
}

int vrt_selftest_reciprocal(vrt_ctx* ctx, uint64_t first, uint64_t count, uint64_t* mismatches) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!mismatches || first >= (1ull << 32)) return fail(VRT_ERR_INVALID, "invalid selftest range");
    unsigned long long* bad = nullptr;
    VRT_HIP(hipMalloc(&bad, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), ctx->stream);
    int rc = e == hipSuccess ? selftest_reciprocal(first, count, bad, ctx->stream) : hip_fail(e, "hipMemsetAsync");
    unsigned long long h = 0;
    if (!rc) {
        e = hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "selftest copy");
    }
    (void)hipFree(bad);
    if (rc) return rc;
    *mismatches = h;
    return VRT_OK;
}

int vrt_selftest_division(vrt_ctx* ctx, uint32_t d_begin, uint32_t d_count, uint64_t* mismatches) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!mismatches || d_begin >= (1u << 23)) return fail(VRT_ERR_INVALID, "invalid selftest range");
    unsigned long long* bad = nullptr;
    VRT_HIP(hipMalloc(&bad, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), ctx->stream);
    int rc = e == hipSuccess ? selftest_division(d_begin, d_count, bad, ctx->stream) : hip_fail(e, "hipMemsetAsync");
    unsigned long long h = 0;
    if (!rc) {
        e = hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "selftest copy");
    }
    (void)hipFree(bad);
    if (rc) return rc;
    *mismatches = h;
    return VRT_OK;
}

int vrt_world_generate_device(vrt_ctx* ctx, const vrt_world_params* p) {
    if (int rc = check_ctx(ctx)) return rc;
    if (!p) return fail(VRT_ERR_INVALID, "null params");
    if (int rc = world_generate_device(ctx->grid, *p, ctx->stream)) return rc;
    if (int rc = solid_top_full(ctx)) return rc;
    if (int rc = refresh_render_all(ctx)) return rc;
    return read_top(ctx);
}

}  // extern "C"
