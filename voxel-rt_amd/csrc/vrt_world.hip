// vrt_world.hip — synthetic worlds for benchmarks and tests (no reference
// counterpart: the reference builds flat layers and trees, src/level.cpp:82-138).
// The plan (cave spheres and boxes) is drawn on the host from a seeded
// std::mt19937_64; each (x, z) column is then written by the same function on
// the host (std::thread over z-slabs) or on the device (one thread per column).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <random>
#include <thread>
#include <vector>

#include "vrt_internal.h"
#include "vrt_worldgen.h"

namespace vrt {

namespace {

constexpr int32_t kBin = 64;

int64_t uniform_int(std::mt19937_64& rng, int64_t lo, int64_t hi) {  // inclusive
    if (hi <= lo) return lo;
    return lo + (int64_t)(rng() % (uint64_t)(hi - lo + 1));
}

struct Bins {
    int32_t bins_x = 0, bins_z = 0;
    std::vector<int32_t> cave_start, cave_items, box_start, box_items;
};

void build_bins(const WorldPlan& plan, Bins& b) {
    const int64_t X = plan.p.size[0], Z = plan.p.size[2];
    b.bins_x = (int32_t)((X + kBin - 1) / kBin);
    b.bins_z = (int32_t)((Z + kBin - 1) / kBin);
    const int64_t nb = (int64_t)b.bins_x * b.bins_z;
    std::vector<std::vector<int32_t>> cv(nb), bv(nb);
    const int64_t nc = (int64_t)plan.caves.size() / 4;
    for (int64_t i = 0; i < nc; i++) {
        const int32_t* c = &plan.caves[4 * i];
        int64_t x0 = std::max<int64_t>(0, c[0] - c[3]) / kBin, x1 = std::min<int64_t>(X - 1, c[0] + c[3]) / kBin;
        int64_t z0 = std::max<int64_t>(0, c[2] - c[3]) / kBin, z1 = std::min<int64_t>(Z - 1, c[2] + c[3]) / kBin;
        for (int64_t bz = z0; bz <= z1; bz++)
            for (int64_t bx = x0; bx <= x1; bx++) cv[bx + b.bins_x * bz].push_back((int32_t)i);
    }
    const int64_t nbx = (int64_t)plan.boxes.size() / 6;
    for (int64_t i = 0; i < nbx; i++) {
        const int32_t* q = &plan.boxes[6 * i];
        int64_t x0 = q[0] / kBin, x1 = (q[3] - 1) / kBin, z0 = q[2] / kBin, z1 = (q[5] - 1) / kBin;
        for (int64_t bz = z0; bz <= z1; bz++)
            for (int64_t bx = x0; bx <= x1; bx++) bv[bx + b.bins_x * bz].push_back((int32_t)i);
    }
    b.cave_start.assign(nb + 1, 0);
    b.box_start.assign(nb + 1, 0);
    for (int64_t i = 0; i < nb; i++) {
        b.cave_start[i + 1] = b.cave_start[i] + (int32_t)cv[i].size();
        b.box_start[i + 1] = b.box_start[i] + (int32_t)bv[i].size();
        b.cave_items.insert(b.cave_items.end(), cv[i].begin(), cv[i].end());
        b.box_items.insert(b.box_items.end(), bv[i].begin(), bv[i].end());
    }
}

// One column, written through `put(y, word)`.
template <typename Put>
VRT_HD void write_column(int64_t x, int64_t z, int64_t Y, const vrt_world_params& p, const int32_t* caves,
                         const int32_t* boxes, const int32_t* cave_idx, int32_t nc, const int32_t* box_idx,
                         int32_t nbx, Put put) {
    const int32_t h = wg::surface_height(x, z, p.seed, p.octaves, p.base_cell, p.height_min, p.height_max);
    for (int64_t y = 0; y < Y; y++) {
        int32_t v = -1;
        if (y < h) v = wg::terrain_colour((int32_t)(h - 1 - y), wg::jitter_of(x, y, z, p.seed, p.jitter));
        put(y, v);
    }
    for (int32_t k = 0; k < nc; k++) {
        int64_t lo, hi;
        if (!wg::cave_span(caves + 4 * cave_idx[k], x, z, lo, hi)) continue;
        if (lo < 0) lo = 0;
        if (hi > Y - 1) hi = Y - 1;
        for (int64_t y = lo; y <= hi; y++) put(y, -1);
    }
    for (int32_t k = 0; k < nbx; k++) {
        int64_t lo, hi;
        if (!wg::box_span(boxes + 6 * box_idx[k], x, z, lo, hi)) continue;
        for (int64_t y = lo; y <= hi; y++) put(y, wg::box_colour(wg::jitter_of(x, y, z, p.seed, p.jitter)));
    }
}

__global__ void __launch_bounds__(256) k_world_columns(GridDesc g, vrt_world_params p, const int32_t* caves,
                                                        const int32_t* boxes, const int32_t* cave_start,
                                                        const int32_t* cave_items, const int32_t* box_start,
                                                        const int32_t* box_items, int32_t bins_x) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t z = blockIdx.y;
    if (x >= g.X) return;
    const int64_t bin = x / kBin + (int64_t)bins_x * (z / kBin);
    int32_t* col = g.vox + x + g.XY * (uint64_t)z;
    const uint64_t X = g.X;
    write_column(x, z, (int64_t)g.Y, p, caves, boxes, cave_items + cave_start[bin], cave_start[bin + 1] - cave_start[bin],
                 box_items + box_start[bin], box_start[bin + 1] - box_start[bin],
                 [&](int64_t y, int32_t v) { col[X * (uint64_t)y] = v; });
}

}  // namespace

int make_world_plan(const vrt_world_params& p, WorldPlan& plan) {
    if (p.size[0] <= 0 || p.size[1] <= 0 || p.size[2] <= 0 || p.size[0] > 0x7fffffff || p.size[1] > 0x7fffffff ||
        p.size[2] > 0x7fffffff)
        return fail(VRT_ERR_INVALID, "world size out of range");
    if (p.octaves < 1 || p.octaves > 16 || p.base_cell < 1 || p.height_max < p.height_min || p.height_min < 0 ||
        p.caves < 0 || p.boxes < 0 || p.cave_r_min < 0 || p.cave_r_max < p.cave_r_min || p.box_min < 1 ||
        p.box_max < p.box_min || p.jitter < 0 || p.jitter > 255)
        return fail(VRT_ERR_INVALID, "world parameters out of range");
    plan.p = p;
    plan.caves.clear();
    plan.boxes.clear();
    const int64_t X = p.size[0], Y = p.size[1], Z = p.size[2];
    std::mt19937_64 rng(p.seed);
    for (int32_t i = 0; i < p.caves; i++) {
        const int64_t cx = uniform_int(rng, 0, X - 1), cz = uniform_int(rng, 0, Z - 1);
        const int64_t r = uniform_int(rng, p.cave_r_min, p.cave_r_max);
        const int32_t h = wg::surface_height(cx, cz, p.seed, p.octaves, p.base_cell, p.height_min, p.height_max);
        const int64_t cy = uniform_int(rng, 1, std::max<int64_t>(1, h - 2));
        plan.caves.insert(plan.caves.end(), {(int32_t)cx, (int32_t)cy, (int32_t)cz, (int32_t)r});
    }
    for (int32_t i = 0; i < p.boxes; i++) {
        const int64_t x0 = uniform_int(rng, 0, X - 1), z0 = uniform_int(rng, 0, Z - 1);
        const int64_t sx = uniform_int(rng, p.box_min, p.box_max), sy = uniform_int(rng, p.box_min, p.box_max),
                      sz = uniform_int(rng, p.box_min, p.box_max);
        const int64_t mx = std::min(X - 1, x0 + sx / 2), mz = std::min(Z - 1, z0 + sz / 2);
        const int32_t h = wg::surface_height(mx, mz, p.seed, p.octaves, p.base_cell, p.height_min, p.height_max);
        const int64_t y0 = std::max<int64_t>(0, h + uniform_int(rng, -3, 8));
        const int64_t x1 = std::min(X, x0 + sx), y1 = std::min(Y, y0 + sy), z1 = std::min(Z, z0 + sz);
        if (y0 >= y1) continue;
        plan.boxes.insert(plan.boxes.end(),
                          {(int32_t)x0, (int32_t)y0, (int32_t)z0, (int32_t)x1, (int32_t)y1, (int32_t)z1});
    }
    return VRT_OK;
}

}  // namespace vrt

using namespace vrt;

// The reference's own world, initVoxels (src/level.cpp:82-138) for an
// X x Y x Z grid (reference 512 x 96 x 512): stone up to y 25, dirt to 33,
// grass to 36 with a colour variation of 5*((x+y+z)%3); then a tree at every
// (x, z) in [10, X-10) x [10, Z-10) with x%30 == 0 and z%25 == 0: a 6-high
// trunk at (x+1+z%7, 36, z) and a radius-6 bush at (x+z%7, 46, z).  Kept
// quirks: placeTrunk's extent is (-length)>>2 .. length>>2 (:60,62), both
// helpers call destroyVoxel with LOCAL offsets (:11,64), clearing cells near
// the origin, and placeTrunk checks only the upper bounds before placeVoxel.
namespace {
struct RefWorld {
    int32_t* v;
    int64_t X, Y, Z;
    bool in(int64_t x, int64_t y, int64_t z) const {
        return x >= 0 && y >= 0 && z >= 0 && x < X && y < Y && z < Z;
    }
    void put(int64_t x, int64_t y, int64_t z, int32_t w) {
        if (in(x, y, z)) v[x + X * y + X * Y * z] = w;
    }
};

int32_t rgb(int32_t r, int32_t g, int32_t b) { return (((r << 8) + g) << 8) + b; }

void ref_trunk(RefWorld& w, int64_t px, int64_t py, int64_t pz) {
    const int length = 6;
    for (int z = (-length) >> 2; z < (length >> 2); z++)
        for (int y = 0; y < length; y++)
            for (int x = (-length) >> 2; x < (length >> 2); x++) {
                if (!(x + px < w.X && y + py < w.Y && z + pz < w.Z)) continue;
                w.put(x, y, z, -1);  // destroyVoxel(x, y, z): local offsets
                const int v = ((x + z) % 2) * 10;
                w.put(x + px, y + py, z + pz, rgb(128 - v, 100 - v, 15));
            }
}

void ref_bush(RefWorld& w, int64_t px, int64_t py, int64_t pz) {
    const int r = 6;
    for (int z = -r; z < r; z++)
        for (int y = -r; y < r; y++)
            for (int x = -r; x < r; x++) {
                if (!w.in(x + px, y + py, z + pz) || x * x + y * y + z * z >= r * r) continue;
                w.put(x, y, z, -1);  // destroyVoxel(x, y, z): local offsets
                w.put(x + px, y + py, z + pz, rgb(15, 128 - ((x + y + z) % 3) * 20, 15));
            }
}
}  // namespace

extern "C" int vrt_world_reference(int64_t X, int64_t Y, int64_t Z, int32_t* out) {
    if (!out || X <= 0 || Y <= 0 || Z <= 0 || X > 0x7fffffff || Y > 0x7fffffff || Z > 0x7fffffff)
        return fail(VRT_ERR_INVALID, "invalid reference world arguments");
    RefWorld w{out, X, Y, Z};
    for (int64_t z = 0; z < Z; z++)
        for (int64_t y = 0; y < Y; y++)
            for (int64_t x = 0; x < X; x++) {
                const int32_t cv = (int32_t)(5 * ((x + y + z) % 3));
                int32_t v = -1;
                if (y <= 25) v = rgb(90 + cv, 90 + cv, 90 + cv);
                else if (y <= 33) v = rgb(120 + cv, 100 + cv, 0);
                else if (y <= 36) v = rgb(10, 130 + cv, 10);
                out[x + X * y + X * Y * z] = v;
            }
    for (int64_t z = 10; z < Z - 10; z++)
        for (int64_t x = 10; x < X - 10; x++)
            if (x % 30 == 0 && z % 25 == 0) {
                ref_trunk(w, x + 1 + z % 7, 36, z);
                ref_bush(w, x + z % 7, 36 + 10, z);
            }
    return VRT_OK;
}

extern "C" int vrt_world_surface(const vrt_world_params* p, int64_t x, int64_t z, int32_t* height) {
    if (!p || !height) return fail(VRT_ERR_INVALID, "null argument");
    *height = wg::surface_height(x, z, p->seed, p->octaves, p->base_cell, p->height_min, p->height_max);
    return VRT_OK;
}

extern "C" int vrt_world_generate(const vrt_world_params* p, int32_t* out, int32_t threads) {
    if (!p || !out) return fail(VRT_ERR_INVALID, "null argument");
    WorldPlan plan;
    int rc = make_world_plan(*p, plan);
    if (rc) return rc;
    Bins b;
    build_bins(plan, b);
    const int64_t X = p->size[0], Y = p->size[1], Z = p->size[2];
    if (threads <= 0) threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    auto work = [&](int64_t z0, int64_t z1) {
        for (int64_t z = z0; z < z1; z++)
            for (int64_t x = 0; x < X; x++) {
                const int64_t bin = x / kBin + (int64_t)b.bins_x * (z / kBin);
                int32_t* col = out + x + X * Y * z;
                write_column(x, z, Y, *p, plan.caves.data(), plan.boxes.data(), b.cave_items.data() + b.cave_start[bin],
                             b.cave_start[bin + 1] - b.cave_start[bin], b.box_items.data() + b.box_start[bin],
                             b.box_start[bin + 1] - b.box_start[bin],
                             [&](int64_t y, int32_t v) { col[X * y] = v; });
            }
    };
    std::vector<std::thread> pool;
    const int64_t per = (Z + threads - 1) / threads;
    for (int32_t t = 0; t < threads; t++) {
        const int64_t z0 = t * per, z1 = std::min(Z, z0 + per);
        if (z0 >= z1) break;
        pool.emplace_back(work, z0, z1);
    }
    for (auto& th : pool) th.join();
    return VRT_OK;
}

namespace vrt {

int world_generate_device(const GridDesc& g, const vrt_world_params& p, hipStream_t s) {
    WorldPlan plan;
    int rc = make_world_plan(p, plan);
    if (rc) return rc;
    if ((uint64_t)p.size[0] != g.X || (uint64_t)p.size[1] != g.Y || (uint64_t)p.size[2] != g.Z)
        return fail(VRT_ERR_INVALID, "world size differs from the context's grid");
    Bins b;
    build_bins(plan, b);
    // one device buffer: caves | boxes | cave_start | cave_items | box_start | box_items
    std::vector<int32_t> blob;
    auto put = [&](const std::vector<int32_t>& v) {
        size_t off = blob.size();
        blob.insert(blob.end(), v.begin(), v.end());
        blob.push_back(0);  // keeps every part non-empty
        return off;
    };
    const size_t o_c = put(plan.caves), o_b = put(plan.boxes), o_cs = put(b.cave_start), o_ci = put(b.cave_items),
                 o_bs = put(b.box_start), o_bi = put(b.box_items);
    int32_t* d = nullptr;
    VRT_HIP(hipMalloc(&d, blob.size() * sizeof(int32_t)));
    hipError_t e = hipMemcpyAsync(d, blob.data(), blob.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        dim3 grid((uint32_t)((g.X + 255) / 256), (uint32_t)g.Z);
        hipLaunchKernelGGL(k_world_columns, grid, dim3(256), 0, s, g, p, d + o_c, d + o_b, d + o_cs, d + o_ci,
                           d + o_bs, d + o_bi, b.bins_x);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "world generation");
    return VRT_OK;
}

}  // namespace vrt
