// vrt_internal.h — declarations shared by the libvrt sources (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vrt.h"

namespace vrt {

// Sets the thread's last error message and returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define VRT_HIP(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return vrt::hip_fail(e_, #call); \
    } while (0)

// Render copy of the grid: every side padded with kGuard cells holding
// kGuardValue, the bits of a NaN.  A walk that steps onto one takes it for a
// depth-field jump of NaN cells, which ends the walk at once as a miss, as
// src/fshader.glsl:123-125 ends it on a cell outside the grid: the traversal
// then needs no bounds test per step.  (A NaN jump inside the grid ends the
// walk the same way in the reference, :114-121 then :83.)
constexpr int32_t kGuard = 3;
constexpr int32_t kGuardValue = (int32_t)0xFFC00000u;

// Grid as the kernels see it.  Cell (x, y, z) is vox[origin + x + SX*y +
// SXY*z]; X, Y, Z are the grid's extents.  The library's grid is dense
// (SX = X, SXY = XY, origin 0); the render copy is padded (kGuard cells on
// every side).
struct GridDesc {
    int32_t* vox;
    uint32_t X, Y, Z;
    uint64_t XY;
    uint64_t count;  // words allocated at vox
    uint32_t SX;
    uint64_t SXY;
    uint64_t origin;
    // Device word: an upper bound on the y of every solid voxel (-1 when there
    // is none).  Every entry point that can add solid voxels raises it; edits
    // that only remove leave it high, which stays an upper bound.  Render
    // kernels stop a shadow ray once it is provably above it; null: they walk
    // every shadow ray to the reference's end.
    int32_t* top;
    // The y and z a walk can index (the grid's and two cells past it on
    // either side) fit signed 24-bit multiplies.
    bool yz24() const { return (uint64_t)Y + 2 < (1ull << 23) && (uint64_t)Z + 2 < (1ull << 23); }
    // Index fits the 32-bit / 24-bit-multiply form (see Idx32).
    bool small() const { return count <= 0xffffffffull && SX < (1u << 23) && SXY < (1ull << 23) && yz24(); }
    // Byte offset fits the 32-bit / 24-bit-multiply form (see IdxB).
    bool bytes32() const {
        return count <= (1ull << 30) && 4ull * SX < (1ull << 23) && 4ull * SXY < (1ull << 23) && yz24();
    }
};

// One frame as the kernels see it: vrt_frame plus derived per-frame values.
struct FrameDesc {
    vrt_frame f;
    float render_lim;   // (float)render_dist
    // Per light: the largest float t with sqrt_rn(t) <= radius, so that
    // length(v) <= radius  <=>  dot(v, v) <= t  (sqrt_rn is monotone).
    float light_dd_max[VRT_MAX_LOCAL_LIGHTS];
    // Per light: the largest float a >= 0 with RN(a * a) <= light_dd_max.
    // dot(v, v) >= RN(v.x * v.x) (a sum of non-negative terms, each rounding
    // monotone), so |v.x| > a already puts the light out of range.
    float light_axis_max[VRT_MAX_LOCAL_LIGHTS];
    // Bit i set when light i is placed: every coordinate >= 0 (unused slots
    // hold -1, src/render.cpp:304-311; NaN compares false).
    uint32_t light_valid;
    float light_rinv[VRT_MAX_LOCAL_LIGHTS];  // RN(1 / radius)
    float inv_w, inv_h;                      // RN(1 / width), RN(1 / height) of the target
    // Primary rays may stop above every solid voxel too: only when nothing
    // shows their step count (no depth view, no step map in the target).
    int32_t primary_cutoff;
    // k_render_plain deals tile rows from this one on, wrapping to row 0
    // (dispatch order only; see plain_row_rotation).
    int32_t row_rot;
};

// Ray/step counters: kCounterSlots slots of one 128-byte line each
// (u64 rays, DDA steps, pixels); a wave adds into slot (its wave id mod
// kCounterSlots), so waves rarely share a line and the atomics never queue
// on one address.  Stats are the sum over slots.
constexpr int kCounterSlots = 4096;
constexpr int kCounterStride = 16;  // u64 per slot

// Plain draws do not add into the counters at the end of every wave: a
// device-scope atomic there keeps the wave's slot (and with it the start of
// the next workgroup, which needs a slot on every SIMD of the CU) until it
// completes, 3% of a 1080p frame.  Each wave stores its tile's totals into
// the context's partial-count array instead, one u64 per tile (DDA iterations
// in the low 32 bits, rays and pixels in 16 bits each above them), at the
// next free entries; the entries are added into the counters by a fold kernel
// when the array is full and before stats are read.  Under stream capture
// the draw keeps the atomics (a replayed graph would rewrite one set of
// entries).  The packed sums are exact while a tile's iteration total stays
// below 2^32: 64 lanes x 18 rays x lim < 2^32 for lim < kTileSum32Lim.
constexpr int64_t kPartsCap = (int64_t)1 << 22;  // entries (32 MiB)
constexpr float kTileSum32Lim = 3.7e6f;

// Queue of the persistent batch kernel: 8 counters (one per XCD), 128 B apart.
constexpr size_t kQueueBytes = 8 * 128;
// Tiles one multi-view launch may hold (queue counters are 32-bit).
constexpr int64_t kMaxBatchTiles = 0xffffffffll - 65536;

// Per-view work description for the persistent batch kernel.
struct ViewDesc {
    FrameDesc frame;
    vrt_target target;
    int32_t tiles_x;       // tiles across
    int32_t tiles;         // tiles in this view
    int64_t first_tile;    // prefix sum of tiles over earlier views
};

struct DepthTableDev {
    const int32_t* packed;  // (dx+8) | (dy+8)<<5 | (dz+8)<<10, sorted by dist descending
    const float* dist;      // dist of each packed entry
    int32_t n;              // entries with dist > -6
    int32_t n_near;         // leading entries with dist > -2
    // The same entries grouped into rows: one per (distance, dy, dz), in the
    // same order of distance, with the mask of their dx (bit dx+7).
    const int32_t* rows;      // mask | (dy+8)<<16 | (dz+8)<<21
    const float* rows_dist;   // dist of each row
    int32_t n_rows;           // rows with dist > -6
    int32_t n_rows_near;      // leading rows with dist > -2
};

struct Ctx;

// ---- kernels launch wrappers (vrt_render.hip / vrt_edit.hip / vrt_world.hip)
// parts == nullptr: the wave totals go into `counters` by atomics; else into
// parts[tile] (see kPartsCap).
int launch_render_plain(const GridDesc& g, const FrameDesc& f, const vrt_target& t,
                        unsigned long long* counters, unsigned long long* parts, hipStream_t s);
// Adds parts[0, n) into the counter slots.
int launch_fold_parts(const unsigned long long* parts, int64_t n, unsigned long long* counters, hipStream_t s);
int launch_cast_rays(const GridDesc& g, int64_t n, const float* start, const float* dir,
                     const int32_t* dist, int32_t render_dist, int64_t* hit, float* hit_pos,
                     uint8_t* face, uint32_t* steps, hipStream_t s);
// parts (chunk_tiles 0 only; nullptr: atomics): one entry per 64-pixel pass
// (batch_items), added into the counters by launch_fold_parts afterwards.
int launch_render_batch(const GridDesc& g, const ViewDesc* views_dev, int32_t n_views,
                        int64_t total_tiles, const vrt_batch_config& cfg, int32_t n_workgroups,
                        unsigned int* queue, unsigned long long* counters, unsigned long long* parts,
                        hipStream_t s);
// Entries a hardware-scheduled multi-view launch writes into `parts`, or -1
// when its counts must go by atomics (a view's render_lim too large for the
// 32-bit pass sums, or a persistent configuration).
int64_t batch_items(const std::vector<ViewDesc>& views, int64_t total_tiles, const vrt_batch_config& cfg);
int launch_queue_reset(unsigned int* queue, hipStream_t s);
int selftest_division(uint32_t d_begin, uint32_t d_count, unsigned long long* bad_dev, hipStream_t s);
int selftest_reciprocal(uint64_t first, uint64_t count, unsigned long long* bad_dev, hipStream_t s);
int batch_occupancy(const vrt_batch_config& cfg, int32_t* waves_per_cu);

int launch_fix_depth_box(const GridDesc& g, const DepthTableDev& t, int32_t x0, int32_t y0, int32_t z0,
                         int32_t x1, int32_t y1, int32_t z1, int32_t sphere_cx, int32_t sphere_cy,
                         int32_t sphere_cz, int32_t sphere_r, hipStream_t s);
// Whole-grid depth field: solidity bitmap (Y*Z rows of solid_bits_words(g)
// words), brick flags (one byte per 8x8x8 brick), then the brick kernel.
uint32_t solid_bits_words(const GridDesc& g);
int launch_build_depth_field(const GridDesc& g, const DepthTableDev& t, uint8_t* bricks, uint32_t* bits,
                             hipStream_t s);
int launch_carve_sphere(const GridDesc& g, int32_t cx, int32_t cy, int32_t cz, int32_t r, hipStream_t s);
int launch_set_voxel(const GridDesc& g, int64_t index, int32_t y, int32_t v, hipStream_t s);
// Raises *g.top to the highest y with a solid voxel in the box [x0,x1) x
// [y0,y1) x [z0,z1) (clipped to the grid); with reset, sets it to -1 first.
int launch_solid_top(const GridDesc& g, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1, int64_t z1,
                     bool reset, hipStream_t s);
int launch_fill(int32_t* p, uint64_t n, int32_t v, hipStream_t s);
// dst's cells of the box [x0,x0+nx) x [y0,y0+ny) x [z0,z0+nz) := src's.
int launch_copy_box(const GridDesc& src, const GridDesc& dst, int64_t x0, int64_t y0, int64_t z0, int64_t nx,
                    int64_t ny, int64_t nz, hipStream_t s);
int launch_copy(int32_t* dst, const int32_t* src, uint64_t n, hipStream_t s);
int launch_world_generate(const GridDesc& g, const vrt_world_params& p, const int32_t* cave_list,
                          int32_t n_caves, const int32_t* box_list, int32_t n_boxes,
                          const int32_t* cave_bins, const int32_t* box_bins, int32_t bin_size,
                          int32_t bins_x, int32_t bins_z, hipStream_t s);

// ---- host helpers
std::vector<float> depth_table_dist();              // reference order
std::vector<int32_t> depth_table_offsets();         // reference order, xyz triples
double now_seconds();

// World generator pieces shared by the host and device generators.
struct WorldPlan {
    vrt_world_params p;
    std::vector<int32_t> caves;  // cx, cy, cz, r
    std::vector<int32_t> boxes;  // x0, y0, z0, x1, y1, z1 (half-open, clipped)
};
int make_world_plan(const vrt_world_params& p, WorldPlan& plan);

}  // namespace vrt

// Context behind the opaque vrt_ctx handle.
struct vrt_ctx {
    int32_t device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    vrt::GridDesc grid{};
    // Padded render copy of `grid` (see kGuard); every entry point that
    // writes the grid refreshes the part it wrote, on the same stream.
    vrt::GridDesc rgrid{};
    vrt_frame frame{};
    bool frame_set = false;
    unsigned long long* counters = nullptr;  // rays, iters, pixels
    unsigned long long* parts = nullptr;     // plain draws' partial counts (kPartsCap entries)
    int64_t parts_used = 0;                  // entries not yet folded into `counters`
    uint64_t draws = 0;
    int32_t* table_packed = nullptr;
    float* table_dist = nullptr;
    int32_t* table_rows = nullptr;
    float* table_rows_dist = nullptr;
    vrt::DepthTableDev table{};
    uint8_t* bricks = nullptr;
    uint32_t* solid_bits = nullptr;  // whole-grid build scratch, allocated by the first build
    int32_t num_cus = 0;
    bool tuned = false;
    vrt_batch_config tuned_cfg{8, 8, 0, 0};
    bool shadow_cutoff = true;  // vrt_set_shadow_cutoff
    // Host copy of *grid.top, read back by the entry points that can raise it
    // (they synchronize anyway or are not on a per-frame path).
    int32_t top_host = -1;
};

namespace vrt {
// The grid as the render kernels get it: the padded copy, with `top` only
// with the shadow cutoff on.
inline GridDesc render_grid(const vrt_ctx* c) {
    GridDesc g = c->rgrid;
    g.top = c->shadow_cutoff ? c->grid.top : nullptr;
    return g;
}
int solid_top_full(vrt_ctx* c);  // recomputes *grid.top over the whole grid
// Copies the box [x0,x1) x [y0,y1) x [z0,z1) (clipped to the grid) of the
// grid into the render copy, on the context's stream.
int refresh_render(vrt_ctx* c, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1, int64_t z1);
int refresh_render_all(vrt_ctx* c);
int refresh_render_on(vrt_ctx* c, hipStream_t s, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1,
                      int64_t z1);
// removeSphere in two parts on stream s: the carve and depth-field repair of
// the grid (render kernels never read it), then the copy of the edited box
// into the render copy (what in-flight renders must not see).
int remove_sphere_grid(vrt_ctx* c, int32_t cx, int32_t cy, int32_t cz, int32_t radius, hipStream_t s);
int remove_sphere_refresh(vrt_ctx* c, int32_t cx, int32_t cy, int32_t cz, int32_t radius, hipStream_t s);
// The same two refreshes into another copy laid out like the render copy.
int refresh_copy_on(vrt_ctx* c, const GridDesc& dst, hipStream_t s, int64_t x0, int64_t y0, int64_t z0, int64_t x1,
                    int64_t y1, int64_t z1);
int remove_sphere_refresh_into(vrt_ctx* c, const GridDesc& dst, int32_t cx, int32_t cy, int32_t cz, int32_t radius,
                               hipStream_t s);
}  // namespace vrt

namespace vrt {
int world_generate_device(const GridDesc& g, const vrt_world_params& p, hipStream_t s);
FrameDesc make_frame_desc(const vrt_frame& f, const vrt_target& target);
int32_t plain_row_rotation(const vrt_frame& f, const vrt_target& t, int32_t top, bool primary_cutoff);
int check_frame(const vrt_frame& f);
// Tile shape in {8x8, 16x4, 32x2, 16x16, 64x1}, chunk_tiles in [0, 4096],
// waves_per_cu in [0, 64]; the batch kernels assume all three.
int check_batch_config(const vrt_batch_config& c);
// One view's descriptor for a multi-view launch, its tiles numbered from
// first_tile; checks the frame and target.
struct ViewDesc;
int make_view_desc(const vrt_view& v, const vrt_batch_config& cfg, int64_t first_tile, ViewDesc& d);
// Workgroups of a multi-view launch: one per four tiles when the hardware
// schedules (chunk_tiles 0), else persistent workgroups from occupancy.
int persistent_workgroups(vrt_ctx* c, const vrt_batch_config& cfg, int64_t total_tiles, int32_t* out);
int check_target(const vrt_target& t);
}  // namespace vrt
