// vrt_render.hip — render kernels for gfx950 (wave64).
//
//   k_render_plain : one frame; a wave shades one 8x8 screen tile (one ray per
//                    lane), a 256-thread workgroup four horizontally adjacent
//                    tiles, in row-major order.  The four share one CU and its
//                    L1: on an 8 GiB grid, one-wave workgroups (tiles spread
//                    over CUs) made 4K frames 38% slower, 2x2, 4x2 or 8x1 tile
//                    groups 3-4% slower (on the 256 MiB grid one-wave groups
//                    were 2% faster, 2x2 even).  The hardware deals
//                    workgroups to the 8 XCDs round-robin, so every XCD gets
//                    an even share of sky, horizon and terrain tiles; giving
//                    each XCD a contiguous band of rows instead (for L2
//                    locality) made the XCDs' loads unequal and the frame 29%
//                    slower.
//   k_render_views : many views in one launch, four consecutive (view, tile)
//                    items per workgroup, scheduled by the hardware.
//   k_render_batch : many views in one persistent launch; every wave pops
//                    chunks of (view, tile) work from a device atomic queue
//                    until the queue is empty (the exit every wave reaches).
//   k_cast_rays    : castRay for an arbitrary list of rays.
// The per-pixel work is src/fshader.glsl main() (vrt_device.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "vrt_device.h"
#include "vrt_internal.h"

namespace vrt {

namespace {

constexpr int kWave = 64;
constexpr int kPlainWaves = 4;                 // waves per workgroup
constexpr int kPlainThreads = kWave * kPlainWaves;
constexpr int kXcds = 8;
constexpr int kQueueStride = 32;  // queue counters 128 bytes apart

// Adds this wave's totals into its counter slot (see kCounterSlots).
// Every lane of the wave must be active.  Per-lane totals are 32-bit sums of
// per-pixel counts, far below 2^58.
__device__ __forceinline__ void flush_counters(unsigned long long* counters, unsigned long long rays,
                                               unsigned long long iters, unsigned long long pixels) {
    if (!counters) return;
    rays = dev::wave_sum_u58(rays);
    iters = dev::wave_sum_u58(iters);
    pixels = dev::wave_sum_u58(pixels);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const uint32_t wave = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        unsigned long long* c = counters + (size_t)(wave % kCounterSlots) * kCounterStride;
        atomicAdd(c + 0, rays);
        atomicAdd(c + 1, iters);
        atomicAdd(c + 2, pixels);
    }
}

// The same for one tile per wave, with wave sums in 32 bits: a lane casts at
// most 18 rays (primary, sun, 16 lights) of at most `lim` iterations each (a
// walk ends at its budget), so 64 * 18 * lim < 2^32 keeps every sum exact;
// rays (at most 1152) and pixels (at most 64) share one word.
__device__ __forceinline__ void flush_counters_tile(unsigned long long* counters, uint32_t rays, uint32_t iters,
                                                    uint32_t pixels, float lim) {
    if (!counters) return;
    if (!(lim < kTileSum32Lim)) {
        flush_counters(counters, rays, iters, pixels);
        return;
    }
    const uint32_t rp = dev::wave_sum_u32(rays | (pixels << 16));
    iters = dev::wave_sum_u32(iters);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const uint32_t wave = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        unsigned long long* c = counters + (size_t)(wave % kCounterSlots) * kCounterStride;
        atomicAdd(c + 0, (unsigned long long)(rp & 0xffffu));
        atomicAdd(c + 1, (unsigned long long)iters);
        atomicAdd(c + 2, (unsigned long long)(rp >> 16));
    }
}

// waves_per_eu(8, 8): without it the build ran at 7 resident waves per SIMD
// (peak measured with per-wave timestamps: 7168 waves on 1024 SIMDs); with it,
// 8 (a few registers of the rare IEEE-division path spill), 4% faster.
template <typename IX>
__global__ void __launch_bounds__(kPlainThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_render_plain(GridDesc g, FrameDesc F, vrt_target t, unsigned long long* counters, unsigned long long* parts) {
    const int32_t top = g.top ? *g.top : INT32_MAX;
    const uint32_t lb = blockIdx.x;
    const int32_t tiles_x = (t.width + 7) >> 3;
    const int32_t tiles_y = (t.height + 7) >> 3;
    // the wave's tile is uniform: the tile arithmetic runs on the scalar unit
    const int32_t tile = (int32_t)lb * kPlainWaves + __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
    const int32_t lane = threadIdx.x & 63;
    uint32_t rays = 0, iters = 0, pixels = 0;
    if (tile < tiles_x * tiles_y) {
        int32_t trow = tile / tiles_x + F.row_rot;
        if (trow >= tiles_y) trow -= tiles_y;
        const int32_t px = (tile % tiles_x) * 8 + (lane & 7);
        const int32_t row = trow * 8 + (lane >> 3);
        if (px < t.width && row < t.height) {
            float vx, vy;
            dev::pixel_vpos(t.width, t.height, F.inv_w, F.inv_h, px, row, vx, vy);
            dev::PixelResult p;
            dev::shade_pixel<IX>(g, F, vx, vy, p, top);
            dev::store_pixel<IX>(g, t, (int64_t)row * t.width + px, p);
            rays = p.rays;
            iters = p.iters;
            pixels = 1;
        }
    }
    if (parts) {
        // the tile's totals by plain store (the host passes parts only when
        // F.render_lim < kTileSum32Lim, so the 32-bit sums are exact)
        const uint32_t rp = dev::wave_sum_u32(rays | (pixels << 16));
        const uint32_t it = dev::wave_sum_u32(iters);
        if (lane == 0 && tile < tiles_x * tiles_y) parts[tile] = (unsigned long long)it | ((unsigned long long)rp << 32);
    } else {
        flush_counters_tile(counters, rays, iters, pixels, F.render_lim);
    }
}

// Adds plain draws' partial counts (iterations | rays << 32 | pixels << 48
// per tile) into the counter slots; every lane of the grid takes part.
__global__ void __launch_bounds__(256) k_fold_parts(const unsigned long long* __restrict__ parts, int64_t n,
                                                     unsigned long long* counters) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long v = i < n ? parts[i] : 0ull;
    flush_counters(counters, (v >> 32) & 0xffffull, v & 0xffffffffull, v >> 48);
}

template <typename IX>
__global__ void __launch_bounds__(256) k_cast_rays(GridDesc g, int64_t n, const float* __restrict__ start,
                                                   const float* __restrict__ dir,
                                                   const int32_t* __restrict__ dist, float render_lim,
                                                   int64_t* hit, float* hit_pos, uint8_t* face, uint32_t* steps) {
    const int32_t top = g.top ? *g.top : INT32_MAX;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lim = (float)dist[i];
    lim = lim < render_lim ? lim : render_lim;
    // the cutoff applies when the caller does not ask for the step counts
    const float dy = dir[3 * i + 1];
    dev::Hit h;
    dev::cast_ray<IX>(g, start[3 * i], start[3 * i + 1], start[3 * i + 2], dir[3 * i], dy, dir[3 * i + 2], lim, h,
                      (steps == nullptr && dy >= 0.0f) ? top : INT32_MAX);
    const bool hitp = h.index != -1;
    if (hit) hit[i] = hitp ? IX::ref_index(g, h.index) : -1;
    if (hit_pos) {
        hit_pos[3 * i] = hitp ? h.pos[0] : 0.0f;
        hit_pos[3 * i + 1] = hitp ? h.pos[1] : 0.0f;
        hit_pos[3 * i + 2] = hitp ? h.pos[2] : 0.0f;
    }
    if (face) face[i] = hitp ? (uint8_t)(h.axis * 3 + (h.step + 1)) : (uint8_t)VRT_FACE_MISS;
    if (steps) steps[i] = h.iters;
}

// One tile of a multi-view launch: `tile` indexes the concatenated tile lists
// of the views (largest view first, each row-major).  Tile shape TW x TH with
// TW a power of two and TW*TH a multiple of 64: the wave shades 64 pixels per
// pass (lane -> (lane % TW, lane / TW)) and TW*TH/64 passes.  The shape is a
// template parameter so that the pass loop has a constant trip count and
// constant lane arithmetic (with runtime shapes the kernels spilled registers
// to scratch).  Per-lane totals stay far below 2^32 for any launch the host
// accepts (tiles per launch < 2^32; 18 rays and at most 18 * 2^24 steps a
// pixel would overflow only past ~14 pixels at the step cap); they are widened
// when the wave flushes.
template <typename IX, int TW, int TH>
__device__ __forceinline__ void shade_view_tile(const GridDesc& g, const ViewDesc* __restrict__ views,
                                                int32_t n_views, int64_t tile, int32_t first_pass,
                                                int32_t n_passes, int32_t lane, int32_t top, uint32_t& rays,
                                                uint32_t& iters, uint32_t& pixels) {
    static_assert((TW & (TW - 1)) == 0 && TW <= kWave && (TW * TH) % kWave == 0, "tile shape");
    constexpr int32_t kRowsPerPass = kWave / TW;
    // view owning this tile (uniform across the wave)
    int32_t lo = 0, hi = n_views - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (views[mid].first_tile <= tile) lo = mid;
        else hi = mid - 1;
    }
    const ViewDesc& V = views[__builtin_amdgcn_readfirstlane(lo)];
    const int32_t local = (int32_t)(tile - V.first_tile);
    const int32_t tx0 = (local % V.tiles_x) * TW;
    const int32_t ty0 = (local / V.tiles_x) * TH;
    const int32_t px = tx0 + (lane & (TW - 1));
    for (int32_t pass = first_pass; pass < first_pass + n_passes; pass++) {
        const int32_t row = ty0 + pass * kRowsPerPass + lane / TW;
        if (px < V.target.width && row < V.target.height) {
            float vx, vy;
            dev::pixel_vpos(V.target.width, V.target.height, V.frame.inv_w, V.frame.inv_h, px, row, vx, vy);
            dev::PixelResult p;
            dev::shade_pixel<IX>(g, V.frame, vx, vy, p, top);
            dev::store_pixel<IX>(g, V.target, (int64_t)row * V.target.width + px, p);
            rays += p.rays;
            iters += p.iters;
            pixels += 1;
        }
    }
}

// Multi-view launch scheduled by the hardware.  Every wave shades one pass of
// 64 pixels: work item i is pass i mod P of tile i / P (P = TW*TH/64), and a
// 256-thread workgroup takes four consecutive items, so its waves share the
// CU's L1 over one 16x16 tile or four adjacent smaller ones.  With one pass
// per wave no value lives across a second walk, which keeps every tile shape
// within 64 VGPRs without spilling.  The dispatcher hands out workgroups in
// order as slots free up, so the views' tiles are balanced over the device
// without any queue.
template <typename IX, int TW, int TH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_render_views(GridDesc g, const ViewDesc* __restrict__ views, int32_t n_views, int64_t total_tiles,
               unsigned long long* counters, unsigned long long* parts) {
    constexpr int32_t kPasses = TW * TH / kWave;
    const int32_t top = g.top ? *g.top : INT32_MAX;
    const int32_t lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
    uint32_t rays = 0, iters = 0, pixels = 0;
    if (item < total_tiles * kPasses)
        shade_view_tile<IX, TW, TH>(g, views, n_views, item / kPasses, (int32_t)(item % kPasses), 1, lane, top, rays,
                                    iters, pixels);
    if (parts) {
        // one pass of 64 pixels has a tile's bounds; the host passes parts
        // only when every view's render_lim < kTileSum32Lim
        const uint32_t rp = dev::wave_sum_u32(rays | (pixels << 16));
        const uint32_t it = dev::wave_sum_u32(iters);
        if (lane == 0 && item < total_tiles * kPasses)
            parts[item] = (unsigned long long)it | ((unsigned long long)rp << 32);
    } else {
        flush_counters(counters, rays, iters, pixels);
    }
}

// Persistent multi-view kernel.  The tiles are split into kXcds contiguous
// ranges with one queue counter each, on its own cache line; a wave drains
// the range of the XCD it runs on (workgroups are dispatched round-robin over
// the XCDs, so blockIdx.x mod 8) and then steals from the others.  A pop takes
// `chunk` tiles; every wave leaves once all ranges are empty.  Pops are
// device-scope atomics on one word per XCD, so they pay off only when tiles
// are large: on one 1080p frame of 8x8 tiles this kernel ran 50% slower than
// k_render_views (2.8x slower with a single counter for all XCDs).
// Held to 8 waves per SIMD, the queue loop's state spills about 28 SGPRs
// into VGPR lanes and one VGPR to scratch; at the compiler's own budget (67
// VGPRs, 7 waves, no spill) the 64-view set took 11.7 ms instead of 10.4.
template <typename IX, int TW, int TH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_render_batch(GridDesc g, const ViewDesc* __restrict__ views, int32_t n_views, int64_t total_tiles, int32_t chunk,
               unsigned int* queue, unsigned long long* counters) {
    constexpr uint32_t kPasses = TW * TH / kWave;
    const int32_t top = g.top ? *g.top : INT32_MAX;
    const int32_t lane = threadIdx.x & 63;
    uint32_t rays = 0, iters = 0, pixels = 0;
    const uint32_t total = (uint32_t)total_tiles;  // below 2^32 (host check)
    const uint32_t home = blockIdx.x % kXcds;
    for (uint32_t k = 0; k < kXcds; k++) {
        const uint32_t q = (home + k) % kXcds;
        const uint32_t r0 = (uint32_t)((uint64_t)total * q / kXcds), r1 = (uint32_t)((uint64_t)total * (q + 1) / kXcds);
        for (;;) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(queue + q * kQueueStride, (unsigned int)chunk);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base >= r1 - r0) break;
            const uint32_t end = r1 - r0 - base > (uint32_t)chunk ? r0 + base + (uint32_t)chunk : r1;
            for (uint64_t item = (uint64_t)(r0 + base) * kPasses; item < (uint64_t)end * kPasses; item++)
                shade_view_tile<IX, TW, TH>(g, views, n_views, (int64_t)(item / kPasses), (int32_t)(item % kPasses), 1,
                                            lane, top, rays, iters, pixels);
        }
    }
    flush_counters(counters, rays, iters, pixels);
}

// div_by against IEEE division for a = 1 + i/2^23 (all i, one per thread)
// and d = 1 + j/2^23 for j in [d_begin, d_end).
__global__ void __launch_bounds__(256) k_selftest_div(uint32_t d_begin, uint32_t d_end,
                                                       unsigned long long* bad) {
    const uint32_t ai = blockIdx.x * 256u + threadIdx.x;
    const float a = __uint_as_float(0x3f800000u | ai);
    uint32_t n = 0;
    const uint32_t j0 = d_begin + blockIdx.y * 64u;
    for (uint32_t k = 0; k < 64u; k++) {
        const uint32_t j = j0 + k;
        if (j >= d_end) break;
        const float d = __uint_as_float(0x3f800000u | j);
        const float y = 1.0f / d;
        n += __float_as_uint(dev::div_by(a, d, y, true)) != __float_as_uint(a / d);
    }
    if (n) atomicAdd(bad, (unsigned long long)n);
}

// rcp_rn against IEEE division 1/d for the float bit patterns
// [first, first + count), one pattern per thread and 16 per thread in turn.
__global__ void __launch_bounds__(256) k_selftest_rcp(uint64_t first, uint64_t end, unsigned long long* bad) {
    uint32_t n = 0;
    for (uint32_t k = 0; k < 16u; k++) {
        const uint64_t i = first + ((uint64_t)blockIdx.x * 16u + k) * 256u + threadIdx.x;
        if (i >= end) break;
        const float d = __uint_as_float((uint32_t)i);
        const float r = 1.0f / d;
        const float y = dev::rcp_rn(d);
        // any NaN equals any NaN (both come from the same division)
        n += (__float_as_uint(y) != __float_as_uint(r)) & !((r != r) & (y != y));
    }
    if (n) atomicAdd(bad, (unsigned long long)n);
}

__global__ void k_queue_reset(unsigned int* q) {
    if (threadIdx.x < kXcds) q[threadIdx.x * kQueueStride] = 0u;
}

// Calls f with the index form the grid admits (IdxB, Idx32 or Idx64).
template <typename F>
void with_index(const GridDesc& g, F&& f) {
    if (g.bytes32()) f(dev::IdxB{});
    else if (g.small()) f(dev::Idx32{});
    else f(dev::Idx64{});
}

}  // namespace

int launch_render_plain(const GridDesc& g, const FrameDesc& f, const vrt_target& t, unsigned long long* counters,
                        unsigned long long* parts, hipStream_t s) {
    const int64_t tiles = (int64_t)((t.width + 7) / 8) * ((t.height + 7) / 8);
    const int64_t blocks = (tiles + kPlainWaves - 1) / kPlainWaves;
    if (blocks <= 0) return VRT_OK;
    if (blocks > 0x7fffffff) return fail(VRT_ERR_INVALID, "frame too large");
    with_index(g, [&](auto ix) {
        hipLaunchKernelGGL(k_render_plain<decltype(ix)>, dim3((uint32_t)blocks), dim3(kPlainThreads), 0, s, g, f, t,
                           counters, parts);
    });
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_fold_parts(const unsigned long long* parts, int64_t n, unsigned long long* counters, hipStream_t s) {
    if (n <= 0) return VRT_OK;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) return fail(VRT_ERR_INVALID, "too many partial counts");
    hipLaunchKernelGGL(k_fold_parts, dim3((uint32_t)blocks), dim3(256), 0, s, parts, n, counters);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_cast_rays(const GridDesc& g, int64_t n, const float* start, const float* dir, const int32_t* dist,
                     int32_t render_dist, int64_t* hit, float* hit_pos, uint8_t* face, uint32_t* steps,
                     hipStream_t s) {
    if (n <= 0) return VRT_OK;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) return fail(VRT_ERR_INVALID, "too many rays");
    with_index(g, [&](auto ix) {
        hipLaunchKernelGGL(k_cast_rays<decltype(ix)>, dim3((uint32_t)blocks), dim3(256), 0, s, g, n, start, dir, dist,
                           (float)render_dist, hit, hit_pos, face, steps);
    });
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

// Calls f with the tile shape as template arguments (check_batch_config's set).
template <typename F>
int with_shape(int32_t tw, int32_t th, F&& f) {
    if (tw == 8 && th == 8) f(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{});
    else if (tw == 16 && th == 4) f(std::integral_constant<int, 16>{}, std::integral_constant<int, 4>{});
    else if (tw == 32 && th == 2) f(std::integral_constant<int, 32>{}, std::integral_constant<int, 2>{});
    else if (tw == 16 && th == 16) f(std::integral_constant<int, 16>{}, std::integral_constant<int, 16>{});
    else if (tw == 64 && th == 1) f(std::integral_constant<int, 64>{}, std::integral_constant<int, 1>{});
    else return fail(VRT_ERR_INVALID, "unsupported tile shape");
    return VRT_OK;
}

int batch_occupancy(const vrt_batch_config& cfg, int32_t* waves_per_cu) {
    (void)cfg;
    int blocks = 0;
    VRT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_batch<dev::Idx64, 8, 8>, 256, 0));
    *waves_per_cu = blocks * 4;
    return VRT_OK;
}

int launch_render_batch(const GridDesc& g, const ViewDesc* views_dev, int32_t n_views, int64_t total_tiles,
                        const vrt_batch_config& cfg, int32_t n_workgroups, unsigned int* queue,
                        unsigned long long* counters, unsigned long long* parts, hipStream_t s) {
    if (parts && cfg.chunk_tiles != 0) return fail(VRT_ERR_INVALID, "partial counts need the hardware-scheduled kernel");
    if (cfg.chunk_tiles == 0) {
        const int64_t wgs = (total_tiles * (cfg.tile_w * cfg.tile_h / 64) + 3) / 4;  // one wave per 64 pixels
        if (wgs <= 0) return VRT_OK;
        if (wgs > 0x7fffffff) return fail(VRT_ERR_INVALID, "too many tiles in one batch");
        if (int rc = with_shape(cfg.tile_w, cfg.tile_h, [&](auto tw, auto th) {
                with_index(g, [&](auto ix) {
                    hipLaunchKernelGGL((k_render_views<decltype(ix), decltype(tw)::value, decltype(th)::value>),
                                       dim3((uint32_t)wgs), dim3(256), 0, s, g, views_dev, n_views, total_tiles,
                                       counters, parts);
                });
            }))
            return rc;
        VRT_HIP(hipGetLastError());
        return VRT_OK;
    }
    if (n_workgroups <= 0) return fail(VRT_ERR_INVALID, "no workgroups");
    if (int rc = with_shape(cfg.tile_w, cfg.tile_h, [&](auto tw, auto th) {
            with_index(g, [&](auto ix) {
                hipLaunchKernelGGL((k_render_batch<decltype(ix), decltype(tw)::value, decltype(th)::value>),
                                   dim3((uint32_t)n_workgroups), dim3(256), 0, s, g, views_dev, n_views, total_tiles,
                                   cfg.chunk_tiles, queue, counters);
            });
        }))
        return rc;
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int selftest_division(uint32_t d_begin, uint32_t d_count, unsigned long long* bad_dev, hipStream_t s) {
    const uint32_t end = (uint32_t)std::min<uint64_t>((uint64_t)d_begin + d_count, 1u << 23);
    for (uint32_t j = d_begin; j < end; j += 64u * 4096u) {
        const uint32_t e = std::min<uint32_t>(end, j + 64u * 4096u);
        dim3 grid((1u << 23) / 256u, (e - j + 63u) / 64u);
        hipLaunchKernelGGL(k_selftest_div, grid, dim3(256), 0, s, j, e, bad_dev);
        VRT_HIP(hipGetLastError());
    }
    return VRT_OK;
}

int selftest_reciprocal(uint64_t first, uint64_t count, unsigned long long* bad_dev, hipStream_t s) {
    const uint64_t end = std::min<uint64_t>(first + count, 1ull << 32);
    const uint64_t per_launch = 16ull * 256ull * 65536ull;  // 2^28 patterns
    for (uint64_t j = first; j < end; j += per_launch) {
        const uint64_t e = std::min<uint64_t>(end, j + per_launch);
        const uint32_t blocks = (uint32_t)((e - j + 16ull * 256ull - 1) / (16ull * 256ull));
        hipLaunchKernelGGL(k_selftest_rcp, dim3(blocks), dim3(256), 0, s, j, e, bad_dev);
        VRT_HIP(hipGetLastError());
    }
    return VRT_OK;
}

int launch_queue_reset(unsigned int* queue, hipStream_t s) {
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, queue);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

}  // namespace vrt
