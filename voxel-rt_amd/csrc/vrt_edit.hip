// vrt_edit.hip — depth field and voxel edits on the device.
//
// fixDepthField (src/render.cpp:226-253) gives an empty voxel (word < 0) the
// largest table distance d* over its SOLID neighbours with d* > -6 (strict,
// :240), and then: d* > -2 -> the word is left as it was (:244-249);
// -6 < d* <= -2 -> the bits of d*; no such neighbour -> the bits of -6.0f.
// The result depends only on which neighbours are solid, and a jump word is
// negative (never solid), so every voxel of a region can be fixed in
// parallel, in place, with the same result as the reference's sequential
// loops.  The table is scanned in order of decreasing distance, so the first
// solid neighbour found is d*.  For whole-grid builds a workgroup owns one
// 8x8x8 brick and first checks the 3x3x3 neighbouring bricks' solid flags:
// with none solid, every empty voxel of the brick takes -6 without a scan;
// otherwise the scan reads the neighbourhood's solidity from LDS.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vrt_internal.h"

namespace vrt {

namespace {

constexpr int32_t kMinus6Bits = (int32_t)0xC0C00000;  // bits of -6.0f

// Solidity bitmap of the grid: row (y, z) is W words, bit x + 32 of the row
// is voxel x (bits of x < 0 and x >= X are 0), so any 22-bit window starting
// at x >= -7 lies in two consecutive words.  One wave per 64 voxels of a row.
__global__ void __launch_bounds__(256) k_solid_bits(GridDesc g, uint32_t* bits, uint32_t W, uint32_t chunks) {
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t row = wave / chunks;  // y + Y * z
    if (row >= (uint64_t)g.Y * g.Z) return;
    const uint32_t j = (uint32_t)(wave % chunks);
    const uint32_t x = j * 64 + lane, y = (uint32_t)(row % g.Y), z = (uint32_t)(row / g.Y);
    bool solid = false;
    if (x < g.X) solid = g.vox[(uint64_t)x + g.X * (uint64_t)y + g.XY * (uint64_t)z] >= 0;
    const uint64_t m = __ballot(solid);
    uint32_t* r = bits + row * W;
    if (lane == 0) r[1 + 2 * j] = (uint32_t)m;
    if (lane == 32) r[2 + 2 * j] = (uint32_t)(m >> 32);
}

// Any solid voxel in each 8x8x8 brick, from the bitmap: one wave per brick,
// one lane per (y, z) row of the brick, 8 bits each.
__global__ void __launch_bounds__(256) k_brick_flags_bits(GridDesc g, const uint32_t* bits, uint32_t W,
                                                           uint8_t* flags, uint32_t bx, uint32_t by, uint64_t nb) {
    const uint64_t b = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nb) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ix = (uint32_t)(b % bx), iy = (uint32_t)((b / bx) % by), iz = (uint32_t)(b / ((uint64_t)bx * by));
    const uint32_t y = iy * 8 + (lane & 7), z = iz * 8 + (lane >> 3);
    uint32_t v = 0;
    if (y < g.Y && z < g.Z) {
        const uint32_t bit = ix * 8 + 32;
        v = (bits[((uint64_t)y + (uint64_t)g.Y * z) * W + (bit >> 5)] >> (bit & 31)) & 0xffu;
    }
    const bool any = __ballot(v != 0u) != 0;
    if (lane == 0) flags[b] = (uint8_t)any;
}

// Whole-grid build, one workgroup per 8x8x8 brick.  A solid neighbour within
// radius 7 lies in the 3x3x3 bricks around this one; with none, every empty
// voxel takes -6.  Otherwise the solidity of the brick's 22x22x22
// neighbourhood (+-7) is loaded once into LDS as one 22-bit word per x-row
// (a ballot per row), and each voxel scans the sorted table against those
// bits instead of gathering the grid per entry.  Writing jump words never
// changes solidity, so the snapshot stays exact while the grid is written.
constexpr int32_t kHalo = 7;
constexpr int32_t kSpan = 8 + 2 * kHalo;  // 22 cells per row of the neighbourhood

// fixDepthField's result for the voxel at (lx, ly, lz) of a brick whose
// 22x22x22 neighbourhood solidity is in `rows` (bit b of row (ry, rz) is
// x = lx - 7 + b).  The table is grouped by (distance, dy, dz): the first
// group with a solid neighbour gives d*, as the first solid entry of the
// sorted table would, and the group's dx mask (bit dx + 7) lines up with the
// row shifted right by lx.  Eight groups per step keep eight LDS reads in
// flight.
__device__ __forceinline__ int32_t scan_groups(const uint32_t* rows, const DepthTableDev& t, int32_t lx, int32_t ly,
                                               int32_t lz, int32_t old) {
    const int32_t my_row = (lz + kHalo) * kSpan + ly + kHalo;
    auto hits = [&](int32_t r) {
        const int32_t dy = ((r >> 16) & 31) - 8, dz = ((r >> 21) & 31) - 8;
        return ((rows[my_row + dz * kSpan + dy] >> lx) & (uint32_t)(r & 0x7fff)) != 0u;
    };
    int32_t found = -1;
    int32_t k = 0;
    for (; k + 8 <= t.n_rows; k += 8) {
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) m |= (uint32_t)hits(t.rows[k + j]) << j;
        if (m) {
            found = k + __builtin_ctz(m);
            break;
        }
    }
    if (found < 0)
        for (; k < t.n_rows; k++)
            if (hits(t.rows[k])) {
                found = k;
                break;
            }
    if (found < 0) return kMinus6Bits;
    return found < t.n_rows_near ? old : __float_as_int(t.rows_dist[found]);
}

// The neighbourhood solidity of the brick at (bx0, by0, bz0) read from the
// grid itself: thread r < 484 gathers the 22 voxels of row r (independent
// loads, all in flight) into one word.
__device__ __forceinline__ void fill_rows_from_grid(const GridDesc& g, uint32_t* rows, int32_t bx0, int32_t by0,
                                                    int32_t bz0, int32_t c) {
    if (c >= kSpan * kSpan) return;
    const int32_t gy = by0 - kHalo + c % kSpan, gz = bz0 - kHalo + c / kSpan;
    uint32_t v = 0;
    if (((uint32_t)gy < g.Y) & ((uint32_t)gz < g.Z)) {
        const int32_t* row = g.vox + (uint64_t)g.X * (uint64_t)gy + g.XY * (uint64_t)gz;
#pragma unroll
        for (int32_t b = 0; b < kSpan; b++) {
            const int32_t gx = bx0 - kHalo + b;
            if ((uint32_t)gx < g.X && row[gx] >= 0) v |= 1u << b;
        }
    }
    rows[c] = v;
}

__global__ void __launch_bounds__(512) k_fix_depth_bricks(GridDesc g, DepthTableDev t, const uint8_t* flags,
                                                           const uint32_t* bits, uint32_t W, uint32_t bx,
                                                           uint32_t by, uint32_t bz) {
    __shared__ uint32_t rows[kSpan * kSpan];
    const uint64_t b = blockIdx.x;
    const int32_t ix = (int32_t)(b % bx), iy = (int32_t)((b / bx) % by), iz = (int32_t)(b / ((uint64_t)bx * by));
    bool near_solid = false;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int32_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
                if (jx >= 0 && jy >= 0 && jz >= 0 && jx < (int32_t)bx && jy < (int32_t)by && jz < (int32_t)bz)
                    near_solid |= flags[(uint64_t)jx + (uint64_t)bx * jy + (uint64_t)bx * by * jz] != 0;
            }
    const int32_t c = threadIdx.x;
    const int32_t x = ix * 8 + (c & 7), y = iy * 8 + ((c >> 3) & 7), z = iz * 8 + (c >> 6);
    const bool in_grid = ((uint32_t)x < g.X) & ((uint32_t)y < g.Y) & ((uint32_t)z < g.Z);
    const uint64_t i = (uint64_t)x + g.X * (uint64_t)y + g.XY * (uint64_t)z;
    const int32_t old = in_grid ? g.vox[i] : 0;
    if (!near_solid) {
        if (in_grid && old < 0 && old != kMinus6Bits) g.vox[i] = kMinus6Bits;
        return;
    }
    if (!__syncthreads_or(in_grid && old < 0)) return;  // no empty voxel in the brick
    // neighbourhood solidity: thread r takes row r, the 22 bits from x = ox
    // (bit ox + 32 >= 25 of the bitmap row) out of two words
    const int32_t ox = ix * 8 - kHalo, oy = iy * 8 - kHalo, oz = iz * 8 - kHalo;
    if (c < kSpan * kSpan) {
        const int32_t gy = oy + c % kSpan, gz = oz + c / kSpan;
        uint32_t v = 0;
        if (((uint32_t)gy < g.Y) & ((uint32_t)gz < g.Z)) {
            const uint32_t* rw = bits + ((uint64_t)gy + (uint64_t)g.Y * (uint64_t)gz) * W;
            const uint32_t b0 = (uint32_t)(ox + 32);
            const uint64_t two = (uint64_t)rw[b0 >> 5] | ((uint64_t)rw[(b0 >> 5) + 1] << 32);
            v = (uint32_t)(two >> (b0 & 31)) & ((1u << kSpan) - 1u);
        }
        rows[c] = v;
    }
    __syncthreads();
    if (!in_grid || old >= 0) return;
    const int32_t nv = scan_groups(rows, t, c & 7, (c >> 3) & 7, c >> 6, old);
    if (nv != old) g.vox[i] = nv;
}

// fixDepthField over the box [x0,x1) x [y0,y1) x [z0,z1) (already clipped);
// with sr >= 0 only voxels with (x-cx)^2+(y-cy)^2+(z-cz)^2 < sr^2.  The box
// is tiled into 8x8x8 bricks from its corner, one workgroup each: the
// brick's neighbourhood solidity is gathered from the grid into LDS and its
// voxels scan the grouped table against it.
__global__ void __launch_bounds__(512) k_fix_depth_box(GridDesc g, DepthTableDev t, int32_t x0, int32_t y0,
                                                        int32_t z0, int32_t w, int32_t h, int32_t d, int32_t nbx,
                                                        int32_t nby, int32_t cx, int32_t cy, int32_t cz, int32_t sr) {
    __shared__ uint32_t rows[kSpan * kSpan];
    const int32_t b = (int32_t)blockIdx.x;
    const int32_t bx0 = x0 + (b % nbx) * 8, by0 = y0 + ((b / nbx) % nby) * 8, bz0 = z0 + (b / (nbx * nby)) * 8;
    const int32_t c = threadIdx.x;
    const int32_t lx = c & 7, ly = (c >> 3) & 7, lz = c >> 6;
    const int32_t x = bx0 + lx, y = by0 + ly, z = bz0 + lz;
    bool mine = (x < x0 + w) & (y < y0 + h) & (z < z0 + d);
    if (sr >= 0) {
        const int32_t ex = x - cx, ey = y - cy, ez = z - cz;
        mine &= ex * ex + ey * ey + ez * ez < sr * sr;
    }
    const uint64_t i = (uint64_t)x + g.X * (uint64_t)y + g.XY * (uint64_t)z;
    const int32_t old = mine ? g.vox[i] : 0;
    if (!__syncthreads_or(mine && old < 0)) return;  // nothing to fix in this brick
    fill_rows_from_grid(g, rows, bx0, by0, bz0, c);
    __syncthreads();
    if (!mine || old >= 0) return;
    const int32_t nv = scan_groups(rows, t, lx, ly, lz, old);
    if (nv != old) g.vox[i] = nv;
}

// destroyVoxel over the clipped box of removeSphere's first loop
// (src/level.cpp:31-42): (x-cx)^2+(y-cy)^2+(z-cz)^2 < r^2.
__global__ void __launch_bounds__(256) k_carve(GridDesc g, int32_t x0, int32_t y0, int32_t z0, int32_t w, int32_t h,
                                                int32_t d, int32_t cx, int32_t cy, int32_t cz, int32_t r) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)w * h * d) return;
    const int32_t x = x0 + (int32_t)(k % w);
    const int32_t y = y0 + (int32_t)((k / w) % h);
    const int32_t z = z0 + (int32_t)(k / ((int64_t)w * h));
    const int32_t ex = x - cx, ey = y - cy, ez = z - cz;
    if (ex * ex + ey * ey + ez * ez >= r * r) return;
    g.vox[(uint64_t)x + g.X * (uint64_t)y + g.XY * (uint64_t)z] = -1;
}

__global__ void k_set_voxel(int32_t* vox, uint64_t i, int32_t v, int32_t* top, int32_t y) {
    if (threadIdx.x == 0) {
        vox[i] = v;
        if (v >= 0 && top) atomicMax(top, y);
    }
}

__global__ void k_set_word(int32_t* p, int32_t v) {
    if (threadIdx.x == 0) *p = v;
}

// Highest y with a solid voxel over rows (y, z) of the box [x0, x0+w) x
// [y0, y0+h) x [z0, z0+d): a wave reads a row's words with its 64 lanes, keeps
// the highest y of its rows with any solid word and raises *top once.
__global__ void __launch_bounds__(256) k_solid_top(GridDesc g, int32_t x0, int32_t w, int32_t y0, int32_t h,
                                                    int32_t z0, int64_t rows, int32_t* top) {
    const int32_t lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int32_t best = -1;
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += waves) {
        const int32_t y = y0 + (int32_t)(r % h);
        const int32_t z = z0 + (int32_t)(r / h);
        if (y <= best) continue;
        const int32_t* row = g.vox + (uint64_t)g.X * (uint64_t)y + g.XY * (uint64_t)z + (uint64_t)x0;
        bool any = false;
        for (int32_t x = lane; x < w; x += 64) any |= row[x] >= 0;
        if (__builtin_amdgcn_ballot_w64(any) != 0) best = y;
    }
    if (lane == 0 && best >= 0) atomicMax(top, best);
}

// Grid copy, 16 bytes per lane per step (a plain streaming copy of known size).
__global__ void __launch_bounds__(256) k_copy(int4* __restrict__ dst, const int4* __restrict__ src, uint64_t n16,
                                               int32_t* __restrict__ dst_tail, const int32_t* __restrict__ src_tail,
                                               uint32_t tail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t; i < n16; i += stride) dst[i] = src[i];
    if (t < tail) dst_tail[t] = src_tail[t];
}

// One wave per x-row segment of the box (four rows per workgroup, the y and z
// loops covering boxes taller or deeper than the launch); each lane moves the
// words lane, lane+64, ... of its row, four loads in flight.
__global__ void __launch_bounds__(256) k_copy_box(GridDesc src, GridDesc dst, int64_t x0, int64_t y0, int64_t z0,
                                                  uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t z = blockIdx.z; z < nz; z += gridDim.z) {
        for (uint32_t y = blockIdx.y * 4 + w; y < ny; y += gridDim.y * 4) {
            const uint64_t gy = (uint64_t)(y0 + y), gz = (uint64_t)(z0 + z);
            const int32_t* __restrict__ s = src.vox + (src.origin + (uint64_t)x0 + src.SX * gy + src.SXY * gz);
            int32_t* __restrict__ d = dst.vox + (dst.origin + (uint64_t)x0 + dst.SX * gy + dst.SXY * gz);
            uint32_t x = lane;
            for (; x + 192 < nx; x += 256) {
                const int32_t a = s[x], b = s[x + 64], c = s[x + 128], e = s[x + 192];
                d[x] = a;
                d[x + 64] = b;
                d[x + 128] = c;
                d[x + 192] = e;
            }
            for (; x < nx; x += 64) d[x] = s[x];
        }
    }
}

__global__ void __launch_bounds__(256) k_fill(int32_t* p, uint64_t n, int32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Clip [lo, hi) to [0, n).
inline void clip(int64_t& lo, int64_t& hi, int64_t n) {
    if (lo < 0) lo = 0;
    if (hi > n) hi = n;
}

}  // namespace

int launch_fix_depth_box(const GridDesc& g, const DepthTableDev& t, int32_t x0, int32_t y0, int32_t z0, int32_t x1,
                         int32_t y1, int32_t z1, int32_t scx, int32_t scy, int32_t scz, int32_t sr, hipStream_t s) {
    int64_t ax = x0, bx_ = x1, ay = y0, by_ = y1, az = z0, bz_ = z1;
    clip(ax, bx_, g.X);
    clip(ay, by_, g.Y);
    clip(az, bz_, g.Z);
    if (bx_ <= ax || by_ <= ay || bz_ <= az) return VRT_OK;
    const int64_t w = bx_ - ax, h = by_ - ay, d = bz_ - az;
    const int64_t nbx = (w + 7) / 8, nby = (h + 7) / 8, nbz = (d + 7) / 8;
    const int64_t blocks = nbx * nby * nbz;
    if (blocks > 0x7fffffff) return fail(VRT_ERR_INVALID, "box too large");
    hipLaunchKernelGGL(k_fix_depth_box, dim3((uint32_t)blocks), dim3(512), 0, s, g, t, (int32_t)ax, (int32_t)ay,
                       (int32_t)az, (int32_t)w, (int32_t)h, (int32_t)d, (int32_t)nbx, (int32_t)nby, scx, scy, scz,
                       sr);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

uint32_t solid_bits_words(const GridDesc& g) { return (g.X + 64u + 31u) / 32u; }

int launch_build_depth_field(const GridDesc& g, const DepthTableDev& t, uint8_t* bricks, uint32_t* bits,
                             hipStream_t s) {
    const uint32_t W = solid_bits_words(g);
    const uint64_t rows = (uint64_t)g.Y * g.Z;
    VRT_HIP(hipMemsetAsync(bits, 0, rows * W * sizeof(uint32_t), s));
    const uint32_t chunks = (g.X + 63u) / 64u;
    const uint64_t bit_blocks = (rows * chunks + 3) / 4;
    const uint32_t nbx = (g.X + 7) / 8, nby = (g.Y + 7) / 8, nbz = (g.Z + 7) / 8;
    const uint64_t nb = (uint64_t)nbx * nby * nbz;
    if (bit_blocks > 0x7fffffffull || nb > 0x7fffffffull) return fail(VRT_ERR_INVALID, "grid too large for the build");
    hipLaunchKernelGGL(k_solid_bits, dim3((uint32_t)bit_blocks), dim3(256), 0, s, g, bits, W, chunks);
    VRT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_brick_flags_bits, dim3((uint32_t)((nb + 3) / 4)), dim3(256), 0, s, g, bits, W, bricks, nbx,
                       nby, nb);
    VRT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fix_depth_bricks, dim3((uint32_t)nb), dim3(512), 0, s, g, t, bricks, bits, W, nbx, nby, nbz);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_carve_sphere(const GridDesc& g, int32_t cx, int32_t cy, int32_t cz, int32_t r, hipStream_t s) {
    if (r <= 0) return VRT_OK;
    int64_t ax = (int64_t)cx - r, bx_ = (int64_t)cx + r, ay = (int64_t)cy - r, by_ = (int64_t)cy + r,
            az = (int64_t)cz - r, bz_ = (int64_t)cz + r;
    clip(ax, bx_, g.X);
    clip(ay, by_, g.Y);
    clip(az, bz_, g.Z);
    if (bx_ <= ax || by_ <= ay || bz_ <= az) return VRT_OK;
    const int64_t w = bx_ - ax, h = by_ - ay, d = bz_ - az;
    const int64_t blocks = (w * h * d + 255) / 256;
    if (blocks > 0x7fffffff) return fail(VRT_ERR_INVALID, "sphere too large");
    hipLaunchKernelGGL(k_carve, dim3((uint32_t)blocks), dim3(256), 0, s, g, (int32_t)ax, (int32_t)ay, (int32_t)az,
                       (int32_t)w, (int32_t)h, (int32_t)d, cx, cy, cz, r);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_set_voxel(const GridDesc& g, int64_t index, int32_t y, int32_t v, hipStream_t s) {
    if (index < 0 || (uint64_t)index >= g.count) return fail(VRT_ERR_INVALID, "voxel index out of range");
    hipLaunchKernelGGL(k_set_voxel, dim3(1), dim3(64), 0, s, g.vox, (uint64_t)index, v, g.top, y);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_solid_top(const GridDesc& g, int64_t x0, int64_t y0, int64_t z0, int64_t x1, int64_t y1, int64_t z1,
                     bool reset, hipStream_t s) {
    if (!g.top) return fail(VRT_ERR_INVALID, "grid without a solid-top word");
    if (reset) {
        hipLaunchKernelGGL(k_set_word, dim3(1), dim3(64), 0, s, g.top, -1);
        VRT_HIP(hipGetLastError());
    }
    clip(x0, x1, g.X);
    clip(y0, y1, g.Y);
    clip(z0, z1, g.Z);
    if (x1 <= x0 || y1 <= y0 || z1 <= z0) return VRT_OK;
    const int64_t rows = (y1 - y0) * (z1 - z0);
    int64_t blocks = (rows + 3) / 4;  // 4 waves per block, one row each to start
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_solid_top, dim3((uint32_t)blocks), dim3(256), 0, s, g, (int32_t)x0, (int32_t)(x1 - x0),
                       (int32_t)y0, (int32_t)(y1 - y0), (int32_t)z0, rows, g.top);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_copy(int32_t* dst, const int32_t* src, uint64_t n, hipStream_t s) {
    if (n == 0) return VRT_OK;
    if (((uintptr_t)dst | (uintptr_t)src) & 15) return fail(VRT_ERR_INVALID, "copy buffers must be 16-byte aligned");
    const uint64_t n16 = n / 4;
    const uint32_t tail = (uint32_t)(n % 4);
    uint64_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, s, (int4*)dst, (const int4*)src, n16,
                       dst + 4 * n16, src + 4 * n16, tail);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_copy_box(const GridDesc& src, const GridDesc& dst, int64_t x0, int64_t y0, int64_t z0, int64_t nx,
                    int64_t ny, int64_t nz, hipStream_t s) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return VRT_OK;
    if (x0 < 0 || y0 < 0 || z0 < 0 || x0 + nx > src.X || y0 + ny > src.Y || z0 + nz > src.Z || src.X != dst.X ||
        src.Y != dst.Y || src.Z != dst.Z)
        return fail(VRT_ERR_INVALID, "copy box outside the grid");
    const uint32_t by = (uint32_t)std::min<int64_t>((ny + 3) / 4, 65535), bz = (uint32_t)std::min<int64_t>(nz, 65535);
    hipLaunchKernelGGL(k_copy_box, dim3(1, by, bz), dim3(256), 0, s, src, dst, x0, y0, z0, (uint32_t)nx,
                       (uint32_t)ny, (uint32_t)nz);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

int launch_fill(int32_t* p, uint64_t n, int32_t v, hipStream_t s) {
    if (n == 0) return VRT_OK;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536ull * 16) blocks = 65536ull * 16;
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    VRT_HIP(hipGetLastError());
    return VRT_OK;
}

}  // namespace vrt
