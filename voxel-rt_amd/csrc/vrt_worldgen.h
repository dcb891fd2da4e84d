// vrt_worldgen.h — the synthetic world as a pure per-column function, used by
// both generators in vrt_world.hip (host threads over z-slabs and one device
// thread per column) so both write the same words.  Integer arithmetic only.
//
// World: value-noise heightfield (octaves of lattice noise, amplitudes
// 2^(octaves-1) .. 1, smoothstep interpolation in Q16) giving heights in
// [height_min, height_max]; solid below the height with colour bands by depth
// under the surface (grass 1 layer, dirt 3 layers, stone) plus a per-voxel
// jitter of +-jitter on all three channels; spherical caves carved out; then
// axis-aligned boxes filled solid (overhangs).  Voxel words use the
// reference's packing (r<<16)|(g<<8)|b, empty = -1 (src/level.cpp:13-19).
#pragma once

#include <cstdint>

#define VRT_HD __host__ __device__ __forceinline__

namespace vrt {
namespace wg {

VRT_HD uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

VRT_HD uint32_t hash4(uint32_t a, uint32_t b, uint32_t c, uint64_t seed) {
    uint32_t h = mix32((uint32_t)seed ^ 0x9e3779b9u);
    h = mix32(h ^ (a * 0x85ebca6bu));
    h = mix32(h ^ (b * 0xc2b2ae35u));
    h = mix32(h ^ (c * 0x27d4eb2fu) ^ (uint32_t)(seed >> 32));
    return h;
}

VRT_HD int64_t lattice(int64_t ix, int64_t iz, int32_t octave, uint64_t seed) {
    return (int64_t)(hash4((uint32_t)ix, (uint32_t)iz, (uint32_t)octave + 101u, seed) & 0xFFFFu);
}

// Smoothstep weight in Q16 of t in Q16.
VRT_HD int64_t smooth_q16(int64_t t) { return (t * t * (3 * 65536 - 2 * t)) >> 32; }

VRT_HD int64_t lerp_q16(int64_t a, int64_t b, int64_t w) { return a + (((b - a) * w) >> 16); }

// Terrain surface: cells y < height are solid before caves/boxes.
VRT_HD int32_t surface_height(int64_t x, int64_t z, uint64_t seed, int32_t octaves, int32_t base_cell,
                              int32_t hmin, int32_t hmax) {
    int64_t sum = 0;
    for (int32_t o = 0; o < octaves; o++) {
        int64_t s = (int64_t)base_cell >> o;
        if (s < 1) s = 1;
        int64_t ix = x / s, iz = z / s;
        int64_t wx = smooth_q16(((x - ix * s) << 16) / s);
        int64_t wz = smooth_q16(((z - iz * s) << 16) / s);
        int64_t a = lerp_q16(lattice(ix, iz, o, seed), lattice(ix + 1, iz, o, seed), wx);
        int64_t b = lerp_q16(lattice(ix, iz + 1, o, seed), lattice(ix + 1, iz + 1, o, seed), wx);
        int64_t v = lerp_q16(a, b, wz);
        sum += v << (octaves - 1 - o);
    }
    int64_t amp = ((int64_t)1 << octaves) - 1;
    return hmin + (int32_t)((sum * (int64_t)(hmax - hmin + 1)) / (amp * 65536));
}

VRT_HD int32_t clamp255(int32_t v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

VRT_HD int32_t pack_rgb(int32_t r, int32_t g, int32_t b, int32_t j) {
    return (clamp255(r + j) << 16) | (clamp255(g + j) << 8) | clamp255(b + j);
}

VRT_HD int32_t jitter_of(int64_t x, int64_t y, int64_t z, uint64_t seed, int32_t jitter) {
    if (jitter <= 0) return 0;
    uint32_t h = hash4((uint32_t)x, (uint32_t)y, (uint32_t)z, seed ^ 0x5bd1e9955bd1e995ull);
    return (int32_t)(h % (uint32_t)(2 * jitter + 1)) - jitter;
}

// Terrain colour of a solid cell `depth` cells below the surface layer.
VRT_HD int32_t terrain_colour(int32_t depth, int32_t j) {
    if (depth == 0) return pack_rgb(70, 150, 40, j);    // grass
    if (depth < 4) return pack_rgb(120, 85, 50, j);     // dirt
    return pack_rgb(115, 115, 120, j);                  // stone
}

VRT_HD int32_t box_colour(int32_t j) { return pack_rgb(150, 75, 60, j); }

// Integer square root (floor) for non-negative v < 2^40.
VRT_HD int64_t isqrt64(int64_t v) {
    if (v <= 0) return 0;
    int64_t r = 0;
    int64_t bit = (int64_t)1 << 40;
    while (bit > v) bit >>= 2;
    while (bit) {
        if (v >= r + bit) {
            v -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
        bit >>= 2;
    }
    return r;
}

// y-interval [lo, hi] (inclusive) where a cave sphere crosses column (x, z);
// returns false when it misses the column.  Cave cells satisfy
// (x-cx)^2 + (y-cy)^2 + (z-cz)^2 <= r^2.
VRT_HD bool cave_span(const int32_t* c, int64_t x, int64_t z, int64_t& lo, int64_t& hi) {
    int64_t dx = x - c[0], dz = z - c[2], r = c[3];
    int64_t rem = r * r - dx * dx - dz * dz;
    if (rem < 0) return false;
    int64_t h = isqrt64(rem);
    lo = c[1] - h;
    hi = c[1] + h;
    return true;
}

VRT_HD bool box_span(const int32_t* b, int64_t x, int64_t z, int64_t& lo, int64_t& hi) {
    if (x < b[0] || x >= b[3] || z < b[2] || z >= b[5]) return false;
    lo = b[1];
    hi = b[4] - 1;
    return true;
}

}  // namespace wg
}  // namespace vrt
