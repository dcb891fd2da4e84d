// vrt_device.h — device-side pieces of the per-pixel path of
// src/fshader.glsl (castRay :59-129 and main :131-190) for gfx950.
//
// Every float operation below is one IEEE binary32 operation in the order
// the shader states it (the library is built with -ffp-contract=off, correctly
// rounded division and square root, denormals kept), so the hit index, face,
// shadow bits and step counts equal a literal scalar restatement of the
// shader.  The code is arranged for wave64 SIMT: one ray per lane, the axis
// choice as selects, the voxel index kept incrementally, and the shadow rays
// of a pixel cast one after another by the same lane.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vrt_internal.h"

namespace vrt {
namespace dev {

// GLSL int(float): truncation toward zero; out-of-range saturates, NaN -> 0.
// That is exactly what v_cvt_i32_f32 does, so it is issued directly (C++'s
// (int) of an out-of-range float is undefined and the guards would branch).
__device__ __forceinline__ int32_t f2i(float v) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// int(sign(v)) (src/fshader.glsl:71); sign(NaN) = 0.
__device__ __forceinline__ int32_t isign(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }

__device__ __forceinline__ float max0(float v) { return v > 0.0f ? v : 0.0f; }

// a / d rounded to nearest, given y = RN(1/d) (one IEEE division per ray) and
// dok = (2^-60 <= |d| <= 2^60).  Markstein's correction: q0 = RN(a*y),
// r = a - q0*d (exact, by FMA), q = RN(q0 + r*y) = RN(a/d).  With a and d
// in [2^-60, 2^60] every intermediate stays in the normal range, so each step
// is the same computation on the significands scaled by a power of two; the
// library's selftest checks all 2^23 x 2^23 significand pairs against IEEE
// division.  Any other input (zero, subnormal, huge, infinite, NaN) takes the
// IEEE division itself.
__device__ __forceinline__ float div_by(float a, float d, float y, bool dok) {
    const uint32_t ua = __float_as_uint(a) & 0x7fffffffu;
    const bool ok = dok & ((ua - 0x21800000u) <= (0x5d800000u - 0x21800000u));
    const float q0 = a * y;
    float q = __builtin_fmaf(__builtin_fmaf(-q0, d, a), y, q0);
    if (__builtin_expect(!ok, 0)) q = a / d;
    return q;
}

// The three quotients (cell boundary - position) / d of one intersect update
// share one range test and one fallback branch.  With |s| < 2^31 the cell
// int(s) does not saturate and every numerator is below 2^33 in magnitude, so
// only the lower end of the fast path's range is tested.  Larger numerators
// come only from a saturated cell (|s| >= 2^31 or infinite), which lies outside
// every grid: the ray's next step then leaves the grid and the quotients are
// never read.  A NaN numerator gives NaN on both paths.
__device__ __forceinline__ void div3_by(float ax, float ay, float az, float dx, float dy, float dz, float rx,
                                        float ry, float rz, bool dok, float& qx, float& qy, float& qz) {
    const float lo = fminf(fminf(fabsf(ax), fabsf(ay)), fabsf(az));
    const bool ok = dok & (lo >= 0x1p-60f);
    const float q0x = ax * rx, q0y = ay * ry, q0z = az * rz;
    qx = __builtin_fmaf(__builtin_fmaf(-q0x, dx, ax), rx, q0x);
    qy = __builtin_fmaf(__builtin_fmaf(-q0y, dy, ay), ry, q0y);
    qz = __builtin_fmaf(__builtin_fmaf(-q0z, dz, az), rz, q0z);
    if (__builtin_expect(!ok, 0)) {
        qx = ax / dx;
        qy = ay / dy;
        qz = az / dz;
    }
}

__device__ __forceinline__ bool div_range_ok(float d) {
    const uint32_t ud = __float_as_uint(d) & 0x7fffffffu;
    return (ud - 0x21800000u) <= (0x5d800000u - 0x21800000u);
}

// a / d for a = +0 or an integer in [1, 2^25] and d in [1, 2^24] with
// y = RN(1/d): the correction of div_by with every quantity normal (or +0),
// so no range test is needed.
__device__ __forceinline__ float div_small_int(float a, float d, float y) {
    const float q0 = a * y;
    return __builtin_fmaf(__builtin_fmaf(-q0, d, a), y, q0);
}

// RN(1/d): v_rcp_f32 (within one ulp) refined by one Newton step with FMAs,
// e = RN(1 - d*y0), y = RN(y0 + y0*e).  For |d| in [2^-60, 2^60] this equals
// IEEE division 1/d for every input: the library's reciprocal selftest
// compares the two for all 2^32 bit patterns.  rcp_fast is that form alone,
// for callers that test the range themselves; rcp_rn gives any other input
// the division itself.
__device__ __forceinline__ float rcp_fast(float d) {
    const float y0 = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, y0, 1.0f), y0, y0);
}
__device__ __forceinline__ float rcp_rn(float d) {
    float y = rcp_fast(d);
    if (__builtin_expect(!div_range_ok(d), 0)) y = 1.0f / d;
    return y;
}

// dot, length, normalize in GLM 1.0.0's order (func_geometric.inl:48-54,88;
// func_exponential.inl:136-139).
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    float x = ax * bx;
    float y = ay * by;
    float z = az * bz;
    return (x + y) + z;
}

__device__ __forceinline__ uint8_t unorm8(float c) {
    float k = c > 0.0f ? (c < 1.0f ? c : 1.0f) : 0.0f;
    return (uint8_t)(int32_t)(k * 255.0f + 0.5f);
}

// Result of one castRay.
struct Hit {
    int64_t index;   // word index of the hit voxel in the padded render copy, or -1
    float pos[3];    // hitPos (valid on a hit)
    int32_t axis;    // axis of the step into the hit voxel (-1 on a miss)
    int32_t step;    // step on that axis
    uint32_t iters;  // loop iterations (stepCount increments)
};

// Clamps a cell coordinate to [-2, n + 1] (one v_med3_i32).  A walk indexes
// the padded copy only through this: a cell two or more outside the grid
// cannot enter it on the next step, and from -2 or n + 1 that step lands on a
// guard cell (kGuard = 3 of them on each side) whatever its direction, as it
// lands outside the grid from the true cell.
__device__ __forceinline__ int32_t clamp_cell(int32_t v, uint32_t n) {
    int32_t r;
    asm("v_med3_i32 %0, %1, -2, %2" : "=v"(r) : "v"(v), "s"((int32_t)n + 1));
    return r;
}

// Index arithmetic on the padded render copy, three forms: IdxB keeps the
// byte offset of the word (4 x index) in 32 bits when the copy is at most
// 2^32 bytes and 4 SX, 4 SXY < 2^23, so the gather adds it to the copy's base
// address in the load itself; Idx32 keeps the word index in 32 bits with
// signed 24-bit multiplies when the copy has at most 2^32 words and SX, SXY <
// 2^23; Idx64 otherwise.  make() takes clamped cells (clamp_cell), so every
// offset is non-negative and inside the copy, and y, z fit 24 bits.
// ref_index() turns the word index of an in-grid cell back into the
// reference's x + X*y + X*Y*z (src/fshader.glsl:36-39).
template <typename U>
__device__ __forceinline__ int64_t ref_index_of(const GridDesc& g, U L) {
    const U z = L / (U)g.SXY;
    const U r = L - z * (U)g.SXY;
    const U y = r / (U)g.SX;
    const U x = r - y * (U)g.SX;
    return (int64_t)x + (int64_t)g.X * (int64_t)y + (int64_t)g.XY * (int64_t)z;
}
struct IdxB {
    using T = uint32_t;  // byte offset
    __device__ __forceinline__ static T make(const GridDesc& g, int32_t x, int32_t y, int32_t z) {
        // a shift-add and two 24-bit multiply-adds
        uint32_t t = ((uint32_t)x << 2) + (uint32_t)g.origin * 4u;
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(t) : "v"(y), "s"((int32_t)(g.SX * 4)), "v"(t));
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(t) : "v"(z), "s"((int32_t)(g.SXY * 4)), "v"(t));
        return t;
    }
    // stride of one step in words; advance() adds it as bytes
    __device__ __forceinline__ static T delta(const GridDesc& g, int32_t sx, int32_t sy, int32_t sz, int axis) {
        return axis == 0 ? (uint32_t)sx : (axis == 1 ? (uint32_t)sy * g.SX : (uint32_t)sz * (uint32_t)g.SXY);
    }
    __device__ __forceinline__ static T advance(T i, T d) { return i + (d << 2); }
    __device__ __forceinline__ static int32_t load(const int32_t* vox, T i) {
        return *(const int32_t*)((const char*)vox + i);
    }
    __device__ __forceinline__ static int64_t word(T i) { return (int64_t)(i >> 2); }
    __device__ __forceinline__ static int64_t ref_index(const GridDesc& g, int64_t w) {
        return ref_index_of<uint32_t>(g, (uint32_t)w - (uint32_t)g.origin);
    }
};
struct Idx32 {
    using T = uint32_t;
    __device__ __forceinline__ static T make(const GridDesc& g, int32_t x, int32_t y, int32_t z) {
        uint32_t t = (uint32_t)x + (uint32_t)g.origin;
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(t) : "v"(y), "s"((int32_t)g.SX), "v"(t));
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(t) : "v"(z), "s"((int32_t)g.SXY), "v"(t));
        return t;
    }
    __device__ __forceinline__ static T delta(const GridDesc& g, int32_t sx, int32_t sy, int32_t sz, int axis) {
        return axis == 0 ? (uint32_t)sx : (axis == 1 ? (uint32_t)sy * g.SX : (uint32_t)sz * (uint32_t)g.SXY);
    }
    __device__ __forceinline__ static T advance(T i, T d) { return i + d; }
    __device__ __forceinline__ static int32_t load(const int32_t* vox, T i) { return vox[i]; }
    __device__ __forceinline__ static int64_t word(T i) { return (int64_t)i; }
    __device__ __forceinline__ static int64_t ref_index(const GridDesc& g, int64_t w) {
        return ref_index_of<uint32_t>(g, (uint32_t)w - (uint32_t)g.origin);
    }
};
struct Idx64 {
    using T = uint64_t;
    __device__ __forceinline__ static T make(const GridDesc& g, int32_t x, int32_t y, int32_t z) {
        return g.origin + (uint64_t)(int64_t)x + (uint64_t)g.SX * (uint64_t)(int64_t)y +
               g.SXY * (uint64_t)(int64_t)z;
    }
    __device__ __forceinline__ static T delta(const GridDesc& g, int32_t sx, int32_t sy, int32_t sz, int axis) {
        return axis == 0 ? (uint64_t)(int64_t)sx
                         : (axis == 1 ? (uint64_t)(int64_t)sy * g.SX : (uint64_t)(int64_t)sz * g.SXY);
    }
    __device__ __forceinline__ static T advance(T i, T d) { return i + d; }
    __device__ __forceinline__ static int32_t load(const int32_t* vox, T i) { return vox[i]; }
    __device__ __forceinline__ static int64_t word(T i) { return (int64_t)i; }
    __device__ __forceinline__ static int64_t ref_index(const GridDesc& g, int64_t w) {
        return ref_index_of<uint64_t>(g, (uint64_t)w - g.origin);
    }
};

// castRay, src/fshader.glsl:59-129.  `lim` = min(float(dist), float(RENDER_DIST)):
// the loop test `t < dist && t < RENDER_DIST` of :83 is `t < lim` for any t.
// The walk reads the padded render copy and keeps only the word offset of
// the current cell, advanced by one step's stride per iteration; the cell
// coordinates are formed at the start and after each jump (clamp_cell), never
// per step.  A step off the grid (per-axis bounds, :42-45) lands on a guard
// cell, whose NaN jump ends the walk in that iteration with the result and
// iteration count of the reference's exit at :123-125.  The axis choice is
// computed with selects so the wave does not branch.
//
// ytop (shadow rays, and primary rays whose step count nothing shows): an
// upper bound on the y of every solid voxel, or INT32_MAX.  Let dy >= 0.  Every cell of a segment has y >= int(start.y),
// and a jump sets start.y += RN(dy * c) with c = the entry intersect plus the
// jump.  A negative entry intersect either belongs to an axis whose step
// leaves the grid (the walk ends before any jump), or comes from a zero
// direction component, which then makes c infinite or NaN at every jump from
// the first one on.  So after a jump that leaves int(start.y) > ytop with c
// finite and >= 0, every later jump keeps c >= 0, start.y never decreases,
// and no cell the walk can still reach is solid: the ray is a miss.  The loop
// stops there instead of walking on to the grid's edge or its budget; the
// result is the reference's, only the iteration count is smaller.  (A jump
// with c = -inf or NaN leaves int(start.y) at INT32_MIN or 0, where the test
// fires only if the grid has no solid voxel; c = +inf makes `travelled`
// infinite, which ends the walk at the next loop test anyway.)
template <typename IX>
__device__ __forceinline__ void cast_ray(const GridDesc& g, float sx, float sy, float sz, float dx, float dy,
                                         float dz, float lim, Hit& h, int32_t ytop = INT32_MAX) {
    using T = typename IX::T;
    const int32_t cx = f2i(sx), cy = f2i(sy), cz = f2i(sz);
    const int32_t stx = isign(dx), sty = isign(dy), stz = isign(dz);
    const int32_t fx = stx > 0, fy = sty > 0, fz = stz > 0;
    // The six reciprocals (the step sizes 1/|d + 1e-6| of :74-76 and 1/d for
    // the intersect quotients) share one range test: every |value| in
    // [2^-60, 2^60] takes rcp_rn's fast form, anything else (a zero, a
    // subnormal, a huge or infinite component) IEEE division for all six and
    // for the quotients.  A NaN component is NaN on either path.
    const float bx_ = fabsf(dx + 0.000001f), by_ = fabsf(dy + 0.000001f), bz_ = fabsf(dz + 0.000001f);
    const float lo = fminf(fminf(fminf(fabsf(dx), fabsf(dy)), fabsf(dz)), fminf(fminf(bx_, by_), bz_));
    const float hi = fmaxf(fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz)), fmaxf(fmaxf(bx_, by_), bz_));
    const bool dok = (lo >= 0x1p-60f) & (hi <= 0x1p60f);
    float ddx, ddy, ddz, rx, ry, rz;
    if (__builtin_expect(dok, 1)) {
        ddx = rcp_fast(bx_);
        ddy = rcp_fast(by_);
        ddz = rcp_fast(bz_);
        rx = rcp_fast(dx);
        ry = rcp_fast(dy);
        rz = rcp_fast(dz);
    } else {
        ddx = 1.0f / bx_;
        ddy = 1.0f / by_;
        ddz = 1.0f / bz_;
        rx = 1.0f / dx;
        ry = 1.0f / dy;
        rz = 1.0f / dz;
    }
    float ix, iy, iz;
    div3_by((float)(cx + fx) - sx, (float)(cy + fy) - sy, (float)(cz + fz) - sz, dx, dy, dz, rx, ry, rz, dok, ix, iy,
            iz);
    const T sdx = IX::delta(g, stx, sty, stz, 0);
    const T sdy = IX::delta(g, stx, sty, stz, 1);
    const T sdz = IX::delta(g, stx, sty, stz, 2);
    T idx = IX::make(g, clamp_cell(cx, g.X), clamp_cell(cy, g.Y), clamp_cell(cz, g.Z));
    const int32_t* __restrict__ vox = g.vox;

    float trav = 0.0f;
    uint32_t iters = 0;
    int32_t axis = -1;
    int64_t result = -1;
    while (trav < lim) {
        iters++;
        trav += 1.0f;
        const bool bx = (ix < iy) & (ix < iz);
        const bool by = (iy < ix) & (iy < iz);
        const bool bz = !(bx | by);
        const T dsel = bx ? sdx : (by ? sdy : sdz);
        idx = IX::advance(idx, dsel);
        const int32_t v = IX::load(vox, idx);
        // currentIntersect: the intersect of the axis just stepped, before its update
        const float cur = bx ? ix : (by ? iy : iz);
        if (v >= 0) {
            h.pos[0] = dx * cur + sx;
            h.pos[1] = dy * cur + sy;
            h.pos[2] = dz * cur + sz;
            axis = bx ? 0 : (by ? 1 : 2);
            result = IX::word(idx);
            break;
        }
        if (v != -1) {
            // the jump recomputes every intersect, so the step's update is skipped
            const float toJump = -__int_as_float(v);
            trav += toJump;
            const float c2 = cur + toJump;
            sx = dx * c2 + sx;
            sy = dy * c2 + sy;
            sz = dz * c2 + sz;
            const int32_t cx = f2i(sx), cy = f2i(sy), cz = f2i(sz);
            // A jump that leaves `travelled` at or past the budget (or NaN:
            // a guard cell) ends the walk at the loop test (:83); nothing the
            // jump recomputes is read after that.  One branch for that and
            // the cutoff.
            if (!(trav < lim) | (cy > ytop)) break;
            div3_by((float)(cx + fx) - sx, (float)(cy + fy) - sy, (float)(cz + fz) - sz, dx, dy, dz, rx, ry, rz, dok,
                    ix, iy, iz);
            idx = IX::make(g, clamp_cell(cx, g.X), clamp_cell(cy, g.Y), clamp_cell(cz, g.Z));
        } else {
            // The axis just taken has step 0 (its direction component is +-0,
            // so its intersect is +-inf or NaN and adding its delta keeps it
            // so): the cell did not move, and every later iteration repeats
            // this one exactly -- same selection, same cell, same -1 word --
            // until the budget ends the walk as a miss.  Only the counters
            // advance, so they are advanced without the loads.  (The
            // reference's own light orbit keeps lightPos.z = X/2: a sun ray
            // from a hit on the plane z = X/2 has dir.z = +0 and a NaN z
            // intersect, so z is always the axis taken, and it sticks here
            // when its cell holds -1.)  The stride just added is 0 exactly
            // when that step is: a stride is the step times 1, SX or SXY,
            // nonzero in the index's width.
            if (dsel == 0) {
                while (trav < lim) {
                    iters++;
                    trav += 1.0f;
                }
                break;
            }
            ix = bx ? ix + ddx : ix;
            iy = by ? iy + ddy : iy;
            iz = bz ? iz + ddz : iz;
        }
    }
    h.index = result;
    h.axis = axis;
    h.step = axis == 0 ? stx : (axis == 1 ? sty : stz);
    h.iters = iters;
}

// Per-pixel outputs kept in registers until the stores.
struct PixelResult {
    uint8_t rgb[3];
    float depth;
    int64_t hit_word;  // word index of the hit in the padded render copy, -1 on a miss
    uint8_t face;
    uint32_t eval, occ;
    uint32_t steps;
    uint32_t rays;
    uint32_t iters;
};

// main(), src/fshader.glsl:131-190, for the fragment at vPos = (vx, vy).
// `top`: the grid's solid-top bound for shadow rays (INT32_MAX: none).
template <typename IX>
__device__ __forceinline__ void shade_pixel(const GridDesc& g, const FrameDesc& F, float vx, float vy,
                                            PixelResult& o, int32_t top) {
    const vrt_frame& f = F.f;
    float col0 = 0.6f, col1 = 0.7f, col2 = 0.8f;

    // normalize(vec3(vPos.x*aspectRatio, vPos.y, focal)) (:136)
    const float ux = vx * f.aspect_ratio, uy = vy, uz = f.focal;
    const float un = rcp_rn(sqrtf(dot3(ux, uy, uz, ux, uy, uz)));
    const float rx = ux * un, ry = uy * un, rz = uz * un;
    // vec3(rotateMatrix * vec4(d, 0)) (:137), GLM's pairwise column sum
    const float* M = f.rotate_matrix;
    const float d0 = (M[0] * rx + M[4] * ry) + (M[8] * rz + M[12] * 0.0f);
    const float d1 = (M[1] * rx + M[5] * ry) + (M[9] * rz + M[13] * 0.0f);
    const float d2 = (M[2] * rx + M[6] * ry) + (M[10] * rz + M[14] * 0.0f);

    // The primary ray may take the cutoff when its step count is not an
    // output of this draw (no depth view, no step map): a ray that stops
    // there misses, as the reference's walk would.
    Hit h;
    cast_ray<IX>(g, f.cam_pos[0], f.cam_pos[1], f.cam_pos[2], d0, d1, d2, F.render_lim, h,
                 (F.primary_cutoff != 0 && d1 >= 0.0f) ? top : INT32_MAX);
    uint32_t iters = h.iters;
    uint32_t rays = 1;
    uint32_t eval = 0, occ = 0;
    const bool hit = h.index != -1;
    const float hx = hit ? h.pos[0] : 0.0f, hy = hit ? h.pos[1] : 0.0f, hz = hit ? h.pos[2] : 0.0f;
    // firstHitNormal = -step on the axis of the last step
    const float nx = h.axis == 0 ? (float)(-h.step) : 0.0f;
    const float ny = h.axis == 1 ? (float)(-h.step) : 0.0f;
    const float nz = h.axis == 2 ? (float)(-h.step) : 0.0f;

    o.hit_word = hit ? h.index : -1;
    o.face = hit ? (uint8_t)(h.axis * 3 + (h.step + 1)) : (uint8_t)VRT_FACE_MISS;
    o.steps = h.iters;
    if (hit) {
        const float ex = hx - f.cam_pos[0], ey = hy - f.cam_pos[1], ez = hz - f.cam_pos[2];
        o.depth = sqrtf(dot3(ex, ey, ez, ex, ey, ez));
    } else {
        o.depth = __builtin_inff();
    }

    if (f.view_depth_field == 1) {
        const float s = div_small_int((float)h.iters, 100.0f, 1.0f / 100.0f);
        col0 = s;
        col1 = s;
        col2 = s;
    } else if (hit) {
        const int32_t vcol = g.vox[h.index];
        // toLight = normalize(lightPos - firstHitPos) (:147)
        const float lx0 = f.light_pos[0] - hx, ly0 = f.light_pos[1] - hy, lz0 = f.light_pos[2] - hz;
        const float ln = rcp_rn(sqrtf(dot3(lx0, ly0, lz0, lx0, ly0, lz0)));
        const float tlx = lx0 * ln, tly = ly0 * ln, tlz = lz0 * ln;
        float m0 = 0.4f, m1 = 0.4f, m2 = 0.4f;  // AMBIENT

        // sun shadow (:152-156)
        {
            Hit s;
            cast_ray<IX>(g, hx + tlx * 0.001f, hy + tly * 0.001f, hz + tlz * 0.001f, tlx, tly, tlz, F.render_lim, s,
                         tly >= 0.0f ? top : INT32_MAX);
            iters += s.iters;
            rays++;
            eval |= 1u;
            if (s.index == -1) {
                const float add = 0.8f * max0(dot3(nx, ny, nz, tlx, tly, tlz));  // DIFFUSE
                m0 += add;
                m1 += add;
                m2 += add;
            } else {
                occ |= 1u;
            }
        }

        // local lights (:158-181).  The loop's first statement clamps to
        // MAX_OVERBRIGHT and leaves once every channel is >= 1.25; the
        // channels change only when a light adds to them, so the test is made
        // after each addition (and once for the sun) instead of on every
        // iteration, and the clamp is applied where the next iteration's test
        // would apply it: always after the sun, after light i only if i < 15
        // (no iteration follows the last).  Unplaced lights do nothing but
        // that test, so the loop visits the placed ones only.
        bool over = (m0 >= 1.25f) & (m1 >= 1.25f) & (m2 >= 1.25f);
        if (over) {
            m0 = 1.25f;
            m1 = 1.25f;
            m2 = 1.25f;
        }
        for (uint32_t rem = F.light_valid; rem != 0u; rem &= rem - 1u) {
            if (over) break;
            const int i = __builtin_ctz(rem);
            const vrt_local_light& L = f.lights[i];
            // length <= radius (:171), decided without the square root; the
            // x distance alone rules most lights out for the whole wave
            const float vx0 = L.pos[0] - hx;
            if (!(fabsf(vx0) <= F.light_axis_max[i])) continue;
            const float vy0 = L.pos[1] - hy, vz0 = L.pos[2] - hz;
            const float dd = dot3(vx0, vy0, vz0, vx0, vy0, vz0);
            if (!(dd <= F.light_dd_max[i])) continue;
            const float dist = sqrtf(dd);
            const float inv = rcp_rn(dist);
            const float ux1 = vx0 * inv, uy1 = vy0 * inv, uz1 = vz0 * inv;
            float lim = (float)f2i(dist + 1.0f);
            lim = lim < F.render_lim ? lim : F.render_lim;
            Hit s;
            cast_ray<IX>(g, hx + ux1 * 0.001f, hy + uy1 * 0.001f, hz + uz1 * 0.001f, ux1, uy1, uz1, lim, s,
                         uy1 >= 0.0f ? top : INT32_MAX);
            iters += s.iters;
            rays++;
            eval |= 2u << i;
            if (s.index == -1) {
                const float lam = max0(dot3(nx, ny, nz, ux1, uy1, uz1));
                const float fall = div_by(L.radius - dist, L.radius, F.light_rinv[i], div_range_ok(L.radius));
                m0 += L.rgb[0] * lam * fall;
                m1 += L.rgb[1] * lam * fall;
                m2 += L.rgb[2] * lam * fall;
                over = (m0 >= 1.25f) & (m1 >= 1.25f) & (m2 >= 1.25f);
                if (over & (i < VRT_MAX_LOCAL_LIGHTS - 1)) {
                    m0 = 1.25f;
                    m1 = 1.25f;
                    m2 = 1.25f;
                }
            } else {
                occ |= 2u << i;
            }
        }

        col0 = div_small_int((float)((vcol & 0x00FF0000) >> 16), 255.0f, 1.0f / 255.0f) * m0;
        col1 = div_small_int((float)((vcol & 0x0000FF00) >> 8), 255.0f, 1.0f / 255.0f) * m1;
        col2 = div_small_int((float)(vcol & 0x000000FF), 255.0f, 1.0f / 255.0f) * m2;
    }
    o.rgb[0] = unorm8(col0);
    o.rgb[1] = unorm8(col1);
    o.rgb[2] = unorm8(col2);
    o.eval = eval;
    o.occ = occ;
    o.rays = rays;
    o.iters = iters;
}

// The hit's reference index x + X*y + X*Y*z (two integer divisions) is formed
// only when the target has a hit map.
template <typename IX>
__device__ __forceinline__ void store_pixel(const GridDesc& g, const vrt_target& t, int64_t k, const PixelResult& p) {
    uint8_t* c = t.rgb + 3 * k;
    c[0] = p.rgb[0];
    c[1] = p.rgb[1];
    c[2] = p.rgb[2];
    if (t.depth) t.depth[k] = p.depth;
    if (t.hit) t.hit[k] = p.hit_word != -1 ? IX::ref_index(g, p.hit_word) : -1;
    if (t.face) t.face[k] = p.face;
    if (t.shadow_eval) t.shadow_eval[k] = p.eval;
    if (t.shadow_occ) t.shadow_occ[k] = p.occ;
    if (t.steps) t.steps[k] = p.steps;
}

// vPos of pixel (px, row), row 0 at the top (GL rows count from the bottom):
// (2px+1)/W - 1 and (2py+1)/H - 1, the divisions through RN(1/W), RN(1/H).
__device__ __forceinline__ void pixel_vpos(int32_t W, int32_t H, float inv_w, float inv_h, int32_t px, int32_t row,
                                           float& vx, float& vy) {
    const int32_t py = H - 1 - row;
    vx = div_small_int((float)(2 * px + 1), (float)W, inv_w) - 1.0f;
    vy = div_small_int((float)(2 * py + 1), (float)H, inv_h) - 1.0f;
}

// Sum of a 32-bit value over a full wave (all 64 lanes active), on the scalar
// unit: prefix sums within each row of 16 lanes by DPP row shifts (zeros
// shifted in), then the row totals carried on by row broadcasts, so lane 63
// holds the total.  Six VALU adds instead of the lane-index arithmetic and
// LDS permutes of a shuffle tree.  The caller keeps the total below 2^32.
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// The same for per-lane values below 2^58: the low 26 bits and the rest are
// summed apart (64 lanes of 26 bits stay below 2^32), then recombined.
__device__ __forceinline__ unsigned long long wave_sum_u58(unsigned long long v) {
    const uint32_t lo = wave_sum_u32((uint32_t)(v & 0x3ffffffull));
    const uint32_t hi = wave_sum_u32((uint32_t)(v >> 26));
    return (unsigned long long)lo + ((unsigned long long)hi << 26);
}


}  // namespace dev
}  // namespace vrt
