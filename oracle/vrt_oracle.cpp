/*
 * vrt_oracle.cpp — CPU oracle for the voxel ray tracer.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only to check or time against it.  The product
 * (libvrt.so) never links, loads or falls back to it.
 *
 * It restates, line by line and in the same float operation order, the
 * reference's per-pixel path and its depth-field / edit code:
 *   src/fshader.glsl:33-190    getVoxelIndex, vec3ToIntVec3, castRay, main
 *   src/render.cpp:66-100      computeDepthIndices
 *   src/render.cpp:226-253     fixDepthField
 *   src/render.cpp:256-271     placeVoxel / destroyVoxel
 *   src/render.cpp:273-286     computeDepthField
 *   src/level.cpp:30-56        removeSphere
 * Build: g++ -O2 -fopenmp -ffp-contract=off -fno-fast-math -fwrapv (no FTZ/DAZ),
 * so every float operation is one IEEE binary32 operation, never fused.
 *
 * Choices where GLSL leaves the result open (the GPU kernel makes the same):
 *   - normalize(v) = v * (1/sqrt(dot(v,v))), dot = (x*x + y*y) + z*z and
 *     mat4*vec4 = (m0*v0 + m1*v1) + (m2*v2 + m3*v3): GLM 1.0.0's orders
 *     (src/glm/detail/func_geometric.inl:48-54,88,
 *      src/glm/detail/func_exponential.inl:136-139,
 *      src/glm/detail/type_mat4x4.inl:560-571);
 *   - int(float) truncates toward zero, saturates out of range, NaN -> 0;
 *   - sign(NaN) = 0; max(0, x) = x > 0 ? x : 0;
 *   - UNORM8 = int(clamp(c,0,1)*255 + 0.5) with NaN -> 0;
 *   - vPos of pixel (px, row) = ((2px+1)/W - 1, (2(H-1-row)+1)/H - 1);
 *   - depth-field neighbours outside the grid are not solid (the reference
 *     reads voxels[-1], src/render.cpp:237-240).
 * Extensions (reduce to the reference at their default values):
 *   - grid X x Y x Z (reference 512 x 96 x 512), 64-bit index;
 *   - render_dist (RENDER_DIST, 384), focal (1.0);
 *   - per-light RGB intensity (reference: scalar .a) and radius (64); the
 *     overbright early-out fires when all three channels are >= 1.25, which
 *     is the scalar rule when r = g = b;
 *   - float32 depth = length(hitPos - camPos), +inf on a miss.
 *
 * ORC_MUTANT (default 0): building with -DORC_MUTANT=k (k >= 1) plants one
 * deliberate, plausible mistake (the MUT(k) sites below).  Only
 * tests/test_oracle_mutants.py builds those variants, to show that the pins
 * in tests/test_oracle_pins.py fail on each of them.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <climits>
#include <vector>
#include <omp.h>

#include "../include/vrt.h"

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(k) (ORC_MUTANT == (k))

namespace {

/* The grid as the shader's storage buffer: word i is at(i).  The functions
 * below take any type with X, Y, Z and at() (the lazily evaluated world
 * further down is the other one). */
struct Grid {
    const int32_t* vox;
    int64_t X, Y, Z;
    int32_t at(int64_t i) const { return vox[i]; }
};

/* getVoxelIndex, src/fshader.glsl:33-52.  The shader's component tests
 * (y*W in [0, W*H), z*W*H in [0, W*H*W), x in [0, W)) are the per-axis
 * bounds; the index is formed in 64 bits (2048x512x2048 has 2^31 cells). */
template <class W>
inline int64_t voxel_index(const W& g, int32_t x, int32_t y, int32_t z) {
    if (x >= 0 && y >= 0 && z >= 0 && x < g.X && y < g.Y && z < g.Z)
        return (int64_t)x + g.X * (int64_t)y + g.X * g.Y * (int64_t)z;
    return -1;
}

/* int(float): vec3ToIntVec3, src/fshader.glsl:54-57. */
inline int32_t f2i(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

inline float glsl_sign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
inline float max0(float v) { return v > 0.0f ? v : 0.0f; }

inline float int_bits_to_float(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

inline float dot3(const float a[3], const float b[3]) {
    float x = a[0] * b[0];
    float y = a[1] * b[1];
    float z = a[2] * b[2];
    return (x + y) + z;
}

inline float length3(const float v[3]) { return std::sqrt(dot3(v, v)); }

inline void normalize3(const float v[3], float out[3]) {
    float inv = 1.0f / std::sqrt(dot3(v, v));
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
    if (MUT(6)) {  /* v / length(v) instead of v * (1 / length(v)) */
        float len = std::sqrt(dot3(v, v));
        out[0] = v[0] / len;
        out[1] = v[1] / len;
        out[2] = v[2] / len;
    }
}

inline uint8_t unorm8(float c) {
    float k = c > 0.0f ? (c < 1.0f ? c : 1.0f) : 0.0f;
    return (uint8_t)(int32_t)(k * 255.0f + (MUT(10) ? 0.0f : 0.5f));
}

/* Globals of the shader that castRay writes (src/fshader.glsl:28-31). */
struct ShaderGlobals {
    float hitPos[3] = {0.0f, 0.0f, 0.0f};
    float hitNormal[3] = {0.0f, 0.0f, 0.0f};
    float stepCount = 0.0f;
    int32_t lastAxis = -1;  /* axis of the last step, for the face code */
    int32_t lastStep = 0;
    uint64_t loopIters = 0; /* all loop iterations, for the byte count */
};

/* castRay, src/fshader.glsl:59-129, transcribed statement by statement. */
template <class W>
int64_t cast_ray(const W& g, const float startIn[3], const float rayDirection[3],
                 int32_t dist, int32_t renderDist, ShaderGlobals& G) {
    float startPosition[3] = {startIn[0], startIn[1], startIn[2]};
    int32_t currCheck[3] = {f2i(startPosition[0]), f2i(startPosition[1]), f2i(startPosition[2])};
    if (MUT(3))  /* floor instead of truncation */
        for (int a = 0; a < 3; a++) currCheck[a] = f2i(std::floor(startPosition[a]));
    int64_t fColorIndex = -1;
    int64_t tempIndex = -1;

    int32_t step[3] = {f2i(glsl_sign(rayDirection[0])), f2i(glsl_sign(rayDirection[1])),
                       f2i(glsl_sign(rayDirection[2]))};
    int32_t forwardSteps[3] = {(int32_t)(step[0] > 0), (int32_t)(step[1] > 0), (int32_t)(step[2] > 0)};
    if (MUT(5))  /* step >= 0 counted as forward */
        for (int a = 0; a < 3; a++) forwardSteps[a] = (int32_t)(step[a] >= 0);

    const float bias = MUT(2) ? -0.000001f : 0.000001f;  /* MUT(2): sign of the 1e-6 flipped */
    float dx = 1.0f / std::fabs(rayDirection[0] + bias);
    float dy = 1.0f / std::fabs(rayDirection[1] + bias);
    float dz = 1.0f / std::fabs(rayDirection[2] + bias);

    float intersect[3];
    for (int a = 0; a < 3; a++)
        intersect[a] = ((float)(currCheck[a] + forwardSteps[a]) - startPosition[a]) / rayDirection[a];

    float currDist = 0.0f;
    float distTravelled = 0.0f;
    while (distTravelled < (float)dist && distTravelled < (float)renderDist) {
        G.stepCount += 1.0f;
        distTravelled += 1.0f;
        G.loopIters++;
        const bool takeX = MUT(1) ? (intersect[0] <= intersect[1] && intersect[0] <= intersect[2])  /* ties to x */
                                  : (intersect[0] < intersect[1] && intersect[0] < intersect[2]);
        if (takeX) {
            currDist = intersect[0];
            currCheck[0] += step[0];
            intersect[0] += dx;
            G.hitNormal[0] = (float)(-step[0]); G.hitNormal[1] = 0.0f; G.hitNormal[2] = 0.0f;
            G.lastAxis = 0; G.lastStep = step[0];
        } else if (intersect[1] < intersect[0] && intersect[1] < intersect[2]) {
            currDist = intersect[1];
            currCheck[1] += step[1];
            intersect[1] += dy;
            G.hitNormal[0] = 0.0f; G.hitNormal[1] = (float)(-step[1]); G.hitNormal[2] = 0.0f;
            G.lastAxis = 1; G.lastStep = step[1];
        } else {
            currDist = intersect[2];
            currCheck[2] += step[2];
            intersect[2] += dz;
            G.hitNormal[0] = 0.0f; G.hitNormal[1] = 0.0f; G.hitNormal[2] = (float)(-step[2]);
            G.lastAxis = 2; G.lastStep = step[2];
        }
        tempIndex = voxel_index(g, currCheck[0], currCheck[1], currCheck[2]);

        const int32_t word = tempIndex >= 0 ? g.at(tempIndex) : 0;
        if (tempIndex >= 0 && word >= 0) {
            for (int a = 0; a < 3; a++) G.hitPos[a] = rayDirection[a] * currDist + startPosition[a];
            fColorIndex = tempIndex;
            break;
        } else if (tempIndex >= 0 && word != -1) {
            float toJump = -int_bits_to_float(word);
            distTravelled += toJump;
            if (!MUT(4)) currDist += toJump;  /* MUT(4): the jump does not move the start */
            if (MUT(13)) G.stepCount += 1.0f;  /* MUT(13): a jump counted as two steps */
            for (int a = 0; a < 3; a++) startPosition[a] = rayDirection[a] * currDist + startPosition[a];
            for (int a = 0; a < 3; a++) currCheck[a] = f2i(MUT(3) ? std::floor(startPosition[a]) : startPosition[a]);
            for (int a = 0; a < 3; a++)
                intersect[a] = ((float)(currCheck[a] + forwardSteps[a]) - startPosition[a]) / rayDirection[a];
        } else if (tempIndex < 0) {
            break;
        }
    }
    return fColorIndex;
}

struct PixelOut {
    uint8_t rgb[3];
    float depth;
    int64_t hit;
    uint8_t face;
    uint32_t eval, occ;
    uint32_t steps;
    uint64_t rays;
    uint64_t iters;
};

/* Primary ray of src/fshader.glsl:136-137 for vPos = (vx, vy):
 * rotateMatrix * vec4(normalize(vec3(vPos.x * aspectRatio, vPos.y, focal)), 0)
 * with the column-major matrix and GLM's pairwise sum order. */
void primary_dir(const vrt_frame& f, float vx, float vy, float out[3]) {
    float unrot[3] = {vx * f.aspect_ratio, vy, f.focal};
    if (MUT(23)) { unrot[0] = vx; unrot[1] = vy * f.aspect_ratio; }  /* aspect on the wrong axis */
    float d[3];
    normalize3(unrot, d);
    if (MUT(11)) { d[0] = unrot[0]; d[1] = unrot[1]; d[2] = unrot[2]; }  /* not normalized */
    const float* M = f.rotate_matrix;
    for (int c = 0; c < 3; c++)
        out[c] = (M[0 + c] * d[0] + M[4 + c] * d[1]) + (M[8 + c] * d[2] + M[12 + c] * 0.0f);
    if (MUT(12))  /* the matrix read row-major (transposed) */
        for (int c = 0; c < 3; c++)
            out[c] = (M[4 * c + 0] * d[0] + M[4 * c + 1] * d[1]) + (M[4 * c + 2] * d[2] + M[4 * c + 3] * 0.0f);
}

/* main(), src/fshader.glsl:131-190, for the fragment at vPos = (vx, vy). */
template <class W>
void shade_pixel(const W& g, const vrt_frame& f, float vx, float vy, PixelOut& o) {
    ShaderGlobals G;
    float fColor[3] = {0.6f, 0.7f, 0.8f};

    float rotatedDir[3];
    primary_dir(f, vx, vy, rotatedDir);

    int64_t fColorIndex = cast_ray(g, f.cam_pos, rotatedDir, f.render_dist, f.render_dist, G);
    float firstHitPos[3] = {G.hitPos[0], G.hitPos[1], G.hitPos[2]};
    float firstHitNormal[3] = {G.hitNormal[0], G.hitNormal[1], G.hitNormal[2]};
    uint32_t primarySteps = (uint32_t)G.stepCount;

    o.rays = 1;
    o.eval = 0;
    o.occ = 0;
    o.hit = fColorIndex;
    o.face = fColorIndex != -1 ? (uint8_t)(G.lastAxis * 3 + (G.lastStep + 1)) : (uint8_t)VRT_FACE_MISS;
    o.steps = primarySteps;
    if (fColorIndex != -1) {
        float d[3] = {firstHitPos[0] - f.cam_pos[0], firstHitPos[1] - f.cam_pos[1],
                      firstHitPos[2] - f.cam_pos[2]};
        o.depth = length3(d);
    } else {
        o.depth = INFINITY;
    }

    if (f.view_depth_field == 1) {
        float s = G.stepCount / (MUT(22) ? 255.0f : 100.0f);  /* MUT(22): wrong scale */
        fColor[0] = s; fColor[1] = s; fColor[2] = s;
    } else {
        float lv[3] = {f.light_pos[0] - firstHitPos[0], f.light_pos[1] - firstHitPos[1],
                       f.light_pos[2] - firstHitPos[2]};
        float toLight[3];
        normalize3(lv, toLight);

        float multiplier[3] = {0.4f, 0.4f, 0.4f};   /* AMBIENT */
        if (fColorIndex != -1 && g.at(fColorIndex) >= 0) {
            float so[3];
            for (int a = 0; a < 3; a++) so[a] = firstHitPos[a] + toLight[a] * 0.001f;
            o.rays++;
            o.eval |= 1u;
            if (cast_ray(g, so, toLight, f.render_dist, f.render_dist, G) == -1) {
                float lam = MUT(17) ? dot3(firstHitNormal, toLight) : max0(dot3(firstHitNormal, toLight));
                for (int c = 0; c < 3; c++) multiplier[c] += 0.8f * lam;   /* DIFFUSE */
            } else {
                o.occ |= 1u;
            }

            for (int i = 0; i < VRT_MAX_LOCAL_LIGHTS; i++) {
                const vrt_local_light& L = f.lights[i];
                const bool over = MUT(7) ? (multiplier[0] > 1.25f && multiplier[1] > 1.25f && multiplier[2] > 1.25f)
                                         : (multiplier[0] >= 1.25f && multiplier[1] >= 1.25f && multiplier[2] >= 1.25f);
                if (over) {
                    multiplier[0] = 1.25f; multiplier[1] = 1.25f; multiplier[2] = 1.25f;  /* MAX_OVERBRIGHT */
                    break;
                } else if (MUT(18) ? (L.pos[0] > 0.0f && L.pos[1] > 0.0f && L.pos[2] > 0.0f)
                                   : (L.pos[0] >= 0.0f && L.pos[1] >= 0.0f && L.pos[2] >= 0.0f)) {
                    float lvec[3] = {L.pos[0] - firstHitPos[0], L.pos[1] - firstHitPos[1],
                                     L.pos[2] - firstHitPos[2]};
                    float localLightDist = length3(lvec);
                    if (MUT(8) ? localLightDist < L.radius : localLightDist <= L.radius) {
                        float toLocalLight[3];
                        normalize3(lvec, toLocalLight);
                        float lo[3];
                        for (int a = 0; a < 3; a++) lo[a] = firstHitPos[a] + toLocalLight[a] * 0.001f;
                        o.rays++;
                        o.eval |= 2u << i;
                        const int32_t budget = MUT(20) ? f.render_dist : f2i(localLightDist + 1.0f);  /* MUT(20): no distance budget */
                        if (cast_ray(g, lo, toLocalLight, budget, f.render_dist, G) == -1) {
                            float lam = max0(dot3(firstHitNormal, toLocalLight));
                            float fall = (L.radius - localLightDist) / L.radius;
                            if (MUT(16)) fall = localLightDist / L.radius;  /* the "r -" term dropped */
                            for (int c = 0; c < 3; c++) multiplier[c] += L.rgb[c] * lam * fall;
                        } else {
                            o.occ |= 2u << i;
                        }
                    }
                }
            }

            if (MUT(24) && multiplier[0] >= 1.25f && multiplier[1] >= 1.25f && multiplier[2] >= 1.25f)
                multiplier[0] = multiplier[1] = multiplier[2] = 1.25f;  /* MUT(24): clamp after the last light too */
            int32_t v = g.at(fColorIndex);
            if (MUT(19)) v = ((v & 0xFF) << 16) | (v & 0xFF00) | ((v >> 16) & 0xFF);  /* red and blue swapped */
            fColor[0] = (float)((v & 0x00FF0000) >> 16) / 255.0f * multiplier[0];
            fColor[1] = (float)((v & 0x0000FF00) >> 8) / 255.0f * multiplier[1];
            fColor[2] = (float)(v & 0x000000FF) / 255.0f * multiplier[2];
        }
    }
    o.rgb[0] = unorm8(fColor[0]);
    o.rgb[1] = unorm8(fColor[1]);
    o.rgb[2] = unorm8(fColor[2]);
    o.iters = G.loopIters;
}

/* computeDepthIndices, src/render.cpp:66-100. */
struct DepthEntry {
    float dist;
    int32_t x, y, z;
};

std::vector<DepthEntry> make_depth_table() {
    const int R = VRT_DEPTH_FIELD_RADIUS;
    std::vector<DepthEntry> out;
    float dist = (float)(-R + 1);
    for (int zCheck = -R; zCheck <= R; zCheck++)
        for (int yCheck = -R; yCheck <= R; yCheck++)
            for (int xCheck = -R; xCheck <= R; xCheck++)
                if (xCheck * xCheck + yCheck * yCheck + zCheck * zCheck <= R * R) {
                    int xDist = xCheck - (xCheck > 0) + (xCheck < 0);
                    int yDist = yCheck - (yCheck > 0) + (yCheck < 0);
                    int zDist = zCheck - (zCheck > 0) + (zCheck < 0);
                    if (MUT(9)) {  /* offsets not shrunk toward 0 */
                        xDist = xCheck;
                        yDist = yCheck;
                        zDist = zCheck;
                    }
                    /* sqrt of an int is the double overload; the result is
                     * stored in the float `dist` (src/render.cpp:67,81). */
                    dist = (float)(-std::sqrt((double)(xDist * xDist + yDist * yDist + zDist * zDist)));
                    if (dist <= (float)R) out.push_back({dist, xCheck, yCheck, zCheck});
                }
    return out;
}

const std::vector<DepthEntry>& depth_table() {
    static const std::vector<DepthEntry> t = make_depth_table();
    return t;
}

/* fixDepthField, src/render.cpp:226-253; returns the new word of (x,y,z)
 * without writing, reading solidity from `g`. */
template <class W>
int32_t fix_depth_value(const W& g, int32_t x, int32_t y, int32_t z) {
    int64_t index = voxel_index(g, x, y, z);
    if (index < 0) return 0;
    int32_t old = g.at(index);
    float dist = (float)(-VRT_DEPTH_FIELD_RADIUS + 1);
    float nearestDist = dist;
    if (old < 0) {
        const auto& T = depth_table();
        for (size_t i = 0; i < T.size(); i++) {
            dist = T[i].dist;
            int64_t indexCheck = voxel_index(g, x + T[i].x, y + T[i].y, z + T[i].z);
            bool solid = indexCheck >= 0 && g.at(indexCheck) >= 0;
            if (solid && dist > nearestDist) {
                if (dist <= -2.0f || MUT(15)) nearestDist = dist;  /* MUT(15): -2 < d* stored too */
                else nearestDist = 0.0f;
            }
        }
        if (nearestDist < 0.0f) {
            int32_t bits;
            std::memcpy(&bits, &nearestDist, 4);
            return bits;
        }
    }
    return old;
}

}  // namespace

extern "C" {

int orc_depth_table(float* dist, int32_t* off, int32_t cap) {
    const auto& T = depth_table();
    int n = (int)T.size();
    for (int i = 0; i < n && i < cap; i++) {
        dist[i] = T[i].dist;
        off[3 * i] = T[i].x;
        off[3 * i + 1] = T[i].y;
        off[3 * i + 2] = T[i].z;
    }
    return n;
}

/* fixDepthField on one voxel, in place. */
void orc_fix_depth_field(int32_t* vox, int64_t X, int64_t Y, int64_t Z, int32_t x, int32_t y, int32_t z) {
    Grid g{vox, X, Y, Z};
    int64_t index = voxel_index(g, x, y, z);
    if (index >= 0) vox[index] = fix_depth_value(g, x, y, z);
}

/* New word for each listed voxel (x,y,z triples), reading `vox` unchanged. */
void orc_fix_depth_values(const int32_t* vox, int64_t X, int64_t Y, int64_t Z, const int32_t* xyz,
                          int64_t n, int32_t* out, int32_t threads) {
    Grid g{vox, X, Y, Z};
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t i = 0; i < n; i++) out[i] = fix_depth_value(g, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

/* computeDepthField over the whole grid (src/render.cpp:273-286).  Every
 * voxel's new word depends only on which neighbours are solid, and writing a
 * jump value never changes solidity, so z-slabs run in parallel. */
void orc_build_depth_field(int32_t* vox, int64_t X, int64_t Y, int64_t Z, int32_t threads) {
    Grid g{vox, X, Y, Z};
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) collapse(2)
    for (int64_t z = 0; z < Z; z++)
        for (int64_t y = 0; y < Y; y++)
            for (int64_t x = 0; x < X; x++) {
                int64_t i = x + X * y + X * Y * z;
                vox[i] = fix_depth_value(g, (int32_t)x, (int32_t)y, (int32_t)z);
            }
}

/* placeVoxel, src/render.cpp:256-262. */
void orc_place_voxel(int32_t* vox, int64_t X, int64_t Y, int64_t Z, int32_t x, int32_t y, int32_t z, int32_t v) {
    Grid g{vox, X, Y, Z};
    int64_t index = voxel_index(g, x, y, z);
    if (index >= 0) vox[index] = v;
}

/* destroyVoxel, src/render.cpp:265-271 (the voxels[-1] write is dropped). */
void orc_destroy_voxel(int32_t* vox, int64_t X, int64_t Y, int64_t Z, int32_t x, int32_t y, int32_t z) {
    Grid g{vox, X, Y, Z};
    int64_t index = voxel_index(g, x, y, z);
    if (index >= 0) vox[index] = -1;
}

/* removeSphere, src/level.cpp:30-56: sequential, in the reference's order. */
void orc_remove_sphere(int32_t* vox, int64_t X, int64_t Y, int64_t Z, int32_t px, int32_t py, int32_t pz,
                       int32_t radius) {
    for (int z = -radius; z < radius; z++)
        for (int y = -radius; y < radius; y++)
            for (int x = -radius; x < radius; x++)
                if (x + px < X && y + py < Y && z + pz < Z && x + px >= 0 && y + py >= 0 && z + pz >= 0)
                    if (MUT(14) ? x * x + y * y + z * z <= radius * radius : x * x + y * y + z * z < radius * radius)
                        orc_destroy_voxel(vox, X, Y, Z, x + px, y + py, z + pz);
    if (!MUT(21)) radius += VRT_DEPTH_FIELD_RADIUS >> 1;  /* MUT(21): repair only the carved ball */
    for (int z = -radius; z < radius; z++)
        for (int y = -radius; y < radius; y++)
            for (int x = -radius; x < radius; x++)
                if (x + px < X && y + py < Y && z + pz < Z && x + px >= 0 && y + py >= 0 && z + pz >= 0)
                    if (x * x + y * y + z * z < radius * radius)
                        orc_fix_depth_field(vox, X, Y, Z, x + px, y + py, z + pz);
}

/* castRay for a list of rays; outputs optional (NULL). */
void orc_cast_rays(const int32_t* vox, int64_t X, int64_t Y, int64_t Z, int64_t n, const float* start,
                   const float* dir, const int32_t* dist, int32_t renderDist, int64_t* hit, float* hitPos,
                   uint8_t* face, uint32_t* steps) {
    Grid g{vox, X, Y, Z};
    for (int64_t i = 0; i < n; i++) {
        ShaderGlobals G;
        int64_t r = cast_ray(g, start + 3 * i, dir + 3 * i, dist[i], renderDist, G);
        if (hit) hit[i] = r;
        if (hitPos) for (int a = 0; a < 3; a++) hitPos[3 * i + a] = r >= 0 ? G.hitPos[a] : 0.0f;
        if (face) face[i] = r >= 0 ? (uint8_t)(G.lastAxis * 3 + (G.lastStep + 1)) : (uint8_t)VRT_FACE_MISS;
        if (steps) steps[i] = (uint32_t)G.stepCount;
    }
}

typedef struct orc_outputs {
    uint8_t* rgb;
    float* depth;
    int64_t* hit;
    uint8_t* face;
    uint32_t* eval;
    uint32_t* occ;
    uint32_t* steps;
} orc_outputs;

static inline void store_pixel(const orc_outputs* o, int64_t k, const PixelOut& p) {
    if (o->rgb) { o->rgb[3 * k] = p.rgb[0]; o->rgb[3 * k + 1] = p.rgb[1]; o->rgb[3 * k + 2] = p.rgb[2]; }
    if (o->depth) o->depth[k] = p.depth;
    if (o->hit) o->hit[k] = p.hit;
    if (o->face) o->face[k] = p.face;
    if (o->eval) o->eval[k] = p.eval;
    if (o->occ) o->occ[k] = p.occ;
    if (o->steps) o->steps[k] = p.steps;
}

static inline void pixel_vpos(int32_t W, int32_t H, int64_t px, int64_t row, float& vx, float& vy) {
    int64_t py = H - 1 - row;  /* GL rows count from the bottom */
    vx = (float)(2 * px + 1) / (float)W - 1.0f;
    vy = (float)(2 * py + 1) / (float)H - 1.0f;
}

}  // extern "C"

namespace {

/* One frame.  pixels == NULL: every pixel, outputs indexed row*W + px
 * (row 0 = top), work split in 8x8 tiles with schedule(dynamic) like
 * fragment quads.  pixels != NULL: the listed pixel ids (row*W + px) only,
 * outputs indexed by list position.  totals[0] += rays, totals[1] += DDA
 * loop iterations. */
template <class Wd>
void render_frame(const Wd& g, const vrt_frame* f, int32_t W, int32_t H, const int64_t* pixels, int64_t npix,
                  const orc_outputs* o, uint64_t* totals, int32_t threads) {
    uint64_t rays = 0, iters = 0;
    if (pixels == nullptr) {
        int64_t tx = (W + 7) / 8, ty = (H + 7) / 8;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : rays, iters)
        for (int64_t t = 0; t < tx * ty; t++) {
            int64_t bx = (t % tx) * 8, by = (t / tx) * 8;
            for (int64_t r = by; r < by + 8 && r < H; r++)
                for (int64_t c = bx; c < bx + 8 && c < W; c++) {
                    float vx, vy;
                    pixel_vpos(W, H, c, r, vx, vy);
                    PixelOut p;
                    shade_pixel(g, *f, vx, vy, p);
                    store_pixel(o, r * W + c, p);
                    rays += p.rays;
                    iters += p.iters;
                }
        }
    } else {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : rays, iters)
        for (int64_t k = 0; k < npix; k++) {
            int64_t id = pixels[k];
            float vx, vy;
            pixel_vpos(W, H, id % W, id / W, vx, vy);
            PixelOut p;
            shade_pixel(g, *f, vx, vy, p);
            store_pixel(o, k, p);
            rays += p.rays;
            iters += p.iters;
        }
    }
    if (totals) {
        totals[0] += rays;
        totals[1] += iters;
    }
}

/* The world a render sees after removeSphere edits, evaluated on demand
 * from the raw world (solid colours and -1, no depth field) and the edit
 * log, for worlds too large to build the depth field of on the CPU.  It
 * stands for: computeDepthField over the raw world (src/render.cpp:273-286),
 * then removeSphere (src/level.cpp:30-56) for each of the first T edits.
 *
 * At version k (k edits applied) a cell is solid iff it is solid in the raw
 * world and no edit j <= k carved it (offset o from the centre with
 * |o|^2 < r_j^2).  fixDepthField reads only solidity and the cell's old
 * word, and the old word of an empty cell it keeps is always -1: removals
 * only lower the nearest-solid distance d*, so a cell with d* > -2 now had
 * d* > -2 at every earlier fix since it became empty (a raw empty cell, or
 * a carved one, which the carve set to -1).  So the word of an empty cell at
 * version T is fixDepthField at version k, with k the last edit <= T whose
 * repair ball (|o|^2 < (r_k + 3)^2) holds the cell (a carved cell lies in
 * its own edit's ball), or k = 0, the initial build.  The same
 * fix_depth_value as the eager build computes it, on a view that answers
 * solid colour or -1 at version k; version-0 words are memoised. */
struct EditRec {
    int32_t cx, cy, cz, r;
};

struct History {
    const int32_t* raw;
    int64_t X, Y, Z;
    const EditRec* edits;
    int32_t n;

    void xyz(int64_t i, int32_t& x, int32_t& y, int32_t& z) const {
        x = (int32_t)(i % X);
        y = (int32_t)((i / X) % Y);
        z = (int32_t)(i / (X * Y));
    }
    static int64_t d2(const EditRec& e, int32_t x, int32_t y, int32_t z) {
        const int64_t dx = (int64_t)x - e.cx, dy = (int64_t)y - e.cy, dz = (int64_t)z - e.cz;
        return dx * dx + dy * dy + dz * dz;
    }
    /* carved by one of the first k edits (removeSphere's first loop) */
    bool carved(int32_t x, int32_t y, int32_t z, int32_t k) const {
        for (int32_t j = 0; j < k; j++)
            if (edits[j].r > 0 && d2(edits[j], x, y, z) < (int64_t)edits[j].r * edits[j].r) return true;
        return false;
    }
    /* last edit <= k whose repair loop (radius r + 3) visits the cell, 1-based; 0 = none */
    int32_t last_repair(int32_t x, int32_t y, int32_t z, int32_t k) const {
        for (int32_t j = k - 1; j >= 0; j--) {
            const int64_t R = (int64_t)edits[j].r + (VRT_DEPTH_FIELD_RADIUS >> 1) - (MUT(25) ? 1 : 0);  /* MUT(25): ball of r + 2 */
            if (R > 0 && d2(edits[j], x, y, z) < R * R) return j + 1;
        }
        return 0;
    }
};

/* Words as fixDepthField sees them at version k: the solid colour, or -1. */
struct SolidAt {
    const History* h;
    int32_t k;
    int64_t X, Y, Z;
    int32_t at(int64_t i) const {
        const int32_t v = h->raw[i];
        if (v < 0) return -1;
        int32_t x, y, z;
        h->xyz(i, x, y, z);
        return h->carved(x, y, z, k) ? -1 : v;
    }
};

/* Words a render sees at version T. */
struct LazyWorld {
    const History* h;
    int32_t T;
    int32_t* memo;  /* version-0 word of each raw-empty cell once computed; 0 = not yet (never a result) */
    int64_t X, Y, Z;
    int32_t at(int64_t i) const {
        const int32_t v = h->raw[i];
        int32_t x, y, z;
        h->xyz(i, x, y, z);
        if (v >= 0 && !h->carved(x, y, z, T)) return v;
        const int32_t k = MUT(26) && v < 0 ? 0 : h->last_repair(x, y, z, T);  /* MUT(26): raw-empty cells keep the first build's word */
        if (k == 0) {
            int32_t m = __atomic_load_n(memo + i, __ATOMIC_RELAXED);
            if (m != 0) return m;
            m = fix_depth_value(SolidAt{h, 0, X, Y, Z}, x, y, z);
            __atomic_store_n(memo + i, m, __ATOMIC_RELAXED);
            return m;
        }
        return fix_depth_value(SolidAt{h, k, X, Y, Z}, x, y, z);
    }
};

}  // namespace

extern "C" {

void orc_render(const int32_t* vox, int64_t X, int64_t Y, int64_t Z, const vrt_frame* f, int32_t W, int32_t H,
                const int64_t* pixels, int64_t npix, const orc_outputs* o, uint64_t* totals, int32_t threads) {
    Grid g{vox, X, Y, Z};
    render_frame(g, f, W, H, pixels, npix, o, totals, threads);
}

/* orc_render on the world after the first T edits of `edits` ((cx, cy, cz,
 * r) each) applied to `raw` with its depth field built first, evaluated on
 * demand.  raw's empty cells must be -1 (no depth field).  memo: X*Y*Z
 * int32 zeros, kept across calls on the same raw world. */
void orc_render_lazy(const int32_t* raw, int64_t X, int64_t Y, int64_t Z, const int32_t* edits, int32_t T,
                     int32_t* memo, const vrt_frame* f, int32_t W, int32_t H, const int64_t* pixels, int64_t npix,
                     const orc_outputs* o, uint64_t* totals, int32_t threads) {
    History h{raw, X, Y, Z, reinterpret_cast<const EditRec*>(edits), T};
    LazyWorld g{&h, T, memo, X, Y, Z};
    render_frame(g, f, W, H, pixels, npix, o, totals, threads);
}

/* The words of the listed cells (linear indices) at version T, as above. */
void orc_lazy_words(const int32_t* raw, int64_t X, int64_t Y, int64_t Z, const int32_t* edits, int32_t T,
                    int32_t* memo, const int64_t* idx, int64_t n, int32_t* out, int32_t threads) {
    History h{raw, X, Y, Z, reinterpret_cast<const EditRec*>(edits), T};
    LazyWorld g{&h, T, memo, X, Y, Z};
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t k = 0; k < n; k++) out[k] = g.at(idx[k]);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

/* The oracle's vector maths, exported so tests can check it against GLM. */
extern "C" {
void orc_normalize(const float* v, float* out) { normalize3(v, out); }
float orc_length(const float* v) { return length3(v); }
float orc_dot(const float* a, const float* b) { return dot3(a, b); }
/* Primary ray of src/fshader.glsl:136-137 for vPos = (vx, vy). */
void orc_primary_dir(const vrt_frame* f, float vx, float vy, float* out) { primary_dir(*f, vx, vy, out); }
}

/* initVoxels, placeTrunk, placeBush: src/level.cpp:4-27,59-138, transcribed
 * with VOXELS_WIDTH -> X (x loops) / Z (z loops) and VOXELS_HEIGHT -> Y. */
namespace {
struct RefCtx {
    int32_t* v;
    int64_t X, Y, Z;
};

void ref_place_bush(RefCtx& c, int px, int py, int pz, int cr, int cg, int cb, int radius) {
    for (int z = -radius; z < radius; z++) {
        for (int y = -radius; y < radius; y++) {
            for (int x = -radius; x < radius; x++) {
                if (x + px < c.X && y + py < c.Y && z + pz < c.Z && x + px >= 0 && y + py >= 0 && z + pz >= 0) {
                    if (x * x + y * y + z * z < radius * radius) {
                        orc_destroy_voxel(c.v, c.X, c.Y, c.Z, x, y, z);
                        int voxel = 0;
                        voxel += cr;
                        voxel = voxel << 8;
                        voxel += cg - ((x + y + z) % 3) * 20;
                        voxel = voxel << 8;
                        voxel += cb;
                        orc_place_voxel(c.v, c.X, c.Y, c.Z, x + px, y + py, z + pz, voxel);
                    }
                }
            }
        }
    }
}

void ref_place_trunk(RefCtx& c, int px, int py, int pz, int cr, int cg, int cb, int length) {
    for (int z = -length >> 2; z < length >> 2; z++) {
        for (int y = 0; y < length; y++) {
            for (int x = -length >> 2; x < length >> 2; x++) {
                if (x + px < c.X && y + py < c.Y && z + pz < c.Z) {
                    orc_destroy_voxel(c.v, c.X, c.Y, c.Z, x, y, z);
                    int voxel = 0;
                    voxel += cr - ((x + z) % 2) * 10;
                    voxel = voxel << 8;
                    voxel += cg - ((x + z) % 2) * 10;
                    voxel = voxel << 8;
                    voxel += cb;
                    orc_place_voxel(c.v, c.X, c.Y, c.Z, x + px, y + py, z + pz, voxel);
                }
            }
        }
    }
}
}  // namespace

extern "C" void orc_world_reference(int32_t* vox, int64_t X, int64_t Y, int64_t Z) {
    RefCtx c{vox, X, Y, Z};
    for (int64_t i = 0; i < X * Y * Z; i++) vox[i] = -1;  /* initRender, src/render.cpp:349-351 */
    int voxel;
    for (int z = 0; z < Z; z++) {
        for (int y = 0; y < Y; y++) {
            for (int x = 0; x < X; x++) {
                int colorVariation = 5 * ((x + y + z) % 3);
                orc_destroy_voxel(vox, X, Y, Z, x, y, z);
                if (y <= 25) {
                    voxel = 0;
                    voxel += 90 + colorVariation;
                    voxel = voxel << 8;
                    voxel += 90 + colorVariation;
                    voxel = voxel << 8;
                    voxel += 90 + colorVariation;
                    orc_place_voxel(vox, X, Y, Z, x, y, z, voxel);
                } else if (y <= 33) {
                    voxel = 0;
                    voxel += 120 + colorVariation;
                    voxel = voxel << 8;
                    voxel += 100 + colorVariation;
                    voxel = voxel << 8;
                    voxel += 0;
                    orc_place_voxel(vox, X, Y, Z, x, y, z, voxel);
                } else if (y <= 36) {
                    voxel = 0;
                    voxel += 10;
                    voxel = voxel << 8;
                    voxel += 130 + colorVariation;
                    voxel = voxel << 8;
                    voxel += 10;
                    orc_place_voxel(vox, X, Y, Z, x, y, z, voxel);
                }
            }
        }
    }
    for (int z = 10; z < Z - 10; z++) {
        for (int x = 10; x < X - 10; x++) {
            if (x % 30 == 0 && z % 25 == 0) {
                ref_place_trunk(c, x + 1 + z % 7, 36, z, 128, 100, 15, 6);
                ref_place_bush(c, x + z % 7, 36 + 10, z, 15, 128, 15, 6);
            }
        }
    }
}
