// vrt_headless — the reference's frame loop (src/main.cpp:58-73) on libvrt,
// without a window: draw, lightUpdate, scripted mouse look and movement,
// destroy edits, placeLocalLight, the depth-view toggle, updateUniforms.
//
//   vrt_headless [--frames N] [--width W] [--height H] [--out DIR] [--record] [--entity]
//
// World: the reference's initVoxels (512 x 96 x 512) with its depth field,
// built on the device.  Input is scripted (fps fixed at 60): the mouse swings
// the view right then down, W is held, a sphere is destroyed every 8th frame
// 15 voxels ahead (src/controls.cpp:100-110), a light is placed on frame 5
// (T, src/controls.cpp:46-49) and the depth view is toggled on frames 20-23
// (SHIFT).  With --record every frame's constants, RGB8 and hit map go to
// DIR/frame_NNNN.bin and the edits to DIR/edits.txt; the last frame is also
// written as DIR/last.ppm.  Prints frame timings and Mrays/s.
//
// --entity turns on the reference's test entity (commented out in
// src/main.cpp:48-50,67-70; src/Entity.cpp): a 10^3 box at (195, 50, 165)
// whose centre voxel moves +z by 2/fps a frame (Entity::move: destroyVoxel,
// then placeVoxel 0xAA0000 on the host array) and whose box is re-uploaded
// every frame (Entity::update: updatePartialGeometry).  The host array is
// kept equal to the device grid as the reference's is: it is refreshed from
// the device after the depth-field build and after each sphere carve.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vrt.h"

namespace {

void check(int rc, const char* what) {
    if (rc != VRT_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, vrt_error_string());
        std::exit(1);
    }
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

struct Camera {  // the reference's camera globals (src/main.cpp:26-29)
    float pos[3] = {195.0f, 55.0f, 155.0f};
    float dir[3] = {0.0f, 0.0f, 1.0f};
    float rot[2] = {0.0f, 0.0f};
    float matrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

// mat4 * vec4 in GLM's pairwise order (src/glm/detail/type_mat4x4.inl:560-571).
void mat_vec(const float* m, const float* v, float* out) {
    for (int r = 0; r < 4; r++) out[r] = (m[r] * v[0] + m[4 + r] * v[1]) + (m[8 + r] * v[2] + m[12 + r] * v[3]);
}

// doMouseLook (src/controls.cpp:112-143) with the mouse at (lookX, lookY).
void mouse_look(Camera& c, float lookX, float lookY, float fps) {
    const float PI = 3.14159f, sensitivity = 7.0f;
    c.rot[0] += lookY * sensitivity / fps;
    c.rot[1] += lookX * sensitivity / fps;
    for (int i = 0; i < 2; i++)
        if (c.rot[i] >= 2 * PI || c.rot[i] <= -2 * PI) c.rot[i] = 0.0f;
    c.rot[0] = std::fmin(PI / 2.0f, std::fmax(-PI / 2.0f, c.rot[0]));
    check(vrt_camera_matrix(c.rot[0], c.rot[1], c.matrix), "vrt_camera_matrix");
    const float z1[4] = {0, 0, 1, 1};
    float d[4];
    mat_vec(c.matrix, z1, d);
    std::memcpy(c.dir, d, sizeof(c.dir));
}

// movementUpdate's W key (src/controls.cpp:22-33), without collision.
void move_forward(Camera& c, float fps) {
    const float speed = 36.0f / fps;
    const float v[3] = {c.dir[0], 0.0f, c.dir[2]};
    const float inv = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int i = 0; i < 3; i++) c.pos[i] += v[i] * inv * speed;
    // doGravity's bounds (src/controls.cpp:81-83)
    c.pos[0] = std::fmin(std::fmax(c.pos[0], 1.0f), 512.0f - 1.0f - 1.0f);
    c.pos[2] = std::fmin(std::fmax(c.pos[2], 1.0f), 512.0f - 1.0f - 1.0f);
}

// The reference's test entity (src/Entity.cpp:3-39).
struct Entity {
    float x = 195.0f, y = 50.0f, z = 165.0f;  // setPos(195, 50, 165) (src/main.cpp:49)
    int xdim = 10, ydim = 10, zdim = 10;      // new Entity(10, 10, 10) (src/main.cpp:48)

    // move(): destroyVoxel at the centre, z += 2/fps, placeVoxel 0x00AA0000
    // at the new centre, on the host array (src/render.cpp:256-271 bounds).
    void move(int32_t* vox, int64_t X, int64_t Y, int64_t Z, float fps, int from[3], int to[3]) {
        from[0] = (int)x + xdim / 2; from[1] = (int)y + ydim / 2; from[2] = (int)z + zdim / 2;
        set(vox, X, Y, Z, from, -1);
        z += 2.0f / fps;
        to[0] = (int)x + xdim / 2; to[1] = (int)y + ydim / 2; to[2] = (int)z + zdim / 2;
        set(vox, X, Y, Z, to, 0x00AA0000);
    }
    // update(): updatePartialGeometry(pos, pos + dims), corners truncated.
    void box(int lo[3], int hi[3]) const {
        lo[0] = (int)x; lo[1] = (int)y; lo[2] = (int)z;
        hi[0] = (int)(x + (float)xdim); hi[1] = (int)(y + (float)ydim); hi[2] = (int)(z + (float)zdim);
    }
    static void set(int32_t* vox, int64_t X, int64_t Y, int64_t Z, const int c[3], int32_t v) {
        if (c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < X && c[1] < Y && c[2] < Z)
            vox[c[0] + X * c[1] + X * Y * c[2]] = v;
    }
};

void write_ppm(const std::string& path, const std::vector<uint8_t>& rgb, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::fwrite(rgb.data(), 1, rgb.size(), f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    int frames = 32, width = 800, height = 600;
    bool record = false;
    bool entity = false;
    std::string out = ".";
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--frames" && i + 1 < argc) frames = std::atoi(argv[++i]);
        else if (a == "--width" && i + 1 < argc) width = std::atoi(argv[++i]);
        else if (a == "--height" && i + 1 < argc) height = std::atoi(argv[++i]);
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--record") record = true;
        else if (a == "--entity") entity = true;
        else {
            std::fprintf(stderr, "usage: %s [--frames N] [--width W] [--height H] [--out DIR] [--record] [--entity]\n",
                         argv[0]);
            return 2;
        }
    }
    const int64_t X = 512, Y = 96, Z = 512;
    vrt_ctx* ctx = nullptr;
    check(vrt_create(&ctx, X, Y, Z, 0), "vrt_create");  // initRender
    std::vector<int32_t> voxels((size_t)(X * Y * Z));
    check(vrt_world_reference(X, Y, Z, voxels.data()), "vrt_world_reference");  // initVoxels
    check(vrt_upload(ctx, voxels.data()), "vrt_upload");
    check(vrt_build_depth_field(ctx), "vrt_build_depth_field");
    if (entity) check(vrt_download(ctx, voxels.data()), "vrt_download");
    Entity ent;

    vrt_frame frame;
    check(vrt_frame_defaults(&frame, X, width, height), "vrt_frame_defaults");
    Camera cam;
    const float fps = 60.0f;
    float lightRotation = -45.0f;
    const float startLight[3] = {(float)X / 2.0f, (float)X * 3.0f, (float)X / 2.0f};

    vrt_target t = {};
    t.width = width;
    t.height = height;
    const size_t px = (size_t)width * height;
    hip_check(hipMalloc((void**)&t.rgb, px * 3), "hipMalloc");
    hip_check(hipMalloc((void**)&t.hit, px * sizeof(int64_t)), "hipMalloc");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    std::vector<uint8_t> rgb(px * 3);
    std::vector<int64_t> hit(px);
    FILE* edits = record ? std::fopen((out + "/edits.txt").c_str(), "w") : nullptr;
    check(vrt_read_stats(ctx, nullptr, 1), "vrt_read_stats");
    double total_ms = 0.0;

    for (int k = 0; k < frames; k++) {
        // glDrawArrays with the uniforms of the previous iteration (src/main.cpp:59)
        check(vrt_set_frame(ctx, &frame), "vrt_set_frame");
        hip_check(hipEventRecord(e0, (hipStream_t)vrt_get_stream(ctx)), "hipEventRecord");
        check(vrt_draw(ctx, &t), "vrt_draw");
        hip_check(hipEventRecord(e1, (hipStream_t)vrt_get_stream(ctx)), "hipEventRecord");
        check(vrt_sync(ctx), "vrt_sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        total_ms += ms;
        if (record || k == frames - 1) {
            hip_check(hipMemcpy(rgb.data(), t.rgb, px * 3, hipMemcpyDeviceToHost), "hipMemcpy");
            hip_check(hipMemcpy(hit.data(), t.hit, px * sizeof(int64_t), hipMemcpyDeviceToHost), "hipMemcpy");
        }
        if (record) {
            char name[64];
            std::snprintf(name, sizeof(name), "/frame_%04d.bin", k);
            FILE* f = std::fopen((out + name).c_str(), "wb");
            if (f) {
                std::fwrite(&frame, sizeof(frame), 1, f);
                std::fwrite(rgb.data(), 1, rgb.size(), f);
                std::fwrite(hit.data(), sizeof(int64_t), hit.size(), f);
                std::fclose(f);
            }
        }

        // lightUpdate (src/render.cpp:388-402)
        float increment = 0.3f / fps;
        if (lightRotation >= 360.0f) lightRotation = 0.0f;
        else if (lightRotation >= 180.0f && lightRotation < 270.0f) increment *= 8.0f;
        lightRotation += increment;
        check(vrt_light_orbit(lightRotation, startLight, frame.light_pos), "vrt_light_orbit");

        // scripted input: look right for 12 frames, then down; W held
        mouse_look(cam, k < 12 ? 0.5f : 0.0f, k < 12 ? 0.0f : 0.3f, fps);
        move_forward(cam, fps);
        if (k == 5) vrt_place_local_light(&frame, cam.pos[0], cam.pos[1] - 1.5f, cam.pos[2], 0.5f);
        frame.view_depth_field = (k >= 20 && k < 24) ? 1 : 0;
        if (k % 8 == 7) {  // doDestroy (src/controls.cpp:100-110)
            const float range = 15.0f;
            int c[3];
            for (int i = 0; i < 3; i++) c[i] = (int)(cam.pos[i] + range * cam.dir[i]);
            const int r = (int)(range / 2);
            check(vrt_remove_sphere(ctx, c[0], c[1], c[2], r), "vrt_remove_sphere");
            if (edits) std::fprintf(edits, "%d %d %d %d %d\n", k, c[0], c[1], c[2], r);
            if (entity) check(vrt_download(ctx, voxels.data()), "vrt_download");
        }
        if (entity) {  // testE->move(); testE->update();
            int from[3], to[3];
            ent.move(voxels.data(), X, Y, Z, fps, from, to);
            int lo[3], hi[3];
            ent.box(lo, hi);
            check(vrt_upload_region(ctx, voxels.data(), lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]), "vrt_upload_region");
            if (edits)
                std::fprintf(edits, "e %d %d %d %d %d %d %d\n", k, from[0], from[1], from[2], to[0], to[1], to[2]);
        }
        // updateUniforms (src/render.cpp:289-302)
        std::memcpy(frame.cam_pos, cam.pos, sizeof(cam.pos));
        std::memcpy(frame.rotate_matrix, cam.matrix, sizeof(cam.matrix));
    }
    if (edits) std::fclose(edits);
    write_ppm(out + "/last.ppm", rgb, width, height);
    vrt_stats st;
    check(vrt_read_stats(ctx, &st, 0), "vrt_read_stats");
    std::printf("frames %d  %dx%d  mean %.3f ms/frame  %.1f Mrays/s  rays %llu  dda_steps %llu\n", frames, width,
                height, total_ms / frames, (double)st.rays / (total_ms * 1e-3) / 1e6, (unsigned long long)st.rays,
                (unsigned long long)st.dda_steps);
    hipFree(t.rgb);
    hipFree(t.hit);
    vrt_destroy(ctx);
    return 0;
}
