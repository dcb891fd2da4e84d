/*
 * vrt.h — C-ABI of the MI355X voxel ray tracer (libvrt.so).
 *
 * This library replaces the fragment-shader render path of Berry2460/voxel-rt:
 * the per-pixel work of src/fshader.glsl (primary-ray DDA through the packed
 * 4-byte voxel grid, hit shading, shadow rays to the global light and up to 16
 * local lights, depth-field view) together with the host entry points of
 * src/render.hpp:30-41 that feed it (upload, partial upload, uniforms, draw,
 * voxel edits, depth-field repair).  Each declaration names the reference
 * interface it stands in for.
 *
 * Conventions
 *   - Every function returns int status: VRT_OK (0) or a negative VRT_ERR_*;
 *     vrt_error_string() describes the last error on the calling thread.
 *     Nothing here calls exit() (the reference exits on shader errors,
 *     src/render.cpp:140-183).
 *   - Voxel word (src/level.cpp:13-19,95-124; src/render.cpp:269):
 *       >= 0   solid, colour (r<<16)|(g<<8)|b in bits 0-23, bits 24-30 free;
 *       == -1  empty;
 *       other negative values: empty cell holding a depth-field jump, the
 *       bits of a negative float (src/render.cpp:249-251).
 *   - Grid is X (width) x Y (height) x Z (depth), x fastest:
 *       index = x + X*y + X*Y*z   (src/render.cpp:193), computed in 64 bits.
 *     The reference is X = Z = 512 (VOXELS_WIDTH), Y = 96 (VOXELS_HEIGHT).
 *   - Pointers marked "device" are HIP device allocations; "host" are
 *     ordinary host memory.  The context owns the voxel grid on the device;
 *     the caller owns frame buffers.  All GPU work of one context is ordered
 *     on that context's stream (vrt_set_stream); calls on one context must
 *     come from one host thread at a time.
 */
#ifndef VRT_H
#define VRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRT_OK 0
#define VRT_ERR_INVALID   (-1)  /* bad argument, shape or state */
#define VRT_ERR_NOMEM     (-2)  /* device or host allocation failed */
#define VRT_ERR_HIP       (-3)  /* a HIP runtime call failed */
#define VRT_ERR_BUDGET    (-4)  /* request larger than the server's framebuffer budget */
#define VRT_ERR_NOTFOUND  (-5)  /* unknown request id */
#define VRT_ERR_NODEVICE  (-6)  /* no HIP device visible */

#define VRT_MAX_LOCAL_LIGHTS 16      /* src/render.hpp:11, src/fshader.glsl:6 */
#define VRT_DEPTH_FIELD_RADIUS 7     /* src/render.hpp:7 */
#define VRT_DEPTH_TABLE_SIZE 1419    /* entries of computeDepthIndices, src/render.cpp:66-100 */
#define VRT_FACE_MISS 255u

/* One local light slot.  Reference: vec4 localLights[16] = (xyz, a)
 * (src/render.hpp:15, src/fshader.glsl:26).  The slot is unused when any
 * position component is negative (src/fshader.glsl:167; initLocalLights sets
 * (-1,-1,-1,0), src/render.cpp:304-311).  rgb is the per-channel intensity
 * (the reference's scalar .a is rgb = (a,a,a)); radius is the range and
 * falloff distance (the reference's LOCAL_LIGHT_DIST = 64, fshader.glsl:7). */
typedef struct vrt_local_light {
    float pos[3];
    float rgb[3];
    float radius;
    float reserved;
} vrt_local_light;

/* Per-frame constants; stands in for updateUniforms (src/render.cpp:289-302)
 * and the shader uniforms (src/fshader.glsl:20-26). */
typedef struct vrt_frame {
    float cam_pos[3];          /* camPos */
    float light_pos[3];        /* lightPos: the global light is a point */
    float rotate_matrix[16];   /* rotateMatrix, column-major as glUniformMatrix4fv(.., GL_FALSE, ..) */
    float aspect_ratio;        /* aspectRatio = width/height (src/render.cpp:404-411) */
    float focal;               /* z of the unnormalised camera ray; 1.0 in the reference (fshader.glsl:136) */
    int32_t view_depth_field;  /* viewDepthField: 1 = step-count grey view (fshader.glsl:143-144) */
    int32_t render_dist;       /* RENDER_DIST step cap; 384 in the reference (fshader.glsl:5) */
    vrt_local_light lights[VRT_MAX_LOCAL_LIGHTS];
} vrt_frame;

/* Where a draw writes.  All buffers are device memory, row 0 is the TOP row
 * of the image (the reference draws with GL's bottom-left origin; pixel
 * (px, row) has vPos = ((2px+1)/W-1, (2(H-1-row)+1)/H-1)).  Only rgb is
 * required; a NULL optional buffer is not written. */
typedef struct vrt_target {
    int32_t width;
    int32_t height;
    uint8_t* rgb;           /* W*H*3, UNORM8 of the shader's fColor.rgb */
    float* depth;           /* W*H, distance camera->primary hit, +inf on miss */
    int64_t* hit;           /* W*H, voxel index of the primary hit, -1 on miss */
    uint8_t* face;          /* W*H, axis*3 + (step+1) of the step that entered the hit cell; 255 on miss */
    uint32_t* shadow_eval;  /* W*H, bit 0 = sun ray cast, bit 1+i = local light i ray cast */
    uint32_t* shadow_occ;   /* W*H, same bits, set when that shadow ray hit a voxel */
    uint32_t* steps;        /* W*H, DDA loop iterations of the primary ray (the depth view's stepCount) */
} vrt_target;

/* Counters accumulated by every draw of a context. */
typedef struct vrt_stats {
    uint64_t rays;       /* primary + shadow rays cast */
    uint64_t dda_steps;  /* DDA loop iterations executed over all rays (one 4-byte voxel read
                          * each); equal to the reference's count with the shadow cutoff off */
    uint64_t pixels;
    uint64_t draws;
} vrt_stats;

typedef struct vrt_ctx vrt_ctx;

const char* vrt_error_string(void);
const char* vrt_version(void);
int vrt_device_count(int32_t* count);

/* ---- context / grid (src/render.cpp:313-372 initRender) ---------------- */

/* Allocates the X*Y*Z grid on `device` and fills it with -1 (empty), as
 * initRender does before initVoxels (src/render.cpp:349-351). */
int vrt_create(vrt_ctx** out, int64_t X, int64_t Y, int64_t Z, int32_t device);
int vrt_destroy(vrt_ctx* ctx);
/* hipStream_t to order all work of this context on (NULL = the context's own
 * stream).  Work queued before the switch stays ahead of work queued after it
 * (the new stream waits for an event recorded on the old one). */
int vrt_set_stream(vrt_ctx* ctx, void* stream);
void* vrt_get_stream(vrt_ctx* ctx);
int vrt_sync(vrt_ctx* ctx);
/* Device pointer of the grid (int32_t[X*Y*Z]).  After writing through it,
 * call vrt_grid_modified before the next draw. */
int vrt_grid_device_ptr(vrt_ctx* ctx, int32_t** out);
/* Re-derives what the context keeps about the grid's contents after writes
 * made through vrt_grid_device_ptr: the highest y of any solid voxel (used by
 * the shadow cutoff and the plain draw's tile order) and the render kernels'
 * padded copy of the grid.  The library's own uploads and edits keep both
 * current.  Returns once the context's stream is idle.  No counterpart in the
 * reference, whose grid has no device-side writers. */
int vrt_grid_modified(vrt_ctx* ctx);
/* Cutoff (on by default): a shadow ray that is above every solid voxel of the
 * grid and not heading down stops there as a miss instead of walking on to the
 * grid's edge or its budget; so does a primary ray when nothing shows its step
 * count (the frame is not the depth view and the target has no step map), and
 * a ray of vrt_cast_rays called without a steps array.
 * Every output equals the reference's either way; only vrt_stats.dda_steps
 * (iterations executed) differs: with the cutoff off it equals the
 * reference's count.  Applies to later draws and to batches created later.
 * No counterpart in the reference. */
int vrt_set_shadow_cutoff(vrt_ctx* ctx, int32_t enable);
int vrt_grid_dims(vrt_ctx* ctx, int64_t dims[3]);

/* ---- uploads: updateGeometry / updatePartialGeometry -------------------- */

/* Full grid host->device; updateGeometry, src/render.cpp:199-202 (and the
 * initial glBufferData, src/render.cpp:366-369).  This and the other calls
 * that can add solid voxels in bulk (vrt_upload_device, vrt_upload_region,
 * vrt_grid_modified, vrt_world_generate_device) return once the context's
 * stream is idle: they read back the grid's highest solid y. */
int vrt_upload(vrt_ctx* ctx, const int32_t* host_voxels);
/* Full grid device->device (grid built or kept elsewhere on the GPU). */
int vrt_upload_device(vrt_ctx* ctx, const int32_t* device_voxels);
/* Row-wise partial upload of the box [x0,x1] x [y0,y1) x [z0,z1) of a full
 * host grid image; updatePartialGeometry, src/render.cpp:204-223, for
 * integral corners (which uploads x-rows of length x1-x0+1 for y in [y0,y1),
 * z in [z0,z1), after swapping the corners when getVoxelIndex of start is
 * above end's).  As there, nothing is uploaded when x0 lies outside the grid
 * (every row's getVoxelIndex is -1) and rows outside the grid are skipped;
 * unlike there, a row is clipped at the grid's last x instead of running on
 * into the next row, so device-side edits there are never overwritten. */
int vrt_upload_region(vrt_ctx* ctx, const int32_t* host_voxels,
                      int32_t x0, int32_t y0, int32_t z0,
                      int32_t x1, int32_t y1, int32_t z1);
/* updatePartialGeometry as the reference writes it, with float corners
 * (src/render.cpp:204-223): the rows its float loops visit (for float i from
 * start.z while i < end.z, i++; likewise j over y), each a run of
 * int(end.x - start.x) + 1 words from getVoxelIndex(int(start.x), int(j),
 * int(i)); runs that start outside the grid are skipped, a run that passes
 * the row's end continues into the next row as glBufferSubData does, and a
 * run past the grid's end is not copied.  VRT_ERR_INVALID for corners that
 * are not finite or reach 2^24 in magnitude (the reference's loops would not
 * end). */
int vrt_upload_region_f(vrt_ctx* ctx, const int32_t* host_voxels, const float start[3], const float end[3]);
/* Full grid device->host (the reference's host array is always current;
 * this returns the device's copy after device-side edits). */
int vrt_download(vrt_ctx* ctx, int32_t* host_voxels);

/* ---- depth field and edits (src/render.cpp:226-286, src/level.cpp:30-56) */

/* Host copy of the 1419-entry offset table of computeDepthIndices
 * (src/render.cpp:66-100): dist[i] and off[3*i..3*i+2], in reference order. */
int vrt_depth_table(float* dist, int32_t* off, int32_t cap, int32_t* count);
/* fixDepthField over every voxel of the grid on the device; stands in for the
 * four computeDepthField threads (src/render.cpp:273-286,354-363). */
int vrt_build_depth_field(vrt_ctx* ctx);
/* fixDepthField (src/render.cpp:226-253) for every voxel of the box
 * [x0,x1) x [y0,y1) x [z0,z1) (clipped). */
int vrt_fix_depth_field_box(vrt_ctx* ctx, int32_t x0, int32_t y0, int32_t z0,
                            int32_t x1, int32_t y1, int32_t z1);
/* removeSphere (src/level.cpp:30-56) as an on-device delta: destroyVoxel over
 * [-r,r)^3 with x*x+y*y+z*z < r*r, then fixDepthField over [-(r+3),r+3)^3
 * with x*x+y*y+z*z < (r+3)^2, both clipped to the grid. */
int vrt_remove_sphere(vrt_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t radius);
/* placeVoxel (src/render.cpp:256-262): out-of-grid writes are ignored. */
int vrt_place_voxel(vrt_ctx* ctx, int32_t x, int32_t y, int32_t z, int32_t voxel);
/* destroyVoxel (src/render.cpp:265-271).  The reference writes voxels[-1] for
 * x >= X, y >= Y or z >= Z; here any out-of-grid coordinate is ignored. */
int vrt_destroy_voxel(vrt_ctx* ctx, int32_t x, int32_t y, int32_t z);

/* ---- frame constants: updateUniforms and its producers ------------------ */

/* Reference start-up values: camPos (195,55,155), identity rotateMatrix,
 * lightPos = startLightPos = (X/2, 3X, X/2) (src/main.cpp:26-39), aspect
 * w/h, focal 1, RENDER_DIST 384, all lights unused with radius 64. */
int vrt_frame_defaults(vrt_frame* f, int64_t grid_width, int32_t width, int32_t height);
/* initLocalLights, src/render.cpp:304-311. */
int vrt_init_local_lights(vrt_frame* f);
/* placeLocalLight, src/render.cpp:375-385: first slot with a negative
 * position component takes (x,y,z), intensity (d,d,d), radius 64.
 * Returns the slot or -1 when all 16 are used. */
int vrt_place_local_light(vrt_frame* f, float x, float y, float z, float diffuse);
/* glm::rotate (src/glm/ext/matrix_transform.inl:18-46) of column-major m. */
int vrt_mat4_rotate(const float m[16], float angle_rad, const float axis[3], float out[16]);
/* glm mat4*mat4 (src/glm/detail/type_mat4x4.inl:630-648). */
int vrt_mat4_mul(const float a[16], const float b[16], float out[16]);
/* doMouseLook's rotateMatrix = rotY(yaw) * rotX(pitch) (src/controls.cpp:136-142). */
int vrt_camera_matrix(float pitch_rad, float yaw_rad, float out[16]);
/* lightUpdate's lightPos = rotZ(radians(deg)) * vec4(start, 1) (src/render.cpp:388-402). */
int vrt_light_orbit(float rotation_deg, const float start[3], float out[3]);

/* Copies the frame constants used by the next vrt_draw; updateUniforms. */
int vrt_set_frame(vrt_ctx* ctx, const vrt_frame* f);

/* ---- draw: glDrawArrays over the fullscreen quad (src/main.cpp:59) ------ */

/* One frame: fshader.glsl main() for every pixel of the target, with the
 * frame set by vrt_set_frame.  Asynchronous on the context's stream. */
int vrt_draw(vrt_ctx* ctx, const vrt_target* target);
/* Same, with explicit frame constants (no vrt_set_frame needed). */
int vrt_draw_frame(vrt_ctx* ctx, const vrt_frame* f, const vrt_target* target);
int vrt_read_stats(vrt_ctx* ctx, vrt_stats* out, int32_t reset);

/* castRay (src/fshader.glsl:59-129) for n independent rays: start[3*i],
 * dir[3*i] (device float), budget dist[i] (device int32) and step cap
 * render_dist.  Outputs (device, optional): hit index or -1, hit position,
 * face code of the last step, loop iterations.  Without a steps array the
 * rays take the cutoff (vrt_set_shadow_cutoff); the other outputs are the
 * same either way. */
int vrt_cast_rays(vrt_ctx* ctx, int64_t n, const float* start, const float* dir,
                  const int32_t* dist, int32_t render_dist,
                  int64_t* hit, float* hit_pos, uint8_t* face, uint32_t* steps);

/* ---- batched form: many views in one launch ------------------------------ */

typedef struct vrt_view {
    vrt_frame frame;
    vrt_target target;
} vrt_view;

typedef struct vrt_batch_config {
    int32_t tile_w;        /* screen tile per work item: 8x8, 16x4, 32x2 or 16x16 */
    int32_t tile_h;
    int32_t chunk_tiles;   /* 0: the hardware schedules workgroups of four tiles (no
                            * queue); >= 1: persistent waves pop this many tiles per
                            * device-queue pop */
    int32_t waves_per_cu;  /* persistent waves per CU (0 = from occupancy) */
} vrt_batch_config;

typedef struct vrt_tune_entry {
    vrt_batch_config config;
    double ms;             /* median over the timed repetitions */
} vrt_tune_entry;

typedef struct vrt_batch vrt_batch;

/* A plan over a fixed set of views: descriptors are uploaded once, the
 * launch (with the queue reset when chunk_tiles >= 1) is captured in a
 * hipGraph.  cfg NULL = the tuned config for this context if
 * vrt_autotune_batch ran, else 8x8 tiles, hardware-scheduled. */
int vrt_batch_create(vrt_ctx* ctx, const vrt_view* views, int32_t n,
                     const vrt_batch_config* cfg, vrt_batch** out);
int vrt_batch_launch(vrt_batch* b);
int vrt_batch_destroy(vrt_batch* b);
/* Renders each view with its own plain launch, one after another. */
int vrt_draw_views_plain(vrt_ctx* ctx, const vrt_view* views, int32_t n);
/* The views drawn like vrt_batch_create + vrt_batch_launch with the same
 * config (cfg NULL = the context's tuned default), outputs bit-identical, but
 * with no device allocation: the view descriptors and queue counters live in
 * a fixed array of about 230 KiB per GPU inside the library, filled by
 * kernels on the context's stream, so the call records only kernel nodes
 * under stream capture.  At most 256 views per call (VRT_ERR_INVALID above).
 * Calls that overlap on one GPU (other streams, other contexts) share that
 * array and race; calls on one stream are ordered.  No counterpart in the
 * reference: it is for callers that cannot allocate (HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). */
#define VRT_NOALLOC_MAX_VIEWS 256
int vrt_draw_views_noalloc(vrt_ctx* ctx, const vrt_view* views, int32_t n, const vrt_batch_config* cfg);
/* Times every tile shape in {8x8,16x4,32x2,16x16} x chunk size in
 * {0,1,2,4,8,16} on these views and keeps the fastest as the context's
 * default.  results (may be NULL) receives up to max_results entries. */
int vrt_autotune_batch(vrt_ctx* ctx, const vrt_view* views, int32_t n, int32_t reps,
                       vrt_batch_config* best, vrt_tune_entry* results,
                       int32_t max_results, int32_t* n_results);

/* ---- served form: resident world, edits between frames, continuous batching */

typedef struct vrt_server_config {
    uint64_t framebuffer_budget;   /* bytes of frame buffers admitted at once; the server reserves
                                    * this much device memory at creation and sub-allocates every
                                    * request's buffers from it (no device allocation or free while
                                    * serving) */
    int32_t max_batch_views;       /* views per batch launch */
    int32_t streams;               /* batches that may run on the device together */
    int32_t want_depth;            /* allocate and fill depth per request */
    int32_t want_aux;              /* allocate and fill hit/face/shadow/steps maps */
    vrt_batch_config batch;        /* tile_w == 0: the context's default */
    int64_t max_batch_pixels;      /* pixels per batch launch (0 = no limit); a batch always takes
                                    * at least one request.  Every request of a batch completes when
                                    * the whole batch does, so smaller batches on more slots cut the
                                    * latency of small requests queued behind large ones */
} vrt_server_config;

typedef struct vrt_result {
    uint64_t id;
    uint64_t world_version;        /* edits applied to the world this frame saw */
    int32_t width, height;
    /* Seconds on the server's clock (vrt_server_clock).  t_submit: as given to
     * vrt_server_submit; t_start: admitted and launched by the host; t_exec:
     * its batch began on the device (after any edit it had to see and earlier
     * work on its stream); t_gpu_done: its batch ended on the device; t_done:
     * the host saw it complete.  The two device times come from HIP events,
     * placed on the host clock through one event synchronised at creation. */
    double t_submit, t_start, t_exec, t_gpu_done, t_done;
    vrt_target target;             /* device buffers, valid until vrt_server_release */
} vrt_result;

typedef struct vrt_server vrt_server;

/* The server takes over the context until vrt_server_destroy: its edits and
 * batches run on streams of its own, ordered after whatever the context had
 * queued at creation but not with later work on the context's stream, so
 * draws, edits or uploads made on the context meanwhile race the server's.
 * vrt_server_destroy waits for the server's work and leaves the context's
 * grid and render copy with every served edit applied. */
int vrt_server_create(vrt_ctx* ctx, const vrt_server_config* cfg, vrt_server** out);
int vrt_server_destroy(vrt_server* s);
/* Queue a render request; it sees every edit submitted before it and none
 * after.  VRT_ERR_BUDGET if it alone exceeds the budget. */
int vrt_server_submit(vrt_server* s, const vrt_frame* f, int32_t width, int32_t height,
                      double t_submit, uint64_t* id);
/* Queue a removeSphere edit; returns the world version it creates. */
int vrt_server_submit_edit(vrt_server* s, int32_t cx, int32_t cy, int32_t cz, int32_t radius,
                           uint64_t* version);
/* One scheduler pass: retire finished batches, apply due edits, admit queued
 * requests under the budget and launch them.  wait != 0 blocks until at
 * least one in-flight batch finishes when nothing else can progress. */
int vrt_server_pump(vrt_server* s, int32_t wait);
/* Completed requests, oldest first; up to max. */
int vrt_server_poll(vrt_server* s, vrt_result* out, int32_t max, int32_t* n);
/* Drops a request that is still queued (not yet admitted); admitted
 * requests run to completion.  VRT_ERR_NOTFOUND otherwise. */
int vrt_server_cancel(vrt_server* s, uint64_t id);
/* Frees a completed request's buffers back to the budget. */
int vrt_server_release(vrt_server* s, uint64_t id);
/* Queued, in-flight and completed-unreleased request counts and bytes in use. */
int vrt_server_status(vrt_server* s, int32_t* queued, int32_t* inflight, int32_t* done,
                      uint64_t* bytes_in_use);
double vrt_server_clock(vrt_server* s);

/* ---- diagnostics (no reference counterpart) ----------------------------- */

/* The DDA divides by a ray's direction component through a reciprocal and
 * one FMA correction instead of a full IEEE division.  This checks that
 * path against IEEE division on this device for a = 1 + i/2^23 (every i)
 * and d = 1 + j/2^23, j in [d_begin, d_begin + d_count): all significand
 * pairs when run over j in [0, 2^23).  Returns the mismatch count. */
int vrt_selftest_division(vrt_ctx* ctx, uint32_t d_begin, uint32_t d_count, uint64_t* mismatches);
/* Reciprocals (step sizes, normalize, 1/distance) use a hardware reciprocal
 * refined by one FMA Newton step.  This compares that path with IEEE
 * division 1/d on this device for the float bit patterns
 * [first, first + count) (clipped to 2^32): every float when run over
 * [0, 2^32).  Returns the mismatch count (any NaN equals any NaN). */
int vrt_selftest_reciprocal(vrt_ctx* ctx, uint64_t first, uint64_t count, uint64_t* mismatches);

/* ---- synthetic worlds (benchmark inputs; no reference counterpart) ------ */

typedef struct vrt_world_params {
    int64_t size[3];         /* X, Y, Z */
    uint64_t seed;
    int32_t height_min;      /* terrain heights in [height_min, height_max] */
    int32_t height_max;
    int32_t octaves;         /* value-noise octaves */
    int32_t base_cell;       /* lattice spacing of octave 0 */
    int32_t caves;           /* spherical caves carved out */
    int32_t cave_r_min, cave_r_max;
    int32_t boxes;           /* axis-aligned solid boxes (overhangs) */
    int32_t box_min, box_max;
    int32_t jitter;          /* per-voxel colour jitter, +-jitter */
} vrt_world_params;

/* Generates the world into host memory (X*Y*Z int32) without depth field. */
int vrt_world_generate(const vrt_world_params* p, int32_t* host_out, int32_t threads);
/* Same world, generated on the device straight into the context's grid. */
int vrt_world_generate_device(vrt_ctx* ctx, const vrt_world_params* p);
/* The reference's own world: initVoxels with placeTrunk / placeBush
 * (src/level.cpp:4-27,59-138) on an X x Y x Z host grid, quirks included;
 * the reference is X = Z = 512, Y = 96. */
int vrt_world_reference(int64_t X, int64_t Y, int64_t Z, int32_t* host_out);
/* Terrain surface height (first empty y of the column before caves/boxes). */
int vrt_world_surface(const vrt_world_params* p, int64_t x, int64_t z, int32_t* height);

#ifdef __cplusplus
}
#endif

#endif /* VRT_H */
