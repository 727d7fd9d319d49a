/*
 * wsrt.h - C ABI of the MI355X-native render loop (photon emit -> photon grid ->
 * spectral camera path trace -> accumulate) of wgpu-spectral-rt.
 *
 * Every entry point is a plain C function over opaque handles, POD structs,
 * pointers and sizes. Each returns an int status (WSRT_OK == 0, negative on
 * error); nothing aborts or unwinds across the boundary. A human-readable
 * message for the last failure on the calling thread is returned by
 * wsrt_last_error().
 *
 * Reference interfaces replaced (paths are in kokutoupan/wgpu-spectral-rt):
 *   Renderer::{new,resize,render}, pub frame_count ....... src/renderer/mod.rs:21-207
 *   SceneBuilder::{new,add_material,add_mesh,add_instance,
 *                  add_light_instance,build} ............... src/scene/builder.rs:47-266
 *   geometry::{create_plane,create_cube,create_sphere} .... src/scene/geometry.rs:12-232
 *   create_cornell_box ..................................... src/scene/cornell_box.rs:6-110
 *   CameraController::build_uniform / CameraUniform ........ src/scene/camera.rs:4-13,178-194
 *   photon_emit / build_grid / shader (camera) passes ...... src/renderer/{photon_emit,build_grid,compute}_pass.rs
 *
 * Ownership: the library owns all device memory. Host pointers passed in are
 * only read (or written, for outputs) during the call. Destroy renderers
 * before the scene they use, and scenes before their device.
 *
 * Threading: a renderer is used by one host thread at a time and owns one HIP
 * stream. wsrt_renderer_render() is asynchronous on that stream; sync and
 * readback calls block until the stream has drained.
 */
#ifndef WSRT_H
#define WSRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WSRT_ABI_VERSION 2

/* ---- status codes ---------------------------------------------------- */
enum {
    WSRT_OK = 0,
    WSRT_E_INVALID = -1,     /* bad handle, bad argument, size out of range */
    WSRT_E_OOM = -2,         /* device or host allocation failed            */
    WSRT_E_HIP = -3,         /* HIP runtime error (message has the detail)  */
    WSRT_E_UNSUPPORTED = -4, /* feature or device not supported             */
    WSRT_E_STATE = -5,       /* call not valid in the object's current state */
    WSRT_E_ABORTED = -6      /* a render stopped before its work was done
                              * (canceled, or past its time limit)         */
};

/* ---- binary contract shared with the reference (byte-identical) ------- */

/* src/scene/geometry.rs:4-9 ; shader.wgsl:22-25. 32 bytes. */
typedef struct wsrt_vertex {
    float pos[4];    /* xyz, w = 1 */
    float normal[4]; /* xyz, w = 0 */
} wsrt_vertex;

/* src/scene/builder.rs:5-10 ; shader.wgsl:17-20. 32 bytes.
 * extra.x = type: 0 Lambert, 1 GGX, 2 dielectric (Cauchy), 3 emitter,
 *                 4 thin film; extensions: 5 thin-film-coated GGX.
 * Lambert/GGX/dielectric: color.rgb = albedo / f0 / tint.
 * GGX: extra.y = roughness alpha.
 * Dielectric: extra.z = Cauchy A, extra.w = Cauchy B [um^2].
 * Emitter: color.r = temperature [K], color.g = intensity, extra.y > 0 = no cull.
 * Thin film: extra.y = thickness [nm], extra.z = film IOR, extra.w = inner IOR.
 * Film-coated GGX (5): extra.y = alpha, extra.z = film IOR, extra.w = substrate
 *   IOR, color.r = film thickness [nm]. */
typedef struct wsrt_material {
    float color[4];
    float extra[4];
} wsrt_material;

/* src/scene/builder.rs:12-20 ; shader.wgsl:33-38. 64 bytes.
 * params = [kind, temperature K, intensity, pad]. The reference never reads
 * params[0] (only triangle lights exist there); here kind 0 = triangle area
 * light (v0,v1,v2 world space), kind 1 = point light at v0.xyz (extension). */
typedef struct wsrt_light_info {
    float v0[4];
    float v1[4];
    float v2[4];
    float params[4];
} wsrt_light_info;

/* src/scene/builder.rs:22-28. 16 bytes. */
typedef struct wsrt_mesh_info {
    uint32_t vertex_offset;
    uint32_t index_offset;
    uint32_t pad[2];
} wsrt_mesh_info;

/* src/scene/camera.rs:4-13 ; shader.wgsl:9-15. 272 bytes, column-major mat4s. */
typedef struct wsrt_camera_uniform {
    float view_inverse[16];
    float proj_inverse[16];
    float view[16];
    float proj[16];
    uint32_t frame_count;
    uint32_t pad[3];
} wsrt_camera_uniform;

/* shader.wgsl:40-45. 32 bytes. direction = -(incoming ray direction). */
typedef struct wsrt_photon {
    float position[3];
    float wavelength;
    float direction[3];
    float energy;
} wsrt_photon;

/* ---- acceleration structure (replaces the driver BLAS/TLAS,
 *      builder.rs:131-254) --------------------------------------------------
 * Compressed 8-wide BVH node, 80 bytes. Child i's box is
 *   lo = origin + qlo[axis][i] * 2^exp[axis],  hi = origin + qhi[axis][i] * 2^exp[axis]
 * (both rounded outward at build time, so the box contains the child).
 * meta[i] == 0: empty slot.  meta[i] & 0x80: internal child, node index
 * child_base + (meta[i] & 0x7F).  Otherwise a leaf child holding
 * (meta[i] >> 5) triangles (1..3) at tri_base + (meta[i] & 0x1F). */
typedef struct wsrt_bvh8_node {
    float origin[3];
    int8_t exp[3];
    uint8_t imask; /* bit i set: child i internal */
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
} wsrt_bvh8_node;

/* Leaf triangle record, 48 bytes: world-space v0, e1 = v1 - v0, e2 = v2 - v0.
 * v0[3] holds the world triangle id (uint32 bits), e1[3] the instance id
 * (uint32 bits), e2[3] the material id (uint32 bits). */
typedef struct wsrt_bvh_tri {
    float v0[4];
    float e1[4];
    float e2[4];
} wsrt_bvh_tri;

/* ---- render-loop configuration ---------------------------------------- */
/* Defaults equal the reference's compile-time constants:
 *   max_depth 12 (compute_pass.rs:26), photon bounces 5 (photon_emit.wgsl:219),
 *   photons 2^20 and hash 2^22 (renderer/mod.rs:18-19), cell 0.02
 *   (build_grid.wgsl:23), gather radius 0.02*(frame+1)^-0.15, spectral radius
 *   20 nm, chain cap 64 (shader.wgsl:274-342), Russian roulette after bounce 3
 *   (shader.wgsl:696). */
typedef struct wsrt_config {
    uint32_t struct_size;         /* = sizeof(wsrt_config)                     */
    uint32_t max_depth;           /* camera bounces                            */
    uint32_t photon_max_depth;    /* photon bounces                            */
    uint32_t photons_per_pass;    /* photon paths traced per frame (<= 2^24)   */
    uint32_t hash_size;           /* grid cells, power of two (<= 2^26)        */
    float cell_size;              /* grid cell edge                            */
    float gather_radius;          /* r0 in r = r0 * (frame+1)^exponent         */
    float gather_radius_exponent; /* -0.15 in the reference                    */
    float spectral_radius;        /* nm                                        */
    uint32_t gather_cap;          /* photons visited per gather (64)           */
    uint32_t rr_after;            /* roulette when bounce index > rr_after (3);
                                   * a guiding renderer plays none          */
    uint32_t enable_photons;      /* 1: photon pass + gather, 0: skip both     */
    uint32_t photon_norm;         /* density divisor (2^20 = MAX_PHOTON)       */
    uint32_t guiding;             /* 0 plain BSDF sampling; 1 learned guiding  */
    float guide_mix;              /* one-sample MIS probability of the guide   */
    float guide_lr;               /* Adam learning rate                        */
    uint32_t guide_train_samples; /* training samples per frame                */
    uint32_t env_map;             /* 1: scene environment map lights misses    */
    uint32_t frame_graphs;        /* 1: submit each batch as one HIP graph     */
    uint32_t frames_in_flight;    /* consecutive frames traced in one wavefront (1..256) */
    uint32_t tail_paths;          /* once fewer camera paths than this enter a
                                   * bounce (bounce 2 on), one launch finishes
                                   * them; 0: every bounce in the wavefront.
                                   * Not used with guiding. Results do not
                                   * depend on it.                             */
    uint32_t quirks;              /* WSRT_FIX_* bits; 0 renders as the
                                   * reference does                          */
    uint32_t trace_rounds;        /* traversal launches: rounds of resident
                                   * waves (1..4; 0 means 1). Results do not
                                   * depend on it.                             */
    uint32_t guide_train_frames;  /* guiding: train online during the first
                                   * this many frames (rounded up to whole
                                   * batches), then keep the network fixed;
                                   * 0: train every frame                     */
    uint32_t extend_refill;       /* k > 0: the closest-hit kernel gives new
                                   * rays to a wave's idle lanes once k of
                                   * them idle (1..64); 0: a wave's rays
                                   * start and end together. Results do not
                                   * depend on it.                             */
    uint32_t ordered_queues;      /* 1: each bounce's rays are queued in path
                                   * order (from per-path tags) instead of in
                                   * the order shading appended them. Results
                                   * do not depend on it.                      */
    uint32_t reserved[2];
} wsrt_config;

/* wsrt_config.quirks: each bit replaces one behaviour of the reference's
 * shader by the textbook estimator (default 0: the reference's behaviour).
 *   WSRT_FIX_EMISSION_MIS  shader.wgsl:601-628 computes a balance-heuristic
 *     weight for emission reached by a cosine-sampled bounce but does not
 *     apply it, so direct light is counted by both strategies. The bit
 *     applies prev_pdf / (prev_pdf + light_pdf), with light_pdf = t^2 /
 *     (A cos_light n_lights) of the hit emitter triangle's world-space area
 *     A (the reference's unapplied weight uses the object-space area); the
 *     two strategies' weights then sum to one. Not with guiding.
 *   WSRT_FIX_RR_CLAMP  shader.wgsl:696-700 divides by a survival probability
 *     max(throughput) that may exceed 1; the bit clamps it to 1.
 *   WSRT_FIX_ONB_SIGN  shader.wgsl:171-188 uses WGSL sign(0) = 0, so the
 *     basis of a normal with z == 0 is infinite and GGX paths there are
 *     absorbed; the bit takes sign(0) = +1. */
#define WSRT_FIX_EMISSION_MIS 1u
#define WSRT_FIX_RR_CLAMP 2u
#define WSRT_FIX_ONB_SIGN 4u

/* Fills *cfg with the reference's defaults. */
int wsrt_config_default(wsrt_config* cfg);

/* ---- opaque handles ---------------------------------------------------- */
typedef struct wsrt_device wsrt_device;
typedef struct wsrt_scene_builder wsrt_scene_builder;
typedef struct wsrt_scene wsrt_scene;
typedef struct wsrt_renderer wsrt_renderer;

/* Message of the last failing call on this thread ("" if none). */
const char* wsrt_last_error(void);
/* Library ABI version (WSRT_ABI_VERSION). */
int wsrt_abi_version(void);

/* ---- device (replaces WgpuContext::new, src/wgpu_ctx.rs:13-73) -------- */
int wsrt_device_create(int ordinal, wsrt_device** out);
int wsrt_device_destroy(wsrt_device* dev);
/* Number of HIP devices visible (0 when none). Never fails. */
int wsrt_device_count(void);
/* Device memory held by the library's objects (scenes, renderers, batches,
 * guides) on this device, and a cap on it (no reference counterpart: the
 * reference leaves allocation to wgpu). Past the cap an allocation fails as
 * if the device were full: the creating call returns WSRT_E_OOM and frees
 * what it had allocated. 0 removes the cap. Temporary buffers of a BVH build
 * are not counted. */
int wsrt_device_set_memory_limit(wsrt_device* dev, uint64_t bytes);
int wsrt_device_get_memory_used(const wsrt_device* dev, uint64_t* out);

/* ---- geometry generators (src/scene/geometry.rs) ------------------------
 * Two-call pattern: pass NULL buffers to get the sizes, then buffers of at
 * least that many elements. */
/* geometry.rs:12-35: unit quad on XZ, normal +Y. 4 vertices, 6 indices. */
int wsrt_geometry_plane(wsrt_vertex* verts, uint32_t* idx, uint32_t* n_verts, uint32_t* n_idx);
/* geometry.rs:38-118: unit cube, 24 vertices, 36 indices. */
int wsrt_geometry_cube(wsrt_vertex* verts, uint32_t* idx, uint32_t* n_verts, uint32_t* n_idx);
/* geometry.rs:121-232: icosphere of radius 0.5; subdiv 4 -> 2562 v, 5120 tris. */
int wsrt_geometry_sphere(uint32_t subdivisions, wsrt_vertex* verts, uint32_t* idx,
                         uint32_t* n_verts, uint32_t* n_idx);

/* ---- scene builder (src/scene/builder.rs:47-266) -----------------------
 * Matrices are column-major float[16] (glam Mat4::to_cols_array order). */
int wsrt_scene_builder_create(wsrt_scene_builder** out);
int wsrt_scene_builder_destroy(wsrt_scene_builder* b);
/* builder.rs:70-74 */
int wsrt_scene_add_material(wsrt_scene_builder* b, const wsrt_material* m, uint32_t* id);
/* builder.rs:76-91 */
int wsrt_scene_add_mesh(wsrt_scene_builder* b, const wsrt_vertex* verts, size_t n_verts,
                        const uint32_t* idx, size_t n_idx, uint32_t* id);
/* builder.rs:93-99 */
int wsrt_scene_add_instance(wsrt_scene_builder* b, uint32_t mesh, uint32_t mat,
                            const float transform[16]);
/* builder.rs:101-129: adds the instance and one world-space light per triangle. */
int wsrt_scene_add_light_instance(wsrt_scene_builder* b, uint32_t mesh, uint32_t mat,
                                  const float transform[16], const float params[4]);
/* Extension: a point light at `position` (params as in wsrt_light_info). */
int wsrt_scene_add_point_light(wsrt_scene_builder* b, const float position[3],
                               const float params[4]);
/* Extension: environment map, equirectangular w x h RGB float radiance
 * (row 0 = +Y pole). Used when wsrt_config.env_map = 1. */
int wsrt_scene_set_environment(wsrt_scene_builder* b, const float* rgb, uint32_t w, uint32_t h);
/* builder.rs:131-266: uploads the buffers and builds the BVH8 on the GPU.
 * The builder may be destroyed afterwards. */
int wsrt_scene_build(wsrt_scene_builder* b, wsrt_device* dev, wsrt_scene** out);
int wsrt_scene_destroy(wsrt_scene* s);
/* cornell_box.rs:6-110 into a fresh builder (not yet built). */
int wsrt_scene_builder_cornell_box(wsrt_scene_builder** out);

typedef struct wsrt_scene_info {
    uint32_t n_triangles;  /* world triangles over all instances            */
    uint32_t n_instances;
    uint32_t n_lights;
    uint32_t n_materials;
    uint32_t n_nodes;      /* BVH8 nodes                                    */
    uint32_t max_depth;    /* BVH8 depth (root = 1)                         */
    uint32_t stack_size;   /* traversal stack entries the depth requires     */
    uint32_t node_bytes;   /* sizeof(BVH8 node) = 80                        */
    uint32_t tri_bytes;    /* sizeof(leaf triangle record) = 48             */
    float build_ms;        /* device time of the BVH build                  */
    float bounds[6];       /* scene AABB min xyz, max xyz                    */
} wsrt_scene_info;
int wsrt_scene_get_info(const wsrt_scene* s, wsrt_scene_info* out);

/* ---- camera (src/scene/camera.rs:178-194) -------------------------------
 * glam look_at_rh(position, position + forward(yaw,pitch), +Y) and
 * perspective_rh(45 deg, aspect, 0.1, 100). */
int wsrt_camera_build_uniform(const float position[3], float yaw, float pitch, float aspect,
                              uint32_t frame_count, wsrt_camera_uniform* out);

/* ---- renderer (src/renderer/mod.rs:21-207) ------------------------------ */
int wsrt_renderer_create(wsrt_device* dev, const wsrt_scene* scene, uint32_t width,
                         uint32_t height, const wsrt_config* cfg, wsrt_renderer** out);
int wsrt_renderer_destroy(wsrt_renderer* r);
/* The Engine writes the camera buffer before each render (engine.rs:158-166). */
int wsrt_renderer_set_camera(wsrt_renderer* r, const wsrt_camera_uniform* cam);
/* One frame shaded with the camera's frame_count (mod.rs:168-207): photon
 * trace, grid build, camera paths, accumulation (reset when frame_count is 0).
 * Asynchronous; increments the renderer's frame counter afterwards. */
int wsrt_renderer_render(wsrt_renderer* r);
/* n frames; frame k is shaded with frame_count = renderer counter, as the
 * Engine does when it rebuilds the uniform every frame (engine.rs:158-166). */
int wsrt_renderer_render_frames(wsrt_renderer* r, uint32_t n);
int wsrt_renderer_get_frame_count(const wsrt_renderer* r, uint32_t* out);
/* Engine resets frame_count to 0 when the camera moves (engine.rs:154-156). */
int wsrt_renderer_set_frame_count(wsrt_renderer* r, uint32_t frame_count);
/* mod.rs:149-166 only refits the blit; the internal resolution is fixed. */
int wsrt_renderer_resize(wsrt_renderer* r, uint32_t window_width, uint32_t window_height);
int wsrt_renderer_sync(wsrt_renderer* r);
/* width*height float4: linear sRGB sum + sample count (shader.wgsl:801-808). */
int wsrt_renderer_read_accum(wsrt_renderer* r, float* out);
/* width*height RGBA8, Reinhard-luma tone mapped, no gamma, row 0 = image bottom
 * (shader.wgsl:809-822). */
int wsrt_renderer_read_rgba8(wsrt_renderer* r, uint8_t* out);
/* The screenshot image of ScreenshotSaver::process_and_save (screenshot.rs:27-73):
 * width*height RGBA8, rows top to bottom, sRGB OETF applied to the tone-mapped
 * RGBA8 channels, alpha 255. Writing the PNG file is left to the caller. */
int wsrt_renderer_read_screenshot(wsrt_renderer* r, uint8_t* out);
/* The presented frame of Renderer::render (mod.rs:193-200): the blit pass
 * (blit_pass.rs:114-180, blit.wgsl) draws the storage image into a
 * win_w x win_h window with its aspect kept (scale_x = target/window when
 * the window is wider, else scale_y = window/target) over a blue clear,
 * nearest-sampled, into an sRGB surface. With overlay_photons set, the
 * photon overlay of debug_photons_pass.rs and debug_photons.wgsl (recorded
 * only in a commented-out line of render) is drawn over it: each stored
 * photon of the last frame as a one-pixel point of (1, 0.6, 0.1), alpha 0.3,
 * blended additively in linear space. out: win_w*win_h RGBA8, rows top to
 * bottom, window size 1..16384 per axis. */
int wsrt_renderer_present(wsrt_renderer* r, uint32_t win_w, uint32_t win_h, int overlay_photons, uint8_t* out);
/* Accumulation state export/import (float4 per pixel) for resume. */
int wsrt_renderer_write_accum(wsrt_renderer* r, const float* in);

/* Counters of the frames rendered since the last reset. */
typedef struct wsrt_frame_stats {
    uint64_t frames;
    uint64_t camera_paths;   /* samples                                   */
    uint64_t extend_rays;    /* closest-hit rays of camera paths           */
    uint64_t shadow_rays;    /* any-hit rays                              */
    uint64_t photon_paths;   /* photon paths traced                        */
    uint64_t photon_rays;    /* closest-hit rays of photon paths           */
    uint64_t photons_stored; /* photons written to the grid                */
    uint64_t guide_samples;  /* vertices the guide network evaluated      */
    /* BVH8 nodes visited and triangles tested, while counting is enabled:
     * [0] photon rays, [1] camera extend rays, [2] shadow rays; trav_rays
     * counts traversals (a photon's first ray is tested against the
     * specular BVH's root, and traversed in full if it may reach it) */
    uint64_t trav_nodes[3];
    uint64_t trav_tris[3];
    uint64_t trav_rays[3];
    /* The same per launching kernel: [0] photon bounce, [1] extend, [2]
     * shadow, [3] photon emission (specular root test), [4] tail closest
     * hit, [5] tail any hit. kernel_rays is always counted, the others
     * while counting is enabled. */
    uint64_t kernel_rays[6];
    uint64_t kernel_nodes[6];
    uint64_t kernel_tris[6];
    uint64_t kernel_counted_rays[6];
    /* The wavefront extend kernel's node visits in lockstep, while counting
     * is enabled: per wave and round of rays, the most visits of any lane
     * times the lanes that traversed, summed. Its ratio to kernel_nodes[1]
     * (>= 1) is the visits a wave spends waiting on its longest ray. */
    uint64_t extend_wave_nodes;
    /* While counting is enabled: the most node visits of a single ray, per
     * launching kernel as kernel_nodes (kept for [1] extend and [2] shadow;
     * 0 elsewhere). */
    uint64_t ray_max_nodes[6];
} wsrt_frame_stats;
int wsrt_renderer_get_stats(wsrt_renderer* r, wsrt_frame_stats* out);
int wsrt_renderer_reset_stats(wsrt_renderer* r);
/* Count BVH8 node visits and triangle tests of every ray (adds one atomic
 * per wave per kernel, and one per wave and round of rays in the extend
 * kernel; off by default). */
int wsrt_renderer_set_traversal_counting(wsrt_renderer* r, int enable);

/* Per-kernel device time, measured with HIP events recorded on the
 * renderer's stream around each launch while profiling is enabled. */
typedef struct wsrt_kernel_stat {
    char name[32];
    uint64_t launches;
    double total_ms;
    uint64_t items;          /* work items (rays, photons, pixels) processed */
} wsrt_kernel_stat;
int wsrt_renderer_set_profiling(wsrt_renderer* r, int enable);
int wsrt_renderer_get_profile(wsrt_renderer* r, wsrt_kernel_stat* out, uint32_t cap, uint32_t* n);
int wsrt_renderer_reset_profile(wsrt_renderer* r);

/* ---- parity / debug taps -------------------------------------------------- */
/* Photon buffer of the last frame in photon-path order: slot i holds photon
 * path i's stored photon, or all zero when it stored none (n =
 * photons_per_pass). *n_stored gets the number of stored photons. */
int wsrt_debug_read_photons(wsrt_renderer* r, wsrt_photon* out, uint32_t cap, uint32_t* n_stored);
/* Grid of the last frame in the form of wsrt_photon_grid_build: cell_id per
 * photon slot (0xFFFFFFFF = skipped), count and exclusive offset per cell
 * (the renderer keeps offsets of non-empty cells only; an empty cell reads
 * the number of photons before it), photon slot per sorted position. */
int wsrt_debug_read_grid(wsrt_renderer* r, uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                         uint32_t* sorted_slot);
/* Per-pixel hero wavelengths and spectral radiance of the last frame
 * (float4 each, pixel = y*width + x). */
int wsrt_debug_read_paths(wsrt_renderer* r, float* wavelengths, float* radiance);
/* The same for n chosen pixels (pixel_ids[k] < width*height): hero
 * wavelengths and spectral radiance of each pixel's path in the last frame,
 * float4 each, in the order of pixel_ids. A parity tap; the reference has no
 * counterpart (its paths live in a shader). */
int wsrt_debug_trace_paths(wsrt_renderer* r, const uint32_t* pixel_ids, uint32_t n, float* wavelengths,
                           float* radiance);

/* Standalone photon grid build on a fixed photon buffer (the build_grid pass,
 * build_grid.wgsl:25-73, as a counting sort). photons: n host records.
 * Outputs (host): cell_id[n], count[hash_size], offset[hash_size],
 * sorted_slot[n] (slots of valid photons in cell order, stable by slot). */
int wsrt_photon_grid_build(wsrt_device* dev, const wsrt_photon* photons, uint32_t n,
                           uint32_t frame_count, float cell_size, uint32_t hash_size,
                           uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                           uint32_t* sorted_slot);

/* No counterpart in the reference. The grid build of wsrt_photon_grid_build
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): it allocates
 * nothing. photons, cell_id, count, offset and sorted_slot are DEVICE
 * pointers of the same meaning and sizes; the call is asynchronous on
 * `stream` (a hipStream_t; null: the device's own stream). Work buffers live
 * in a file-scope device array of 32 MiB per GPU (held until the process
 * ends), which limits n to 2^20 photons and hash_size to 2^22 cells
 * (WSRT_E_INVALID beyond). Calls on one GPU share that array: calls that
 * overlap in time on one GPU race. */
int wsrt_photon_grid_build_noalloc(wsrt_device* dev, const wsrt_photon* photons, uint32_t n,
                                   uint32_t frame_count, float cell_size, uint32_t hash_size,
                                   uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                                   uint32_t* sorted_slot, void* stream);

/* Closest-hit query for n rays: rays = n x 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax).
 * hits = n x 4 words: t (float bits, +inf on miss), triangle id (0xFFFFFFFF on
 * miss), u (float), v (float) ; flags[n]: bit0 hit, bit1 front face.
 * The triangle id indexes the scene's world triangles (instance order, then
 * mesh order). tmin must be >= 0: the traversal's padded box test assumes
 * distances ahead of the origin, and hits behind it (t < 0) can be missed. */
int wsrt_debug_closest_hit(const wsrt_scene* s, const float* rays, uint32_t n, uint32_t* hits,
                           uint32_t* flags);
/* Any-hit query (terminate on first hit): occluded[n] = 0/1. */
int wsrt_debug_any_hit(const wsrt_scene* s, const float* rays, uint32_t n, uint32_t* occluded);
/* Per-ray count of BVH8 nodes and triangles the closest-hit traversal visits. */
int wsrt_debug_traversal_counts(const wsrt_scene* s, const float* rays, uint32_t n,
                                uint32_t* nodes, uint32_t* tris);
/* 30-bit Morton code per world triangle (triangle order), and the build
 * order: triangles sorted by (size class, code), where a triangle is large
 * when its box spans more than 1/8 of the centroid bounds on some axis and
 * large triangles come after all others. */
int wsrt_debug_read_morton(const wsrt_scene* s, uint32_t* codes, uint32_t* sorted_tris);
/* BVH8 arrays: nodes (n_nodes * node_bytes), leaf triangle records
 * (n_triangles * tri_bytes). Either pointer may be NULL. */
int wsrt_debug_read_bvh(const wsrt_scene* s, void* nodes, void* tris);
/* World triangles (n_triangles x 9 floats: v0,v1,v2) and per-triangle
 * (instance, material, mesh-local primitive) triples. */
int wsrt_debug_read_world(const wsrt_scene* s, float* tris, uint32_t* ids);

/* ---- autotuner (no reference counterpart: the reference renders one frame
 *      per submission) -------------------------------------------------------
 * Times candidate frames_in_flight values (1, 2, 4, .. 256: frames traced in
 * one wavefront, one graph per batch) for a scene, camera (null: the default
 * camera), resolution and config by
 * rendering 2 warm batches each, within a device-memory budget (bytes; 0 =
 * a quarter of the device), and writes the fastest into cfg->frames_in_flight.
 * Then, at that setting, times tail_paths in {0, 2^16, 2^18, 2^20} x
 * trace_rounds in {1, 2} and writes the fastest into cfg. Results do not
 * depend on any of them. */
typedef struct wsrt_autotune_result {
    uint32_t n_candidates;
    uint32_t frames_in_flight[12];
    float ms_per_frame[12];       /* 0: not run (over budget or allocation failed) */
    uint64_t device_bytes[12];    /* device memory the candidate's renderer took */
    uint32_t best;
    /* second stage, at the best frames_in_flight: the queue size below which
     * one launch finishes the remaining bounces (tail_paths) and the
     * traversal grids (trace_rounds) */
    uint32_t n_stage2;
    uint32_t tail_paths[8];
    uint32_t trace_rounds[8];
    float stage2_ms_per_frame[8]; /* 0: not run */
    uint32_t best_tail_paths, best_trace_rounds;
    /* third stage, at the second stage's choice: the closest-hit kernel's
     * lane refill threshold (extend_refill 0, 16, 24 and 32) with and without
     * path-ordered queues (ordered_queues) */
    uint32_t n_stage3;
    uint32_t extend_refill[8];
    uint32_t ordered_queues[8];
    float stage3_ms_per_frame[8]; /* 0: not run */
    uint32_t best_extend_refill, best_ordered_queues;
} wsrt_autotune_result;
int wsrt_autotune(wsrt_device* dev, const wsrt_scene* scene, uint32_t width, uint32_t height,
                  const wsrt_camera_uniform* cam, wsrt_config* cfg, uint64_t mem_budget,
                  wsrt_autotune_result* out);

/* ---- persistent multi-scene batch (no reference counterpart: the reference
 *      renders one scene per window) ------------------------------------------
 * Jobs of uneven scene, view, resolution and sample count rendered by one
 * persistent launch pulling 64-pixel work items from a device queue. Each
 * pixel's samples are traced in frame order with the renderer's per-path
 * code, so a job's accumulation equals a wsrt_renderer's with the same
 * scene, camera, size and frame count (guiding must be 0). With
 * enable_photons, each window of frames first runs every job's photon pass
 * and grid build for its frames, then one persistent launch traces them.
 * Every render restarts at frame 0. */
typedef struct wsrt_batch wsrt_batch;
typedef struct wsrt_batch_job {
    const wsrt_scene* scene;
    wsrt_camera_uniform camera;  /* frame_count is ignored: frames 0 .. spp-1 */
    uint32_t width, height;
    uint32_t spp;
    uint32_t pad;
} wsrt_batch_job;
typedef struct wsrt_batch_tuning {
    uint32_t tile_mode;      /* 0: 64 x 1 pixel work items, 1: 8 x 8 */
    uint32_t blocks_per_cu;  /* persistent blocks of 256 threads per CU */
} wsrt_batch_tuning;
int wsrt_batch_create(wsrt_device* dev, const wsrt_batch_job* jobs, uint32_t n_jobs, const wsrt_config* cfg,
                      wsrt_batch** out);
int wsrt_batch_destroy(wsrt_batch* b);
int wsrt_batch_set_tuning(wsrt_batch* b, const wsrt_batch_tuning* t);
/* Times every tile_mode x blocks_per_cu in {1, 2, 4, 8} candidate on a
 * one-sample-per-pixel pass of the batch and keeps the fastest. */
int wsrt_batch_autotune(wsrt_batch* b, wsrt_batch_tuning* best, float* best_ms);
/* All jobs, all samples: one persistent launch (captured as a graph after
 * the first call), asynchronous. */
int wsrt_batch_render(wsrt_batch* b);
/* Waits for the render; WSRT_E_ABORTED when it stopped early (its images are
 * then partial; the next render starts from empty images). */
int wsrt_batch_sync(wsrt_batch* b);
/* Stops the render in flight (callable from any thread): every wave drops
 * the next work item it takes and exits. The next render clears the flag. */
int wsrt_batch_cancel(wsrt_batch* b);
/* A wall-clock budget for each render from its start, in microseconds
 * (0: none): past it, waves stop taking work items. */
int wsrt_batch_set_time_limit(wsrt_batch* b, uint64_t microseconds);
int wsrt_batch_read_accum(wsrt_batch* b, uint32_t job, float* out);
int wsrt_batch_read_rgba8(wsrt_batch* b, uint32_t job, uint8_t* out);

/* ---- learned path guiding (no reference counterpart: the reference has no
 *      guiding; this is the online-trained proposal of wsrt_config.guiding) --
 * A network of 7 input scalars in [0,1] (position over the scene box, normal
 * mapped from [-1,1], hero wavelength over 380..780 nm), 12-band sin/cos
 * encoding, 4 hidden layers of 64 ReLU units and 40 outputs: 8 von
 * Mises-Fisher lobes of (mu xyz, log kappa, logit). Parameters are
 * WSRT_GUIDE_PARAMS floats: W1[64][192] b1[64] W2[64][64] b2 W3 b3 W4 b4
 * W5[48][64] b5[48], row-major [out][in]; W1 columns >= 168 and output rows
 * >= 40 are padding and stay zero. Forward, backward and Adam run as BF16
 * MFMA kernels with FP32 accumulation. */
#define WSRT_GUIDE_PARAMS 27952
#define WSRT_GUIDE_OUTPUTS 40
typedef struct wsrt_guide wsrt_guide;
/* A standalone network for up to max_samples samples per call, initialised
 * from seed. */
int wsrt_guide_create(wsrt_device* dev, uint32_t max_samples, uint32_t seed, wsrt_guide** out);
int wsrt_guide_destroy(wsrt_guide* g);
int wsrt_guide_get_params(wsrt_guide* g, float* out);
/* Sets the parameters (padding forced to zero) and resets the Adam state. */
int wsrt_guide_set_params(wsrt_guide* g, const float* in);
/* Outputs (n x 40) for input scalars (n x 7). */
int wsrt_guide_forward(wsrt_guide* g, const float* scalars, uint32_t n, float* out);
/* Gradient of (1/n) sum_i -weights[i] log p_guide(dirs[i] | scalars[i]) with
 * respect to the parameters (WSRT_GUIDE_PARAMS floats). */
int wsrt_guide_gradients(wsrt_guide* g, const float* scalars, const float* dirs, const float* weights,
                         uint32_t n, float* grads);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, learning rate lr) on that loss. */
int wsrt_guide_train_step(wsrt_guide* g, const float* scalars, const float* dirs, const float* weights,
                          uint32_t n, float lr);
/* The network a guiding renderer trains (owned by the renderer). */
int wsrt_renderer_get_guide(wsrt_renderer* r, wsrt_guide** out);
/* Checkpoint and resume of training (no reference counterpart: the
 * reference's only state is the accumulation and frame_count, which
 * wsrt_renderer_read/write_accum and get/set_frame_count cover). The state
 * is the parameters, both Adam moments (WSRT_GUIDE_PARAMS floats each) and
 * the Adam step, and the running average of the parameters the tracer
 * samples with (each Adam step moves it by 0.5 % toward the new parameters);
 * set_state restores the first three exactly and restarts the average at the
 * parameters, and set_average then restores the average. With the frames a
 * guiding renderer has trained on (guide_train_frames counts them), a
 * renderer restored at a batch boundary renders the frames that follow bit
 * for bit as the original would have. */
int wsrt_guide_get_state(wsrt_guide* g, float* params, float* m1, float* m2, uint32_t* step);
int wsrt_guide_set_state(wsrt_guide* g, const float* params, const float* m1, const float* m2, uint32_t step);
int wsrt_guide_get_average(wsrt_guide* g, float* out);
int wsrt_guide_set_average(wsrt_guide* g, const float* in);
int wsrt_renderer_get_guide_frames_trained(const wsrt_renderer* r, uint64_t* out);
int wsrt_renderer_set_guide_frames_trained(wsrt_renderer* r, uint64_t frames);

#ifdef __cplusplus
}
#endif

#endif /* WSRT_H */
