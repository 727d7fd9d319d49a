/*
 * wsrt_oracle.h - CPU oracle of the render loop of wgpu-spectral-rt.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's WGSL
 * shaders (src/renderer/shaders/{shader,photon_emit,build_grid}.wgsl) and
 * scene code (the src/scene Rust files) in plain C, f32, with every operation rounded
 * as written (-ffp-contract=off). Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py load it, and only as the checker of the
 * HIP path; the product library never links or calls it.
 *
 * The header of each test module states what pins each function.
 */
#ifndef WSRT_ORACLE_H
#define WSRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/wsrt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* ---- RNG / spectral KAT entry points ---- */
uint32_t or_pcg_hash(uint32_t x);
/* n successive rand() values of the stream seeded with pcg_hash(seed_input). */
void or_rand_stream(uint32_t seed_input, uint32_t n, float* out);
void or_cie_xyz_f(float lambda, float* out3);
float or_rgb_to_spectrum_f(const float* rgb3, float lambda);
float or_blackbody_f(float lambda, float temp_k);
float or_thin_film_f(float cos1, float n1, float n2, float n3, float d, float lambda);
float or_schlick_f(float cosine, float ref_idx);
void or_cauchy_constants(float* cdf_min, float* cdf_range, float* pdf_norm);
float or_cauchy_sample_f(float u01);
void or_grid_jitter_f(uint32_t frame, float cell, float* out3);
uint32_t or_hash_position_f(const float* p3, uint32_t frame, float cell, uint32_t hash_size);
/* Hero wavelengths of a camera path after jitter (shader.wgsl:734-767):
 * consumes rand for u and rot from the given state. */
void or_hero_wavelengths(float u_rand, float rot_rand, float* out4);

/* ---- geometry (src/scene/geometry.rs) ---- */
int or_geometry_plane(wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni);
int or_geometry_cube(wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni);
int or_geometry_sphere(uint32_t subdiv, wsrt_vertex* v, uint32_t* idx, uint32_t* nv, uint32_t* ni);

/* ---- glam-style matrices (column-major float[16]) ---- */
void or_mat4_identity(float* m);
void or_mat4_mul(const float* a, const float* b, float* out);
void or_mat4_translation(float x, float y, float z, float* m);
void or_mat4_scale(float x, float y, float z, float* m);
void or_mat4_rotation_x(float a, float* m);
void or_mat4_rotation_y(float a, float* m);
void or_mat4_rotation_z(float a, float* m);
void or_camera_uniform(const float* pos3, float yaw, float pitch, float aspect, uint32_t frame,
                       wsrt_camera_uniform* out);

/* ---- scene ---- */
or_scene* or_scene_create(void);
void or_scene_destroy(or_scene* s);
uint32_t or_scene_add_material(or_scene* s, const wsrt_material* m);
uint32_t or_scene_add_mesh(or_scene* s, const wsrt_vertex* v, uint32_t nv, const uint32_t* idx,
                           uint32_t ni);
void or_scene_add_instance(or_scene* s, uint32_t mesh, uint32_t mat, const float* m16);
void or_scene_add_light_instance(or_scene* s, uint32_t mesh, uint32_t mat, const float* m16,
                                 const float* params4);
void or_scene_add_point_light(or_scene* s, const float* pos3, const float* params4);
void or_scene_set_environment(or_scene* s, const float* rgb, uint32_t w, uint32_t h);
/* Transforms instances to world triangles and builds the oracle's own BVH. */
int or_scene_build(or_scene* s);
uint32_t or_scene_num_triangles(const or_scene* s);
uint32_t or_scene_num_lights(const or_scene* s);
void or_scene_get_lights(const or_scene* s, wsrt_light_info* out);
void or_scene_get_world(const or_scene* s, float* tris9, uint32_t* ids3);
/* cornell_box.rs:6-110 */
void or_scene_cornell_box(or_scene* s);

/* ---- ray queries ---- */
/* rays: n x 8 (o, tmin, d, tmax). hits: n x 4 (t bits, tri, u bits, v bits);
 * flags: bit0 hit, bit1 front face. brute = 1 tests every triangle. */
void or_closest_hit(const or_scene* s, const float* rays, uint32_t n, int brute, uint32_t* hits,
                    uint32_t* flags);
void or_any_hit(const or_scene* s, const float* rays, uint32_t n, int brute, uint32_t* occluded);
/* Traverses a BVH8 produced by the HIP path (node/tri arrays as in wsrt.h)
 * with the same visiting order, returning hits and per-ray visit counts. */
void or_bvh8_closest(const wsrt_bvh8_node* nodes, uint32_t n_nodes, const wsrt_bvh_tri* tris,
                     const float* rays, uint32_t n, uint32_t* hits, uint32_t* flags,
                     uint32_t* node_visits, uint32_t* tri_tests);
/* Morton codes of the world triangles (centroid of the triangle AABB over
 * the centroid bounds, 10 bits per axis). */
void or_morton_codes(const or_scene* s, uint32_t* codes);

/* ---- photon pass + grid ---- */
/* Traces photon paths [first, first+n) of frame `frame`; out[k] is the
 * stored photon of path first+k or energy 0. Returns the number stored. */
uint32_t or_photon_trace(const or_scene* s, const wsrt_config* cfg, uint32_t frame,
                         uint32_t first, uint32_t n, wsrt_photon* out);
/* Counting-sort view of build_grid: cell_id[n], count/offset[hash_size],
 * sorted_slot[n] (stable by slot). Returns the number of valid photons. */
uint32_t or_grid_counting(const wsrt_photon* ph, uint32_t n, uint32_t frame, float cell,
                          uint32_t hash_size, uint32_t* cell_id, uint32_t* count,
                          uint32_t* offset, uint32_t* sorted_slot);
/* Linked-list view (build_grid.wgsl:54-73) with inserts in slot order. */
void or_grid_linked(const wsrt_photon* ph, uint32_t n, uint32_t frame, float cell,
                    uint32_t hash_size, uint32_t* head, uint32_t* next);

/* ---- camera paths ---- */
typedef struct or_frame or_frame;
/* Prepares a frame: photon pass (all photons_per_pass paths) and grid. */
or_frame* or_frame_create(const or_scene* s, const wsrt_config* cfg,
                          const wsrt_camera_uniform* cam, uint32_t width, uint32_t height);
/* The same with the photon buffer given (n slots) instead of traced. */
or_frame* or_frame_create_with_photons(const or_scene* s, const wsrt_config* cfg,
                                       const wsrt_camera_uniform* cam, uint32_t width,
                                       uint32_t height, const wsrt_photon* photons, uint32_t n);
void or_frame_destroy(or_frame* f);
uint32_t or_frame_photons_stored(const or_frame* f);
/* Traces the camera paths of the given pixels: wavelengths[4], radiance[4]
 * per path and linear sRGB[3]. */
void or_frame_trace_pixels(const or_frame* f, const uint32_t* pixels, uint32_t n, float* wl,
                           float* radiance, float* rgb);
/* Full frame: accum (width*height*4) updated as shader.wgsl:801-808;
 * rgba8 optional. */
void or_frame_render(const or_frame* f, float* accum, uint8_t* rgba8);
/* Tone map of n accumulated pixels (shader.wgsl:809-822) to RGBA8. */
void or_tone_map_f(const float* accum, uint32_t n, uint8_t* rgba8);
/* Photon gather at a point (shader.wgsl:297-342) against the frame's grid. */
void or_frame_gather(const or_frame* f, const float* pos3, const float* n3, const float* wl4,
                     float* out4);

/* ---- taps of single steps (for the pins of the transport code) ---- */
/* n directions of random_unit_vector (shader.wgsl:162-169) from the stream
 * seeded with pcg_hash(seed). */
void or_unit_vectors_f(uint32_t seed, uint32_t n, float* out3);
/* sample_bsdf (shader.wgsl:347-471) for incoming direction rdir at a hit
 * with forward-facing normal ffn; rng is the state before the call; quirks
 * as wsrt_config.quirks. Returns 1 when the path is absorbed. */
int or_sample_bsdf_f(const float* rdir, const float* ffn, const wsrt_material* m, int front,
                     const float* wl4, const float* tp4, uint32_t rng, uint32_t quirks,
                     float* dir_out3, float* att_out4, uint32_t* rng_out);
/* sample_random_light (shader.wgsl:477-528) from point hp. Returns 1 for a
 * point light (a delta distribution: pdf is then the distance squared over
 * the light count). pdf 0: no sample. */
int or_sample_light_f(const or_scene* s, const float* hp3, const float* wl4, uint32_t rng,
                      float* dir3, float* dist, float* pdf, float* rad4, uint32_t* rng_out);
/* Emission of photon path id of a frame (photon_emit.wgsl:159-208): point,
 * direction, wavelength, energy (energy 0 when the scene has no light). */
void or_photon_emit_f(const or_scene* s, uint32_t frame, uint32_t id, float* pos3, float* dir3,
                      float* wl, float* energy);
/* Russian roulette of a path's throughput (shader.wgsl:696-700). Returns 1
 * when the path survives. */
int or_roulette_f(const float* tp4, uint32_t rng, uint32_t quirks, float* tp_out4, uint32_t* rng_out);
/* ray_color (shader.wgsl:547-703) of one ray with given hero wavelengths
 * and rng state, in the frame's scene and photon grid. */
void or_ray_color_f(const or_frame* f, const float* o3, const float* d3, const float* wl4,
                    uint32_t rng, float* out4);

#ifdef __cplusplus
}
#endif

#endif
