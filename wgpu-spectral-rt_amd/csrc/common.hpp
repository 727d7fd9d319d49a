// common.hpp - structures shared by the host side of libwsrt and its HIP
// kernels: device views of a scene, of a frame and of the wavefront state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wsrt.h"

namespace wsrt {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Per-instance record (the TLAS instance of builder.rs:237-248 plus what the
// driver derives from it: the world-to-object linear part).
struct Instance {
    uint32_t mesh;
    uint32_t material;
    uint32_t first_tri;  // first world triangle of this instance
    uint32_t n_tris;
    float m[16];         // object-to-world, column-major
    float w2o[9];        // inverse of the linear part, column-major
    uint32_t pad[3];
};

// Device view of a built scene (all pointers are device memory).
struct SceneView {
    const wsrt_bvh8_node* nodes;
    const float4* tris;          // 3 float4 per leaf triangle record
    const wsrt_vertex* verts;
    const uint32_t* indices;
    const wsrt_mesh_info* meshes;
    const wsrt_material* materials;
    const wsrt_light_info* lights;
    const Instance* instances;
    const uint32_t* tri_instance;  // world triangle -> instance
    const float* env;              // equirect RGB radiance or null
    uint32_t n_nodes;
    uint32_t n_tris;
    uint32_t n_lights;
    uint32_t n_instances;
    uint32_t env_w, env_h;
    uint32_t stack_size;           // traversal stack entries required
    uint32_t pad;
};

// Per-frame constants uploaded before a frame (kernels read them from device
// memory so a captured frame graph stays valid across frames).
struct FrameParams {
    wsrt_camera_uniform cam;  // cam.frame_count is the frame being shaded
    float M[16];              // view_inverse * proj_inverse
    uint32_t width, height;
    uint32_t pad[2];
};

// Device-side counters of one frame in flight.
struct Counters {
    uint32_t active;        // paths in the current extend queue
    uint32_t next;          // paths queued for the next bounce
    uint32_t shade[8];      // per material type (sort-by-material bins)
    uint32_t shadow;        // shadow rays queued
    uint32_t stored;        // photons stored this frame
    uint32_t valid;         // photons in the grid
    uint32_t pad[3];
    unsigned long long extend_rays;
    unsigned long long shadow_rays;
    unsigned long long photon_rays;
    unsigned long long photons_stored;
    unsigned long long guide_samples;
    unsigned long long trav_nodes[3];  // photon, extend, shadow
    unsigned long long trav_tris[3];
    unsigned long long trav_rays[3];
};

constexpr int kMatBins = 8;

// Wavefront path state, structure of arrays indexed by path id (= pixel).
struct PathState {
    float4* o;          // origin xyz
    float4* d;          // direction xyz
    float4* tp;         // throughput per hero wavelength
    float4* L;          // accumulated spectral radiance
    float4* wl;         // hero wavelengths
    uint32_t* rng;      // pcg state
    uint32_t* flags;    // bit0 has_hit_diffuse, bit1 caustic path
    uint4* hit;         // t bits, tri | front<<31, u bits, v bits
    uint32_t* queue_a;  // active path ids
    uint32_t* queue_b;  // next bounce
    uint32_t* shade_q;  // kMatBins * capacity
    float4* shadow;     // 3 float4 per shadow ray: o+tmax, d+path, contribution
    uint32_t capacity;
};

// Photon grid (counting sort view of build_grid.wgsl).
struct GridView {
    const uint32_t* count;
    const uint32_t* offset;
    const wsrt_photon* sorted;  // photons in cell order
    uint32_t hash_size;
    float cell_size;
};

}  // namespace wsrt
