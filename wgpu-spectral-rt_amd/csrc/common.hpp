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
    uint32_t stack_size;           // traversal stack entries required (either BVH)
    // BVH over the triangles of metal, glass and thin-film instances (the
    // surfaces a photon must reach first to ever become caustic); its leaf
    // triangle ids are not world ids
    uint32_t spec_n_nodes;
    const wsrt_bvh8_node* spec_nodes;
    const float4* spec_tris;
};

// The scene seen through its specular BVH (for any-hit queries only).
__host__ __device__ inline SceneView specular_view(const SceneView& sv) {
    SceneView v = sv;
    v.nodes = sv.spec_nodes;
    v.tris = sv.spec_tris;
    v.n_nodes = sv.spec_n_nodes;
    return v;
}

// Per-frame constants uploaded before a frame (kernels read them from device
// memory so a captured frame graph stays valid across frames).
struct FrameParams {
    wsrt_camera_uniform cam;  // cam.frame_count is the frame being shaded
    float M[16];              // view_inverse * proj_inverse
    uint32_t width, height;
    uint32_t pad[2];
};

constexpr int kMatBins = 8;
constexpr int kMaxBounces = 64;
constexpr int kStatBlocks = 4096;  // blocks of the photon bounce launch (kTraceBlocks)

// Running statistics of a renderer. Totals over a whole launch are added
// once per launch; per-path totals are kept per producer block (plain
// read-modify-write by that block only) and summed on the host, so no
// kernel issues one atomic per wave on a shared address.
struct Counters {
    unsigned long long extend_rays;
    unsigned long long shadow_rays;
    unsigned long long guide_samples;
    // per launching kernel (RayKernel): rays traversed, always counted; nodes
    // visited and triangles tested, while counting
    unsigned long long kernel_rays[6];
    unsigned long long trav_nodes[6];
    unsigned long long trav_tris[6];
    unsigned long long trav_rays[6];
    // wavefront extend, while counting: per wave and round, the most node
    // visits of any lane times the lanes that traversed, summed
    unsigned long long extend_wave_nodes;
    // while counting: the most node visits of one ray, per launching kernel
    unsigned long long ray_max_nodes[6];
    unsigned long long photon_rays_part[kStatBlocks];
    unsigned long long photons_stored_part[kStatBlocks];
};

// Kernels that traverse the BVH, as indices of the per-kernel counters.
enum RayKernel : int {
    kRkPhotonBounce = 0,  // k_photon_bounce: closest hit
    kRkExtend = 1,        // k_extend: closest hit
    kRkShadow = 2,        // k_shadow: any hit
    kRkPhotonEmit = 3,    // k_photon_emit: the specular BVH root
    kRkTailExtend = 4,    // k_tail: closest hit
    kRkTailShadow = 5,    // k_tail: any hit
};

// Wavefront path state, structure of arrays indexed by path id (= pixel).
// Path flags of the guided tracer, for the Lambert vertices of bounces
// 0 .. kGuideRecordBounces - 1: kFlagRecorded << b marks bounce b's vertex as
// recorded for training; kFlagNeeVisible << b marks its light sample as
// unoccluded (its contribution is then excluded from the training target).
constexpr uint32_t kGuideRecordBounces = 4;
constexpr uint32_t kFlagRecorded = 4u, kFlagNeeVisible = 64u;

struct PathState {
    float4* o;          // origin xyz
    float4* d;          // direction xyz
    float4* tp;         // throughput per hero wavelength
    float4* L;          // accumulated spectral radiance
    float4* wl;         // hero wavelengths
    uint32_t* rng;      // pcg state
    uint32_t* flags;    // bit0 has_hit_diffuse, bit1 caustic path
    uint4* hit;         // t bits, tri | front<<31, u bits, v bits
    uint32_t* queue_a;  // path ids entering a bounce (segmented, 64 * seg_cap)
    uint32_t* queue_b;  // and the next bounce
    uint32_t* shade_q;  // hits per material type (kMatBins * 64 * seg_cap)
    float4* shadow;     // 3 float4 per shadow ray: o+tmax, d+path, contribution
    uint32_t* guide_q;  // Lambert vertices for guided scattering (segmented, 64 * seg_cap)
    float4* gn;         // forward-facing shading normal at a guided vertex
    float4* galb;       // spectral albedo at a guided vertex
    // ordered queues (null: off): tag[p] = b when path p enters bounce b;
    // the active queue of b is then built from the tags in path order
    uint8_t* tag;
    uint32_t capacity;  // path slots (frames in flight x pixels)
};

// Photon paths between bounces (structure of arrays indexed by photon path
// id).
// The photon paths of one depth's queue, in queue order: entry e holds the
// path id (also its photon slot) and the state of the ray it traces next,
// so the queue's producer writes and its consumer reads whole lines.
struct PhotonQueue {
    uint32_t* id;
    float4* pos;        // origin xyz, wavelength
    float4* dir;        // direction xyz, energy
    uint32_t* rng;
    uint32_t* flags;    // 1: caustic path
};
struct PhotonState {
    PhotonQueue q[2];   // depth d reads q[d & 1] and fills q[(d + 1) & 1]
    uint32_t* stored;   // bit i (word i / 32): photon path i stored its photon in slot i
};

// Photon grid (counting sort view of build_grid.wgsl).
struct GridView {
    const uint32_t* count;
    const uint32_t* offset;
    const wsrt_photon* sorted;  // photons in cell order
    uint32_t hash_size;
    float cell_size;
};

// Constants of one frame computed on the host (so the transcendental values
// are those of the host libm, identical frame to frame).
struct FrameConst {
    float cdf_min, cdf_range;  // hero-wavelength Cauchy CDF bounds
    float radius, radius_sq;   // gather radius r0 * (frame + 1)^exponent
    float density;             // 3/(pi r^2) * 1/spectral_radius * 1/photon_norm
    float jitter[3];           // grid jitter of this frame
    float spectral_radius;
    float cell_size;
    uint32_t hash_size;
    uint32_t gather_cap;
    uint32_t frame;
    uint32_t width, height;
    uint32_t max_depth, photon_max_depth, rr_after;
    uint32_t enable_photons;
    uint32_t n_photons;
    float M[16];               // view_inverse * proj_inverse
    float origin[3];           // camera position (view_inverse column 3)
    uint32_t count_traversal;  // 1: count node visits / triangle tests
    uint32_t quirks;           // wsrt_config.quirks (WSRT_FIX_* bits)
    // guided next-event estimation: probability of the uniform light sample
    // against the guide's light proposal (1: uniform only)
    float guide_light_mix;
};

constexpr uint32_t kFrameTable = 4096;  // entries; frames wrap modulo this
struct FrameArgs {
    const FrameConst* table;
    const uint32_t* frame;  // device frame counter
    uint32_t mask;          // table entries - 1
};

}  // namespace wsrt
