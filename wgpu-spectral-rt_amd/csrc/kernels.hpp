// kernels.hpp - host-callable launchers of the HIP kernels of libwsrt.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dev_math.hpp"
#include "guide.hpp"
#include "queues.hpp"

namespace wsrt {


// Traversal launches use a fixed grid of at most kTraceBlocks x kTraceThreads
// lanes that loop over their work (so the per-lane HBM stack spill area has a
// fixed size): 4096 x 128 lanes = 32 waves on each of the 256 CUs.
constexpr int kTraceThreads = 128;
constexpr int kTraceBlocks = 4096;
// Compute units of the current device (cached per device ordinal).
int device_cus();
// Grid of a traversal kernel whose register budget keeps `waves` waves per
// SIMD resident: exactly one round of waves on the device's CUs x 4 SIMDs
// (blocks of 2 waves; 256 CUs on MI355X), so no partly filled second round
// trails the launch. With 5 waves per SIMD this runs the caustic workload's
// extend kernel 4 % faster than the full 4096-block grid. The block count is
// a multiple of the queue segment count, which the segment capacity bound
// needs, and at most kTraceBlocks (the stack spill area).
inline int trace_grid(int waves, uint32_t rounds = 1) {
    const int g = device_cus() * 4 * waves * (int)(rounds ? rounds : 1) / (kTraceThreads / 64) / kSegs * kSegs;
    return g < kSegs ? kSegs : (g > kTraceBlocks ? kTraceBlocks : g);
}
constexpr uint64_t kTraceLanes = (uint64_t)kTraceThreads * kTraceBlocks;
// Register budgets of the traversal kernels, as waves per SIMD the compiler
// must fit (gfx950: 512 VGPRs per lane and SIMD). Left free, the compiler
// gives the extend kernel 122 VGPRs (4 waves); with 5 it fits 96 with a few
// spills outside the node loop and runs 5 % faster on the caustic workload,
// while 6 or 8 waves spill inside the loop and run slower. The photon bounce
// fits 4 waves (128 VGPRs) without spills, but with its grid sized to
// resident waves 5 waves (96 VGPRs, a few spills) run it 5 % faster. The
// any-hit shadow kernel runs 4 % faster at 6 waves (80 VGPRs) than at the
// compiler's 5.
#define WSRT_WAVES_PER_SIMD(n) __attribute__((amdgpu_waves_per_eu(n)))

// ---- primitives (sort.hip) -------------------------------------------------
size_t scan_temp_words(uint32_t n);
int scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* temp,
                   uint32_t* total_out, hipStream_t s);
size_t radix_temp_words(uint32_t n);
// Stable LSD sort of (key, value) pairs by the low `bits` bits of the keys.
// n_dev (optional): device-resident element count <= n.
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits, uint32_t* temp,
                     hipStream_t s, bool* result_in_temp, const uint32_t* n_dev = nullptr);
void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
void iota_u32(uint32_t* p, uint32_t n, hipStream_t s);

// ---- BVH build (bvh.hip) ------------------------------------------------------
struct BvhInputs {
    const Instance* instances;
    uint32_t n_instances;
    const wsrt_vertex* verts;
    const uint32_t* indices;
    const wsrt_mesh_info* meshes;
    uint32_t n_tris;
};
struct BvhOutputs {
    wsrt_bvh8_node* nodes;
    float4* tris;
    float* wv;           // world vertices, 9 floats per triangle
    uint32_t* tri_inst;  // world triangle -> instance
    uint32_t* codes;     // Morton code per triangle
    uint32_t* sorted;    // triangles in Morton order
    uint32_t n_nodes;
    uint32_t max_depth;
    uint32_t placed_tris;
};
int bvh_build(const BvhInputs& in, BvhOutputs* out, hipStream_t s);

// ---- ray queries (query.hip) ------------------------------------------------
// rays: n x 8 floats (o, tmin, d, tmax). mode 0 closest, 1 any, 2 counts.
void launch_query(const SceneView& sv, const float* rays, uint32_t n, int mode, uint32_t* out0,
                  uint32_t* out1, uint2* gstack, hipStream_t s);

// ---- photon pass + grid (photon.hip) ---------------------------------------
struct GridBuffers {
    uint32_t* count;        // hash_size
    uint32_t* offset;       // hash_size
    uint32_t* cell_id;      // n
    uint32_t* sorted_slot;  // n
    wsrt_photon* sorted;    // n (null: grid_build writes the sorted slots only)
    uint32_t* n_valid;      // 1
    uint32_t* temp;         // grid_temp_words(n, hash_size)
    // renderer grid: the cell of each photon of the last build (n_valid of
    // them), so the next build clears only those counts
    uint32_t* touched = nullptr;
};
// Photon paths of nf consecutive frames, np per frame: path i is photon
// i % np of frame slot i / np, stored at slot i with bit i of ps.stored set.
// Emission (which queues the paths whose first ray meets specular geometry),
// then one launch per bounce depth.
void launch_photon_emit(const SceneView& sv, const FrameArgs& fa, uint32_t np, uint32_t nf, const Queues& qs,
                        const PhotonState& ps, uint32_t count_traversal, Counters* ctr, hipStream_t s);
void launch_photon_bounce(const SceneView& sv, uint32_t depth, uint32_t count_traversal, const PhotonState& ps,
                          const Queues& qs, wsrt_photon* photons, Counters* ctr, uint2* gstack, hipStream_t s);
size_t grid_temp_words(uint32_t n, uint32_t cells);
// Writes fc to *dst and sets *frame = 0, on the device (no host copy).
void launch_store_frame(const FrameConst& fc, FrameConst* dst, uint32_t* frame, hipStream_t s);
// p[k] = v for *start <= k < n (start read on the device).
void launch_fill_after(uint32_t* p, const uint32_t* start, uint32_t n, uint32_t v, hipStream_t s);
// One counting sort over the grids of nf frames: cell (frame slot, hash)
// = slot * hash_size + hash; count/offset hold nf * hash_size cells.
int grid_build(const wsrt_photon* photons, uint32_t np, uint32_t nf, uint32_t hash_size,
               const FrameArgs& fa, GridBuffers& g, hipStream_t s);
// The same grid over the slots whose `stored` bit is set (the photon pass's
// output; other slots are not read), with offsets written for non-empty
// cells only: the gather reads a cell's offset only when its count is not 0.
// The counts must be zero except at g.touched[0 .. *g.n_valid) (all zero on
// allocation); each build leaves them so for the next.
int grid_build_stored(const wsrt_photon* photons, const uint32_t* stored, uint32_t np, uint32_t nf,
                      uint32_t hash_size, const FrameArgs& fa, GridBuffers& g, hipStream_t s);

// ---- path guiding (guide.hip) ------------------------------------------------
// Device view of the guiding network used while tracing.
struct GuideView {
    const uint16_t* wf;     // packed forward operands (BF16 bits)
    const uint16_t* wt;     // packed backward operands
    const float* params;    // fp32 parameters (biases are read from here)
    float mix;              // one-sample MIS probability of the guide
    float bmin[3], binv[3]; // scene box for the position inputs
    guide::Records rec;     // training records (null: none)
};
// Buffers of a training step.
struct GuideTrain {
    uint32_t* act;          // guide_act_words(ns) words
    float* partial;         // (ns / chunk) x kParams
    float* params;          // fp32 master copy
    float* m1;              // Adam moments
    float* m2;
    float* avg;             // running average of params (inference), or null
    uint16_t* wf_avg;       // its packed forward operands
    uint32_t* step;         // Adam step counter (device)
    uint32_t chunk;         // samples per partial sum (divides ns, multiple of 32)
    float lr;
};
size_t guide_act_words(uint32_t ns);
void launch_guide_scatter(const SceneView& sv, const PathState& ps, const Queues& qs, const GuideView& gv,
                          const FrameArgs& fa, uint32_t bounce, uint32_t* next_queue, Counters* ctr,
                          hipStream_t s);
// lights the guided light proposal tests one by one (more: uniform light
// sampling only)
constexpr uint32_t kGuideLightsMax = 64;
// Writes the packed operands of params (after an upload or an initialisation).
void launch_guide_pack(const GuideView& gv, float* params, hipStream_t s);
// One step over ns samples (a multiple of 64 and of gt.chunk), the loss
// averaged with factor norm: gradients into gt.partial; with update, Adam
// and repacking.
void launch_guide_train_step(const GuideView& gv, const GuideTrain& gt, const float4* in0, const float4* in1,
                             const float4* dir, uint32_t ns, float norm, bool update, hipStream_t s);
void launch_guide_prepare(const PathState& ps, const GuideView& gv, const FrameArgs& fa, uint32_t npx,
                          uint32_t slot, uint32_t ns, uint32_t stride, float4* in0, float4* in1, float4* dir,
                          hipStream_t s);
void launch_guide_infer(const GuideView& gv, const float4* in0, const float4* in1, uint32_t n, float* out,
                        hipStream_t s);

// ---- persistent multi-scene batch (batch.hip) ---------------------------------
constexpr uint32_t kMaxBatchJobs = 256;
struct BatchJob {
    SceneView sv;          // env null unless the config lights misses with it
    GridView grid;         // the job's photon grid of the frame being traced (count null: no photon pass)
    float4* accum;         // width x height
    uint8_t* rgba;         // width x height x 4
    uint32_t width, height;
    uint32_t frames;       // samples per pixel rendered by one batch launch
    uint32_t table_base;   // FrameConst of frame f at table[table_base + f]
    uint32_t first_item;   // first work item of the job
    uint32_t tiles_x;      // 8 x 8 tiles per row
};
struct BatchArgs {
    const BatchJob* jobs;
    uint32_t njobs;
    const FrameConst* table;
    uint32_t n_items;      // work items over all jobs
    uint32_t* head;        // next work item (device counter)
    uint2* gstack;         // per-lane traversal stack spill
    uint32_t spill;        // spill entries per lane
    uint32_t max_depth;
    uint32_t tile_mode;    // 0: 64 x 1 pixel items, 1: 8 x 8 pixel items
    uint32_t one_frame;    // 1: trace frames [frame, frame + frame_span) only (those a job has)
    uint32_t frame;
    uint32_t frame_span;   // frames of the window; the job's grid holds one per frame, in order
    // stopping a render early: a host-written flag in pinned memory (null:
    // none) and a wall-clock budget from the render's start (0: none). A
    // wave that takes an item after either fires drops it and counts it in
    // abort[0]; abort[2..3] hold the start time (wall-clock ticks).
    const volatile uint32_t* stop;
    uint32_t* abort;
    uint64_t time_limit;
};
// restart: the launch starts a render (clears abort[0], stamps the start).
void launch_batch(const BatchArgs& a, uint32_t blocks, hipStream_t s, bool restart);

// ---- wavefront camera paths (path.hip) ----------------------------------------
struct PathLaunch {
    SceneView sv;
    PathState ps;
    GridView grid;
    Counters* ctr;
    uint2* gstack;
    float4* accum;
    uint8_t* rgba;
    uint32_t count_traversal;
    FrameArgs fa;
    Queues qs;
    uint32_t guided;  // 1: Lambert vertices scatter through k_guide_scatter
    uint32_t extend_refill;  // > 0: k_extend_refill, refilling when this many lanes idle
    uint32_t guide_mis;      // guided: the shadow kernel flags the light samples it finds visible
    GuideView gv;
};
void launch_clear_queues(const Queues& qs, uint32_t depths, hipStream_t s);
void launch_advance_frame(uint32_t* frame, uint32_t nf, hipStream_t s);
void launch_set_frame(uint32_t* frame, uint32_t value, hipStream_t s);
// A batch of nf consecutive frames of npx pixels shares one wavefront.
void launch_generate(const PathLaunch& pl, uint32_t npx, uint32_t nf, hipStream_t s);
// bounce 0 reads path ids 0 .. n0-1 directly
void launch_extend(const PathLaunch& pl, uint32_t bounce, uint32_t n0, const uint32_t* queue, hipStream_t s);
void launch_shade(const PathLaunch& pl, uint32_t npx, uint32_t bounce, uint32_t* next_queue, hipStream_t s);
// Active queue of bounce `next` in path order from the path tags (ps.tag
// set), after the shading of bounce next - 1; n = path slots in use.
void launch_order_active(const PathLaunch& pl, uint32_t next, uint32_t n, uint32_t* queue, hipStream_t s);
void launch_shadow(const PathLaunch& pl, uint32_t bounce, hipStream_t s);
// The camera paths the wavefront left to the tail launch (Queues::tail),
// each finished by one lane.
void launch_tail(const PathLaunch& pl, uint32_t npx, uint32_t max_depth, hipStream_t s);
void launch_accumulate(const PathLaunch& pl, uint32_t npx, uint32_t nf, hipStream_t s);

// ---- presented frame (present.hip) -------------------------------------------
struct PresentArgs {
    float view[16], proj[16];  // camera matrices, column-major
    float sx, sy;              // quad scale of the blit (aspect fit)
    uint32_t win_w, win_h, img_w, img_h;
    // 4 x 256 bytes: the sRGB encoding of an 8-bit linear value, then for
    // r, g, b the 8-bit sRGB value after one overlay point is blended in
    const uint8_t* lut;
    uint8_t point_alpha;
};
// Blit of the RGBA8 image (img_w x img_h) into out (win_w x win_h RGBA8,
// row 0 = window top); with count (win_w x win_h zeroed words) the photons
// of slots 0 .. n_photons-1 whose stored bit (first + slot) is set are drawn
// over it.
void launch_present(const uint8_t* image, const wsrt_photon* photons, const uint32_t* stored, uint32_t first,
                    uint32_t n_photons, const PresentArgs& a, uint32_t* count, uint8_t* out, hipStream_t s);

}  // namespace wsrt
