// batch.hip - many scenes and views of uneven size in one persistent launch.
//
// A work item is one 64-pixel tile of one job (a scene, a camera, a
// resolution and a sample count); items are numbered job by job. A fixed
// grid of blocks stays resident and each wave pulls the next item from a
// device counter until all are taken, so large and small jobs, cheap and
// expensive pixels, spread over the whole chip without a launch per job.
// Each lane traces its pixel's samples one after the other with the same
// per-path steps as the wavefront renderer (path_core.hpp), accumulating in
// frame order, so a job's image equals the wavefront renderer's bit for bit.
// With the photon pass, the launch traces a window of frames of every job,
// after the host has run each job's photon pass and grid build for those
// frames (one grid per frame of the window, cell_base apart).
#include "kernels.hpp"
#include "path_core.hpp"
#include "traverse.hpp"

namespace wsrt {
namespace {

constexpr int kBatchThreads = 256;

// Whether the render must stop taking work items: the time budget at every
// item, the host flag (a read over the host link, ~1 us) at every 32nd item
// a wave takes.
__device__ __forceinline__ bool batch_stopping(const BatchArgs& a, uint32_t taken) {
    if (a.stop && (taken & 31u) == 0u && *a.stop != 0u) return true;
    if (a.time_limit) {
        const uint64_t t0 = *reinterpret_cast<const volatile uint64_t*>(a.abort + 2);
        return (uint64_t)wall_clock64() - t0 > a.time_limit;
    }
    return false;
}

__global__ __launch_bounds__(kBatchThreads) void k_batch(BatchArgs a) {
    __shared__ uint2 lds_stack[kShortStack * kBatchThreads];
    __shared__ uint32_t first[kMaxBatchJobs + 1];
    __shared__ TriShare share_all[kBatchThreads / 64];
    TriShare& sh = share_all[threadIdx.x >> 6];
    for (uint32_t j = threadIdx.x; j <= a.njobs; j += blockDim.x)
        first[j] = j < a.njobs ? a.jobs[j].first_item : a.n_items;
    __syncthreads();
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = __lane_id();
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kBatchThreads;
    st.glob = a.gstack + (uint64_t)gtid * a.spill;
    RayTally tally;
    uint32_t taken = 0;  // items this wave has taken
    while (true) {
        uint32_t item = 0;
        if (lane == 0) {
            item = atomicAdd(a.head, 1u);
            if (item < a.n_items && (a.stop || a.time_limit) && batch_stopping(a, taken++)) {
                atomicAdd(a.abort, 1u);  // the item is left untraced
                item = 0xFFFFFFFFu;
            }
        }
        item = __shfl(item, 0, 64);
        if (item >= a.n_items) break;
        // job of the item: last j with first[j] <= item
        uint32_t lo = 0, hi = a.njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (first[mid] <= item) lo = mid;
            else hi = mid;
        }
        const BatchJob& J = a.jobs[lo];
        if (a.one_frame && a.frame >= J.frames) continue;  // the job has no such sample
        const uint32_t local = item - first[lo];
        uint32_t x, y;
        if (a.tile_mode == 0u) {  // 64 x 1 pixels
            const uint32_t pix = local * 64u + (uint32_t)lane;
            x = pix % J.width;
            y = pix / J.width;
        } else {  // 8 x 8 pixels
            const uint32_t tx = local % J.tiles_x, ty = local / J.tiles_x;
            x = tx * 8u + ((uint32_t)lane & 7u);
            y = ty * 8u + ((uint32_t)lane >> 3);
        }
        // lanes past the image's edge only help the others' traversals
        const bool valid = x < J.width && y < J.height;
        const uint32_t pix = valid ? y * J.width + x : 0u;
        const SceneView& sv = J.sv;
        float4 acc = valid ? J.accum[pix] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const uint32_t f1 = a.one_frame ? min(a.frame + a.frame_span, J.frames) : J.frames;
        for (uint32_t f = a.one_frame ? a.frame : 0u; f < f1; ++f) {
            const FrameConst& fc = a.table[J.table_base + f];
            PathRegs r;
            camera_path(fc, x, y, r);
            // the photon gather reads the job's grid built for this frame,
            // the (f - frame)-th of the window's
            const FrameConst* fs = (fc.enable_photons && J.grid.count) ? &fc : nullptr;
            const uint32_t cell_base = a.one_frame ? (f - a.frame) * fc.hash_size : 0u;
            trace_bounces_wave(sv, J.grid, fs, cell_base, fc.rr_after, fc.quirks, 0u, a.max_depth, r, st, sh, valid,
                               tally);
            acc = accumulate_sample(acc, fc, r.wl, r.L);
        }
        if (valid) {
            J.accum[pix] = acc;
            reinterpret_cast<uchar4*>(J.rgba)[pix] = tone_map(acc);
        }
    }
}

__global__ void k_batch_reset(uint32_t* head, uint32_t* abort, uint32_t restart) {
    if (threadIdx.x == 0) {
        *head = 0u;
        if (restart && abort) {
            abort[0] = 0u;
            *reinterpret_cast<uint64_t*>(abort + 2) = (uint64_t)wall_clock64();
        }
    }
}

}  // namespace

void launch_batch(const BatchArgs& a, uint32_t blocks, hipStream_t s, bool restart) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.head, a.abort, restart ? 1u : 0u);
    hipLaunchKernelGGL(k_batch, dim3(blocks), dim3(kBatchThreads), 0, s, a);
}

}  // namespace wsrt
