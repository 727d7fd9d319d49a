// path.hip - the camera pass of shader.wgsl:547-823 as wavefront kernels.
//
// The reference runs one megakernel thread per pixel. Here a frame is
//   generate  -> [extend -> shade -> shadow] x max_depth -> accumulate
// with the path state in HBM (structure of arrays indexed by pixel):
//   extend  closest hit of every queued path; hits are binned by material
//           type (sort-by-material) into segmented queues (queues.hpp);
//   shade   emission, photon gather, light sample, BSDF sample, roulette of
//           each bin, so a wave mostly runs one material's code; surviving
//           paths are compacted into the next queue;
//   shadow  any-hit test of the deferred light samples, adding their
//           contribution when unoccluded.
// The per-path random stream is consumed in the reference's order (the
// shadow test draws no numbers), and contributions are added to a path's
// radiance in the reference's order (gather, light sample, then the next
// bounce), so each path's result matches the per-pixel reference.
#include "kernels.hpp"
#include "queues.hpp"
#include "path_core.hpp"
#include "shading.hpp"

namespace wsrt {

namespace {

constexpr int kShadeThreads = 256;
constexpr int kShadeBlocks = (int)(kGridLanes / kShadeThreads);
static_assert((uint64_t)kTraceBlocks * kTraceThreads <= kGridLanes, "queue producers use at most kGridLanes");

// Clears the queue counters of a batch (a fill kernel, so it can be
// captured in a graph).
__global__ void k_clear_queues(uint32_t* counts, uint32_t words, uint32_t* tail) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
        counts[i] = 0u;
    if (tail && blockIdx.x == 0 && threadIdx.x == 0) *tail = 0u;
}

// Last node of a batch: the next batch starts at frame_count + nf.
__global__ void k_advance_frame(uint32_t* frame, uint32_t nf) {
    if (threadIdx.x == 0) *frame = *frame + nf;
}

__global__ void k_set_frame(uint32_t* frame, uint32_t value) {
    if (threadIdx.x == 0) *frame = value;
}

// Path p of a batch of frames belongs to frame slot p / npx, pixel p % npx.
// The camera ray, radiance (0), wavelengths and flags are stored for the
// kernels that read them at bounce 0 (extend, the guide's record selection)
// or at the end (accumulate); the throughput (1) and the RNG state after the
// camera's draws are not: the shade of bounce 0 recomputes the whole camera
// path from its pixel and frame instead of reading it back (camera_regs).
__global__ __launch_bounds__(256) void k_generate(PathState ps, FrameArgs fa, uint32_t npx, uint32_t nf) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx * nf) return;
    const uint32_t slot = p / npx, pix = p - slot * npx;
    const FrameConst& fc = fa.table[(*fa.frame + slot) & fa.mask];
    PathRegs r;
    camera_path(fc, pix % fc.width, pix / fc.width, r);
    ps.o[p] = r.o4;
    ps.d[p] = r.d4;
    ps.L[p] = r.L;
    ps.wl[p] = r.wl;
    ps.flags[p] = r.flags;
    if (ps.tag) ps.tag[p] = 0;  // tags of an earlier batch are stale
}

// The state of camera path p before its first vertex, as k_generate made it.
__device__ __forceinline__ void camera_regs(const FrameArgs& fa, uint32_t base_frame, uint32_t npx, uint32_t p,
                                            PathRegs& r) {
    const uint32_t slot = p / npx, pix = p - slot * npx;
    const FrameConst& fc = fa.table[(base_frame + slot) & fa.mask];
    camera_path(fc, pix % fc.width, pix / fc.width, r);
}

// Radiance of path p after its ray escapes to the environment (radiance
// env along it): L + tp env, where a camera ray (bounce 0) has L = 0 and
// tp = 1, so its result is env itself and tp is not read.
__device__ __forceinline__ float4 env_gain(const PathState& ps, uint32_t p, uint32_t bounce, float4 env) {
    return bounce == 0 ? env : add4(ps.L[p], mul4(ps.tp[p], env));
}

// Bounce 0 reads path ids 0 .. n0-1 directly (every path starts); later
// bounces read the compacted queue of the previous shade.
// kCount: the traversal-counting build, a kernel of its own.
template <bool kCount>
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(5) void k_extend(SceneView sv, PathState ps, Queues qs,
                                                          Counters* __restrict__ ctr, uint32_t bounce,
                                                          uint32_t n0, const uint32_t* __restrict__ queue,
                                                          uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    __shared__ TriShare share_all[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t n = bounce == 0 ? n0 : seg_prefix(qs, bounce, kQActive, 1, pre);
    if (qs.tail && bounce >= kTailFirstBounce && (*qs.tail != 0u || n < qs.tail_paths)) {
        if (gtid == 0 && *qs.tail == 0u) *qs.tail = bounce + 1u;  // the tail launch takes these paths
        return;
    }
    if (gtid == 0) {
        atomicAdd(&ctr->extend_rays, (unsigned long long)n);
        atomicAdd(&ctr->kernel_rays[kRkExtend], (unsigned long long)n);
    }
    TriShare& sh = share_all[threadIdx.x >> 6];
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool count = kCount;
    uint32_t my_rays = 0, my_nodes = 0, my_tris = 0, my_max = 0;
    // wave-uniform trip count, so the binning ballots see the whole wave.
    // Bounce 0 takes path ids in pixel order; a plain grid stride would give
    // a wave the same pixel chunk of every frame in flight, so wave w takes
    // chunk (w + 977 k) mod W of round k instead (W waves), which spreads
    // expensive image regions over all waves.
    const uint32_t nw = stride >> 6, wave = gtid >> 6;
    for (uint32_t k = 0, base0 = 0; base0 < n; ++k, base0 += stride) {
        const uint32_t chunk = bounce == 0 ? (wave + 977u * k) % nw : wave;
        const uint32_t base = base0 + (chunk << 6);
        if (base >= n) continue;
        const uint32_t i = base + __lane_id();
        bool have = i < n;
        uint32_t p = 0, bin = 0, ray_nodes = 0, ray_tris = 0;
        const uint32_t s0 = bounce == 0 ? 0u : seg_of_wave_base(pre, base);  // whole wave
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        if (have) {
            p = bounce == 0 ? i : queue[seg_entry_from(pre, s0, i, qs.seg_cap)];
            o = ps.o[p];
            d = ps.d[p];
        }
        // the wave traverses its rays together, triangle tests shared
        const Hit h = trace_wave<kCount>(sv, xyz(o), xyz(d), 0.0001f, 100.0f, have, st, sh, ray_nodes, ray_tris);
        if (have) {
            my_nodes += ray_nodes;
            my_tris += ray_tris;
            my_max = max(my_max, ray_nodes);
            my_rays++;
            if (h.tri == kInvalid) {
                have = false;  // miss: the path ends (shader.wgsl:566)
                if (sv.env) ps.L[p] = env_gain(ps, p, bounce, env_radiance(sv, xyz(d), ps.wl[p]));
            } else {
                ps.hit[p] = make_uint4(__float_as_uint(h.t), h.tri | (h.front ? 0x80000000u : 0u),
                                       __float_as_uint(h.u), __float_as_uint(h.v));
                uint32_t mat = sv.instances[sv.tri_instance[h.tri]].material;
                uint32_t type = (uint32_t)sv.materials[mat].extra[0];
                bin = type < (uint32_t)kMatBins ? type : 0u;
            }
        }
        if (count) {
            const uint32_t w = wave_lockstep_nodes(ray_nodes, i < n);
            if (__lane_id() == 0 && w) atomicAdd(&ctr->extend_wave_nodes, (unsigned long long)w);
        }
        // sort-by-material: one segmented queue per material type
        uint64_t pending = __ballot(have);
        while (pending) {
            const uint32_t leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t b = __shfl(bin, leader);
            const bool mine = have && bin == b;
            const uint32_t e = seg_append(qs, bounce, kQShade0 + b, qs.seg_cap, mine);
            if (mine) {
                ps.shade_q[(size_t)b * kSegs * qs.seg_cap + e] = p;
                have = false;
            }
            pending &= ~__ballot(mine);
        }
    }
    if (count) flush_traversal_counts(ctr, kRkExtend, my_rays, my_nodes, my_tris, my_max);
}

// The same extend with lanes refilled as their rays end. A wave owns the
// rays k_extend gives it (round k: 64 queue entries at k * stride + chunk),
// taken in that order from a wave-uniform cursor; when at least `idle_min`
// lanes idle, the idle lanes take the next rays, so a wave no longer waits
// on its longest ray before starting new ones. Each loop iteration is one
// traversal step of every lane holding a ray, with the step's triangle tests
// spread over all 64 lanes, idle ones included (trav_step_wave). A finished ray's
// hit is written at once; its (path, material bin) waits in the wave's LDS
// buffer until 64 have finished, then the 64 are sorted by path and binned
// together (one segmented append per bin, as in k_extend), so each bin's
// queue stays in path order within a group. Hits are the same as
// k_extend's; only the order of paths within the shade queues differs, and
// paths are independent.
struct DoneBuf {
    uint32_t p[128];
    uint32_t tri[128];  // the hit triangle; its material bin is looked up when the group is binned
};

// Material bin of a hit triangle (one queue per material type).
__device__ __forceinline__ uint32_t tri_bin(const SceneView& sv, uint32_t tri) {
    const uint32_t mat = sv.instances[sv.tri_instance[tri]].material;
    const uint32_t type = (uint32_t)sv.materials[mat].extra[0];
    return type < (uint32_t)kMatBins ? type : 0u;
}

__device__ __forceinline__ void bin_paths(const Queues& qs, const PathState& ps, uint32_t bounce, bool have,
                                          uint32_t p, uint32_t bin) {
    uint64_t pending = __ballot(have);
    while (pending) {
        const uint32_t leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t b = __shfl(bin, leader);
        const bool mine = have && bin == b;
        const uint32_t e = seg_append(qs, bounce, kQShade0 + b, qs.seg_cap, mine);
        if (mine) {
            ps.shade_q[(size_t)b * kSegs * qs.seg_cap + e] = p;
            have = false;
        }
        pending &= ~__ballot(mine);
    }
}

// Sorts the wave's 64 (key, value) pairs by key, ascending across lanes
// (bitonic network over lane shuffles).
__device__ __forceinline__ void wave_sort64(uint32_t& key, uint32_t& val) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t ok = (uint32_t)__shfl_xor((int)key, (int)j, 64);
            const uint32_t ov = (uint32_t)__shfl_xor((int)val, (int)j, 64);
            const bool up = (lane & k) == 0;       // this pair sorts ascending
            const bool low = (lane & j) == 0;      // this lane holds the lower position
            const bool take = low == up ? ok < key : ok > key;
            if (take) {
                key = ok;
                val = ov;
            }
        }
    }
}

template <bool kCount>
__device__ __forceinline__ void extend_refill_loop(const SceneView& sv, const PathState& ps, const Queues& qs,
                                                   Counters* __restrict__ ctr, uint32_t bounce, uint32_t n,
                                                   const uint32_t* __restrict__ queue, const uint32_t* pre,
                                                   Stack st, DoneBuf& db, TriShare& sh, uint32_t idle_min) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t nw = stride >> 6, wave = gtid >> 6, lane = __lane_id();
    const uint32_t seq_len = (n + stride - 1) / stride * 64u;  // the wave's rounds x 64 positions
    uint32_t cursor = 0, ndone = 0, iters = 0;                 // wave-uniform
    uint32_t my_rays = 0, my_nodes = 0, my_tris = 0, my_max = 0;
    bool active = false;
    uint32_t p = 0;
    TravState ts;
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (cursor < seq_len && (nidle >= idle_min || nidle == 64u)) {
            if (!active) {
                const uint32_t q = cursor + (uint32_t)__popcll(idle & lanemask_lt());
                const uint32_t k = q >> 6;
                const uint32_t chunk = bounce == 0 ? (wave + 977u * k) % nw : wave;
                const uint32_t i = k * stride + (chunk << 6) + (q & 63u);
                if (q < seq_len && i < n) {
                    p = bounce == 0 ? i : queue[seg_entry(pre, kSegs, i, qs.seg_cap)];
                    const float4 o = ps.o[p], d = ps.d[p];
                    trav_init(sv, ts, xyz(o), xyz(d), 0.0001f, 100.0f);
                    active = true;
                    my_rays++;
                }
            }
            cursor += nidle;
        }
        if (__ballot(active) == 0ull) {
            if (cursor >= seq_len) break;
            continue;  // positions past the queue's end: take the next ones
        }
        iters++;
        bool to_bin = false;
        uint32_t tri = 0;
        if (!trav_step_wave<kCount>(sv, ts, st, active, sh) && active) {
            active = false;
            my_nodes += ts.nvis;
            my_tris += ts.ntest;
            if (kCount) my_max = max(my_max, ts.nvis);
            const Hit& h = ts.best;
            if (h.tri == kInvalid) {  // miss: the path ends (shader.wgsl:566)
                if (sv.env) ps.L[p] = env_gain(ps, p, bounce, env_radiance(sv, ts.d, ps.wl[p]));
            } else {
                ps.hit[p] = make_uint4(__float_as_uint(h.t), h.tri | (h.front ? 0x80000000u : 0u),
                                       __float_as_uint(h.u), __float_as_uint(h.v));
                tri = h.tri;
                to_bin = true;
            }
        }
        // finished rays wait in the wave's buffer; 64 are binned together,
        // their material lookups made then, once per group
        const uint64_t tb = __ballot(to_bin);
        if (tb) {
            if (to_bin) {
                const uint32_t slot = ndone + (uint32_t)__popcll(tb & lanemask_lt());
                db.p[slot] = p;
                db.tri[slot] = tri;
            }
            ndone += (uint32_t)__popcll(tb);
            if (ndone >= 64u) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();  // other lanes' entries are read below
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                uint32_t key = db.p[lane], val = db.tri[lane];
                wave_sort64(key, val);
                bin_paths(qs, ps, bounce, true, key, tri_bin(sv, val));
                ndone -= 64u;
                if (lane < ndone) {
                    const uint32_t pp = db.p[64 + lane], tt = db.tri[64 + lane];
                    db.p[lane] = pp;
                    db.tri[lane] = tt;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
    // the rest, sorted with empty lanes last
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t key = lane < ndone ? db.p[lane] : 0xFFFFFFFFu, val = lane < ndone ? db.tri[lane] : 0u;
    wave_sort64(key, val);
    bin_paths(qs, ps, bounce, lane < ndone, key, lane < ndone ? tri_bin(sv, val) : 0u);
    if (kCount) {
        // lane slots the wave spent in the traversal loop
        if (lane == 0 && iters) atomicAdd(&ctr->extend_wave_nodes, (unsigned long long)iters * 64ull);
        flush_traversal_counts(ctr, kRkExtend, my_rays, my_nodes, my_tris, my_max);
    }
}

// kCount: the traversal-counting build, a kernel of its own so that its
// counters do not take registers from the plain one.
template <bool kCount>
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(5) void k_extend_refill(
    SceneView sv, PathState ps, Queues qs, Counters* __restrict__ ctr, uint32_t bounce,
    uint32_t n0, const uint32_t* __restrict__ queue, uint2* __restrict__ gstack, uint32_t idle_min) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    __shared__ DoneBuf done_all[kTraceThreads / 64];
    __shared__ TriShare share_all[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t n = bounce == 0 ? n0 : seg_prefix(qs, bounce, kQActive, 1, pre);
    if (qs.tail && bounce >= kTailFirstBounce && (*qs.tail != 0u || n < qs.tail_paths)) {
        if (gtid == 0 && *qs.tail == 0u) *qs.tail = bounce + 1u;  // the tail launch takes these paths
        return;
    }
    if (gtid == 0) {
        atomicAdd(&ctr->extend_rays, (unsigned long long)n);
        atomicAdd(&ctr->kernel_rays[kRkExtend], (unsigned long long)n);
    }
    DoneBuf& db = done_all[threadIdx.x >> 6];
    extend_refill_loop<kCount>(sv, ps, qs, ctr, bounce, n, queue, pre, st, db, share_all[threadIdx.x >> 6],
                               idle_min);
}

__global__ __launch_bounds__(kShadeThreads) void k_shade(SceneView sv, PathState ps, GridView grid,
                                                         FrameArgs fa, uint32_t npx, Queues qs,
                                                         uint32_t bounce, uint32_t guided,
                                                         uint32_t* __restrict__ next_queue) {
    __shared__ uint32_t pre[kMatBins * kSegs + 1];
    if (tail_started(qs, bounce)) return;
    const uint32_t base_frame = *fa.frame;
    const uint32_t rr_after = fa.table[base_frame & fa.mask].rr_after;
    const uint32_t quirks = fa.table[base_frame & fa.mask].quirks;
    const float light_mix = fa.table[base_frame & fa.mask].guide_light_mix;
    const uint32_t photons_on = fa.table[base_frame & fa.mask].enable_photons;
    const uint32_t total = seg_prefix(qs, bounce, kQShade0, kMatBins, pre);
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = gtid & ~63u; base < total; base += stride) {
        const uint32_t i = base + __lane_id();
        const bool have = i < total;
        uint32_t p = 0;
        VertexOut vo;
        vo.alive = vo.shadow = vo.deferred = false;
        if (have) {
            p = ps.shade_q[seg_entry(pre, kMatBins * kSegs, i, qs.seg_cap)];
            const uint4 hr = ps.hit[p];
            PathRegs r;
            if (bounce == 0) {
                // a camera path: recomputed from its pixel and frame, 88
                // bytes of path state a lane does not read back
                camera_regs(fa, base_frame, npx, p, r);
            } else {
                r.o4 = ps.o[p];
                r.d4 = ps.d[p];
                r.tp = ps.tp[p];
                r.L = ps.L[p];
                r.wl = ps.wl[p];
                r.rng = ps.rng[p];
                r.flags = ps.flags[p];
            }
            const uint32_t slot = p / npx;
            const FrameConst* fs = photons_on ? &fa.table[(base_frame + slot) & fa.mask] : nullptr;
            const uint32_t cell_base = photons_on ? slot * fs->hash_size : 0u;
            vo = shade_vertex(sv, grid, fs, cell_base, rr_after, quirks, bounce, guided != 0u,
                              __uint_as_float(hr.x), hr.y & 0x7FFFFFFFu, (hr.y >> 31) != 0u,
                              __uint_as_float(hr.z), __uint_as_float(hr.w), r, light_mix);
            ps.L[p] = r.L;
            if (vo.deferred) {
                ps.galb[p] = vo.galb;
                ps.o[p] = r.o4;
                ps.rng[p] = r.rng;
                ps.flags[p] = r.flags;
                if (bounce == 0) ps.tp[p] = r.tp;  // the guided scatter reads it; k_generate does not store it
            }
            if (vo.alive) {
                ps.o[p] = r.o4;
                ps.d[p] = r.d4;
                ps.tp[p] = r.tp;
                ps.rng[p] = r.rng;
                ps.flags[p] = r.flags;
            }
        }
        const bool shadow = vo.shadow, alive = vo.alive, deferred = vo.deferred;
        const f3 hp = vo.hp, sdir = vo.sdir;
        const float stmax = vo.stmax;
        const float4 contrib = vo.contrib;
        const uint32_t se = seg_append(qs, bounce, kQShadow, qs.seg_cap, shadow);
        if (shadow) {
            float4* r = ps.shadow + 3ull * se;
            r[0] = f4(hp, stmax);
            r[1] = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(p));
            r[2] = contrib;
        }
        if (ps.tag) {
            if (alive) ps.tag[p] = (uint8_t)(bounce + 1);
        } else {
            const uint32_t ne = seg_append(qs, bounce + 1, kQActive, qs.seg_cap, alive);
            if (alive) next_queue[ne] = p;
        }
        const uint32_t ge = seg_append(qs, bounce, kQGuide, qs.seg_cap, deferred);
        if (deferred) {
            ps.guide_q[ge] = p;
            ps.gn[p] = f4(vo.ff, 0.0f);
        }
    }
}

// Active queue of bounce `next` (>= 1) in path order, from the tags the
// previous bounce's shading wrote: path range [s R, (s + 1) R) fills segment
// s (R = the paths over kSegs, rounded up to whole 256-path tiles, so R <=
// seg_cap); a wave takes a tile, each lane 4 consecutive paths, and appends
// the tile's entering paths in order at one wave-aggregated offset.
__global__ __launch_bounds__(256) void k_order_active(PathState ps, Queues qs, uint32_t next, uint32_t n,
                                                      uint32_t* __restrict__ queue) {
    if (tail_started(qs, next - 1u)) return;
    const uint32_t lane = __lane_id();
    const uint32_t R = ((n + kSegs - 1) / kSegs + 255u) & ~255u;
    const uint32_t tiles = (n + 255u) / 256u;
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < tiles; t += nwaves) {
        const uint32_t p0 = t * 256u + lane * 4u;  // tag array padded to whole tiles
        const uint32_t w = *reinterpret_cast<const uint32_t*>(ps.tag + p0);
        uint32_t m = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
            if (p0 + j < n && ((w >> (8 * j)) & 0xFFu) == next) m |= 1u << j;
        const uint32_t c = (uint32_t)__popc(m);
        uint32_t inc = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)inc, off, 64);
            if ((int)lane >= off) inc += v;
        }
        const uint32_t total = (uint32_t)__shfl((int)inc, 63, 64);
        if (total == 0u) continue;
        const uint32_t seg = (t * 256u) / R;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(qcount(qs, next, kQActive, seg), total);
        base = (uint32_t)__shfl((int)base, 0, 64) + inc - c;
        uint32_t* out = queue + (size_t)seg * qs.seg_cap + base;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j)
            if (m & (1u << j)) *out++ = p0 + j;
    }
}

// mark_nee (guided tracer): an unoccluded light sample of bounce b < 4 sets
// the path's kFlagNeeVisible << b, which keeps its contribution out of the
// training target of the vertex.
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(6) void k_shadow(SceneView sv, PathState ps, Queues qs,
                                                          uint32_t count_traversal,
                                                          Counters* __restrict__ ctr, uint32_t bounce,
                                                          uint2* __restrict__ gstack, uint32_t mark_nee) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    if (tail_started(qs, bounce)) return;
    const uint32_t n = seg_prefix(qs, bounce, kQShadow, 1, pre);
    if (gtid == 0) {
        atomicAdd(&ctr->shadow_rays, (unsigned long long)n);
        atomicAdd(&ctr->kernel_rays[kRkShadow], (unsigned long long)n);
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool count = count_traversal != 0u;
    uint32_t my_rays = 0, my_nodes = 0, my_tris = 0, my_max = 0;
    for (uint32_t i = gtid; i < n; i += stride) {
        const float4* r = ps.shadow + 3ull * seg_entry(pre, kSegs, i, qs.seg_cap);
        const float4 a = r[0], b = r[1];
        const uint32_t n_before = my_nodes;
        Hit h = trace_ray<true>(sv, xyz(a), xyz(b), 0.001f, a.w, st, count, my_nodes, my_tris);
        my_max = max(my_max, my_nodes - n_before);
        my_rays++;
        if (h.tri == kInvalid) {
            const uint32_t p = __float_as_uint(b.w);
            ps.L[p] = add4(ps.L[p], r[2]);
            if (mark_nee && bounce < kGuideRecordBounces) ps.flags[p] |= kFlagNeeVisible << bounce;
        }
    }
    if (count) flush_traversal_counts(ctr, kRkShadow, my_rays, my_nodes, my_tris, my_max);
}

// The paths entering bounce t-1 = *qs.tail - 1 (if the wavefront handed
// them over), each carried to its end by one lane with the per-path steps of
// the wavefront kernels (trace_bounces), photon gather included.
// 4 waves per SIMD (128 VGPRs, a few spills) keep 2^18 paths resident at
// once, which measured faster than 2 waves without spills.
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(4) void k_tail(SceneView sv, PathState ps, GridView grid, FrameArgs fa,
                                                         uint32_t npx, Queues qs, uint32_t max_depth,
                                                         uint32_t count_traversal, Counters* __restrict__ ctr,
                                                         uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    const uint32_t ta = *qs.tail;
    if (ta == 0u || ta - 1u >= max_depth) return;
    const uint32_t first = ta - 1u;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t n = seg_prefix(qs, first, kQActive, 1, pre);
    const uint32_t* queue = (first & 1u) ? ps.queue_b : ps.queue_a;
    const uint32_t base_frame = *fa.frame;
    const uint32_t rr_after = fa.table[base_frame & fa.mask].rr_after;
    const uint32_t quirks = fa.table[base_frame & fa.mask].quirks;
    const uint32_t photons_on = fa.table[base_frame & fa.mask].enable_photons;
    const bool count = count_traversal != 0u;
    RayTally tally;
    for (uint32_t i = gtid; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = queue[seg_entry(pre, kSegs, i, qs.seg_cap)];
        PathRegs r;
        r.o4 = ps.o[p];
        r.d4 = ps.d[p];
        r.tp = ps.tp[p];
        r.L = ps.L[p];
        r.wl = ps.wl[p];
        r.rng = ps.rng[p];
        r.flags = ps.flags[p];
        const uint32_t slot = p / npx;
        const FrameConst* fs = photons_on ? &fa.table[(base_frame + slot) & fa.mask] : nullptr;
        const uint32_t cell_base = photons_on ? slot * fs->hash_size : 0u;
        trace_bounces(sv, grid, fs, cell_base, rr_after, quirks, first, max_depth, r, st, count, tally);
        ps.L[p] = r.L;
    }
    // statistics: rays per wave, traversal work while counting
    unsigned long long er = tally.extend_rays, sr = tally.shadow_rays;
    for (int off = 32; off > 0; off >>= 1) {
        er += __shfl_xor(er, off, 64);
        sr += __shfl_xor(sr, off, 64);
    }
    if (__lane_id() == 0) {
        if (er) {
            atomicAdd(&ctr->extend_rays, er);
            atomicAdd(&ctr->kernel_rays[kRkTailExtend], er);
        }
        if (sr) {
            atomicAdd(&ctr->shadow_rays, sr);
            atomicAdd(&ctr->kernel_rays[kRkTailShadow], sr);
        }
    }
    if (count) {
        flush_traversal_counts(ctr, kRkTailExtend, tally.extend_rays, tally.extend_nodes, tally.extend_tris);
        flush_traversal_counts(ctr, kRkTailShadow, tally.shadow_rays, tally.shadow_nodes, tally.shadow_tris);
    }
}

// shader.wgsl:771-822, for the frames of the batch in frame order, so the
// sum per pixel is the one frame-by-frame accumulation produces.
__global__ __launch_bounds__(256) void k_accumulate(PathState ps, FrameArgs fa, uint32_t npx,
                                                    uint32_t nf, float4* __restrict__ accum,
                                                    uint8_t* __restrict__ rgba) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npx) return;
    const uint32_t base_frame = *fa.frame;
    float4 na = accum[pix];
    // eight frames' samples are loaded and converted together (the loads of
    // one frame do not wait for the previous one), then added in frame order
    constexpr uint32_t kGroup = 8;
    for (uint32_t s0 = 0; s0 < nf; s0 += kGroup) {
        float4 col[kGroup];
        bool reset[kGroup];
#pragma unroll
        for (uint32_t k = 0; k < kGroup; ++k) {
            col[k] = splat4(0.0f);
            reset[k] = false;
            if (s0 + k < nf) {
                const FrameConst& fc = fa.table[(base_frame + s0 + k) & fa.mask];
                const uint64_t p = (uint64_t)(s0 + k) * npx + pix;
                col[k] = sample_color(fc, ps.wl[p], ps.L[p]);
                reset[k] = fc.frame == 0u;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kGroup; ++k)
            if (s0 + k < nf) na = add4(reset[k] ? splat4(0.0f) : na, col[k]);
    }
    accum[pix] = na;
    const uchar4 o = tone_map(na);
    reinterpret_cast<uchar4*>(rgba)[pix] = o;
}

}  // namespace

void launch_clear_queues(const Queues& qs, uint32_t depths, hipStream_t s) {
    const uint32_t words = (uint32_t)queue_count_words(depths);
    hipLaunchKernelGGL(k_clear_queues, dim3((words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024), dim3(256),
                       0, s, qs.counts, words, qs.tail);
}
void launch_advance_frame(uint32_t* frame, uint32_t nf, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_frame, dim3(1), dim3(64), 0, s, frame, nf);
}
void launch_set_frame(uint32_t* frame, uint32_t value, hipStream_t s) {
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(64), 0, s, frame, value);
}
void launch_generate(const PathLaunch& pl, uint32_t npx, uint32_t nf, hipStream_t s) {
    uint32_t n = npx * nf;
    hipLaunchKernelGGL(k_generate, dim3((n + 255) / 256), dim3(256), 0, s, pl.ps, pl.fa, npx, nf);
}
void launch_extend(const PathLaunch& pl, uint32_t bounce, uint32_t n0, const uint32_t* queue, hipStream_t s) {
    if (pl.extend_refill)
        hipLaunchKernelGGL(pl.count_traversal ? k_extend_refill<true> : k_extend_refill<false>,
                           dim3(trace_grid(5, pl.qs.trace_rounds)), dim3(kTraceThreads), 0, s, pl.sv, pl.ps, pl.qs,
                           pl.ctr, bounce, n0, queue, pl.gstack, pl.extend_refill);
    else
        hipLaunchKernelGGL(pl.count_traversal ? k_extend<true> : k_extend<false>,
                           dim3(trace_grid(5, pl.qs.trace_rounds)), dim3(kTraceThreads), 0, s, pl.sv, pl.ps, pl.qs,
                           pl.ctr, bounce, n0, queue, pl.gstack);
}
void launch_shade(const PathLaunch& pl, uint32_t npx, uint32_t bounce, uint32_t* next_queue, hipStream_t s) {
    hipLaunchKernelGGL(k_shade, dim3(kShadeBlocks), dim3(kShadeThreads), 0, s, pl.sv, pl.ps, pl.grid,
                       pl.fa, npx, pl.qs, bounce, pl.guided, next_queue);
}
void launch_order_active(const PathLaunch& pl, uint32_t next, uint32_t n, uint32_t* queue, hipStream_t s) {
    hipLaunchKernelGGL(k_order_active, dim3(2048), dim3(256), 0, s, pl.ps, pl.qs, next, n, queue);
}
void launch_shadow(const PathLaunch& pl, uint32_t bounce, hipStream_t s) {
    hipLaunchKernelGGL(k_shadow, dim3(trace_grid(6, pl.qs.trace_rounds)), dim3(kTraceThreads), 0, s, pl.sv, pl.ps, pl.qs,
                       pl.count_traversal, pl.ctr, bounce, pl.gstack, pl.guide_mis);
}
void launch_tail(const PathLaunch& pl, uint32_t npx, uint32_t max_depth, hipStream_t s) {
    hipLaunchKernelGGL(k_tail, dim3(trace_grid(4, pl.qs.trace_rounds)), dim3(kTraceThreads), 0, s, pl.sv, pl.ps, pl.grid, pl.fa, npx,
                       pl.qs, max_depth, pl.count_traversal, pl.ctr, pl.gstack);
}
void launch_accumulate(const PathLaunch& pl, uint32_t npx, uint32_t nf, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3((npx + 255) / 256), dim3(256), 0, s, pl.ps, pl.fa, npx, nf,
                       pl.accum, pl.rgba);
}

}  // namespace wsrt
