// photon.hip - the caustic photon pass (photon_emit.wgsl:154-331) and the
// photon grid (build_grid.wgsl:25-73) rebuilt as a counting sort.
//
// Photon path i writes its stored photon to slot i and sets its stored
// bit, so the photon buffer is deterministic; the reference appends with
// atomicAdd in arbitrary order, which holds the same multiset. The grid of a
// dense photon buffer (wsrt_photon_grid_build; energy 0 marks an empty slot):
//   hash    cell_id[i] and an atomic histogram count[cell]
//   scan    offset = exclusive prefix sum of count (wave-level scans)
//   sort    stable radix sort of (cell, slot), so cells list slots ascending
//   copy    photons in cell order, so a gather reads one contiguous range.
// The renderer's grid hashes only the slots whose stored bit is set and
// writes the offsets of non-empty cells from the run starts of the sorted
// keys (the gather reads an offset only when the cell's count is not 0).
#include "kernels.hpp"
#include "queues.hpp"
#include "shading.hpp"

namespace wsrt {

static_assert(kTraceBlocks <= kStatBlocks, "per-block statistics of the photon bounce launch");

namespace {

constexpr int kConeLights = 64;  // lights with an emission cone in LDS
constexpr int kCandCap = 128;     // candidates a wave buffers (< 64 left + 64 new)

// Bounding sphere (centre, radius) of the child boxes of the root of sv's
// BVH. The radius is padded by 1e-3 of the centre's distance from the
// origin plus the radius, far beyond the slab test's own padding.
__device__ float4 root_sphere(const SceneView& sv) {
    const uint4* np = reinterpret_cast<const uint4*>(sv.nodes);
    const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
    const float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    const uint32_t lo_l[3] = {w2.x, w2.z, w3.x}, lo_h[3] = {w2.y, w2.w, w3.y};
    const uint32_t hi_l[3] = {w3.z, w4.x, w4.z}, hi_h[3] = {w3.w, w4.y, w4.w};
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int c = 0; c < 8; ++c) {
        if (byte_of(w1.z, w1.w, c) == 0u) continue;
        for (int a = 0; a < 3; ++a) {
            const float sc = __uint_as_float((uint32_t)((int)(int8_t)((w0.w >> (8 * a)) & 0xFFu) + 127) << 23);
            lo[a] = fminf(lo[a], org[a] + (float)byte_of(lo_l[a], lo_h[a], c) * sc);
            hi[a] = fmaxf(hi[a], org[a] + (float)byte_of(hi_l[a], hi_h[a], c) * sc);
        }
    }
    const f3 c = mk3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
    const float r = 0.5f * length(mk3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]));
    return f4(c, r + 1e-3f * (length(c) + r) + 1e-4f);
}

// Cone (unit axis, cosine of the half-angle) holding the direction u of
// every ray from light L's emission points (pos = point + 0.001 u) that
// meets the sphere sph at t > 0: from a point within r_l of the light's
// centre, a ray meeting the sphere passes within R = r_l + r of the centre
// d away, so its angle to the axis has sine <= R / d. Cosine -2 (no test)
// when the light is too close for a cone.
__device__ float4 light_cone(const wsrt_light_info& L, float4 sph) {
    const float4 v0 = *reinterpret_cast<const float4*>(L.v0);
    const float4 prm = *reinterpret_cast<const float4*>(L.params);
    f3 cl = xyz(v0);
    float rl = 0.0f;
    if (prm.x != 1.0f) {  // triangle: centroid and farthest corner
        const f3 a = xyz(v0), b = xyz(*reinterpret_cast<const float4*>(L.v1));
        const f3 c = xyz(*reinterpret_cast<const float4*>(L.v2));
        cl = (a + b + c) * (1.0f / 3.0f);
        rl = fmaxf(length(a - cl), fmaxf(length(b - cl), length(c - cl)));
    }
    const f3 D = xyz(sph) - cl;
    const float d = length(D);
    const float R = rl * 1.001f + 0.0011f + sph.w + 1e-3f * d;
    if (!(d > 1.01f * R)) return make_float4(0.0f, 0.0f, 1.0f, -2.0f);
    const float s = R / d;
    return f4(D * (1.0f / d), sqrtf(1.0f - s * s) - 1e-4f);
}

// The first ray of photon paths buffered in a wave's LDS (one entry per
// candidate): id, origin, direction, rng state after the direction, light.
struct Cand {
    uint32_t id[kCandCap];
    float px[kCandCap], py[kCandCap], pz[kCandCap];
    float dx[kCandCap], dy[kCandCap], dz[kCandCap];
    uint32_t rng[kCandCap], li[kCandCap];
};

// Root test and the rest of the emission (photon_emit.wgsl:199-216) for
// buffered candidates [0, cnt), one per lane; the paths that may reach
// specular geometry enter the depth-0 queue.
__device__ __forceinline__ void emit_candidates(const SceneView& sv, const SceneView& spec, const Cand& cb,
                                                uint32_t cnt, const CauchyConst& cc, const Queues& qs,
                                                const PhotonState& ps) {
    const uint32_t l = __lane_id();
    bool go = false;
    float4 sp, sd;
    uint32_t srng = 0u, id = 0u;
    if (l < cnt) {
        id = cb.id[l];
        const f3 pos = mk3(cb.px[l], cb.py[l], cb.pz[l]);
        const f3 dir = mk3(cb.dx[l], cb.dy[l], cb.dz[l]);
        if (meets_root_child(spec, pos, dir, 0.0001f, 100.0f)) {
            uint32_t rng = cb.rng[l];
            const wsrt_light_info& L = sv.lights[cb.li[l]];
            const float4 prm = *reinterpret_cast<const float4*>(L.params);
            float area = 4.0f;
            if (prm.x != 1.0f) {
                const f3 a = xyz(*reinterpret_cast<const float4*>(L.v0));
                const f3 e1 = xyz(*reinterpret_cast<const float4*>(L.v1)) - a;
                const f3 e2 = xyz(*reinterpret_cast<const float4*>(L.v2)) - a;
                area = length(cross(e1, e2)) * 0.5f;
            }
            const float wl = cauchy_sample(rand01(rng), cc);
            const float pdf = cauchy_pdf(wl, cc);
            const float radiance = blackbody(wl, prm.y);
            const float flux = (radiance / pdf) * area * kPi;
            const float energy = flux * prm.z;
            go = true;
            sp = f4(pos, wl);
            sd = f4(dir, energy);
            srng = rng;
        }
    }
    const uint32_t e = seg_append(qs, 0, kQPhoton, qs.pseg_cap, go);
    if (go) {
        const PhotonQueue& q = ps.q[0];
        q.id[e] = id;
        q.pos[e] = sp;
        q.dir[e] = sd;
        q.rng[e] = srng;
        q.flags[e] = 0u;
    }
}

// Emission of photon path i (photon_emit.wgsl:159-216): light choice, point
// on the light, cosine direction, hero wavelength and flux; every path's
// stored bit is cleared (its photon slot holds a photon only once the bit
// is set). A photon is stored only after a specular interaction, so a path
// whose first ray meets no metal, glass or thin-film triangle ends at its
// first hit without storing anything. Such rays are found in two steps:
// a cone per light around the specular BVH's root sphere rejects most of
// them with one dot product, and the others (buffered per wave in LDS until
// 64 fill a wave) take the slab test against the root's child boxes. Only
// the paths that pass it enter the depth-0 queue, with their state in queue
// order; a ray outside the cone meets no child box, so the queue holds the
// same paths as with the slab test alone. Grid-stride over at most
// kGridLanes lanes, and a wave queues at most as many paths as it emits,
// which bounds each queue segment.
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(5) void k_photon_emit(SceneView sv, FrameArgs fa, uint32_t np,
                                                                uint32_t nf, Queues qs, PhotonState ps,
                                                                uint32_t count_traversal,
                                                                Counters* __restrict__ ctr) {
    __shared__ unsigned long long red[kTraceThreads / 64];
    __shared__ float4 cone[kConeLights];
    __shared__ Cand cand_all[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const SceneView spec = specular_view(sv);
    const uint32_t n = np * nf, nl = sv.n_lights;
    const bool use_cone = spec.n_nodes != 0u && nl <= (uint32_t)kConeLights;
    if (use_cone) {
        const float4 sph = root_sphere(spec);
        for (uint32_t k = threadIdx.x; k < nl; k += blockDim.x) cone[k] = light_cone(sv.lights[k], sph);
    }
    __syncthreads();
    Cand& cb = cand_all[threadIdx.x >> 6];
    const uint32_t base_frame = *fa.frame;
    const FrameConst& f0 = fa.table[base_frame & fa.mask];
    const CauchyConst cc{f0.cdf_min, f0.cdf_range};
    const bool count = count_traversal != 0u;
    uint32_t my_rays = 0, my_nodes = 0, my_tris = 0;
    uint32_t ncand = 0;  // wave-uniform
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = gtid & ~63u; base < n; base += stride) {
        const uint32_t id = base + __lane_id();
        bool c = false;
        f3 pos, emit_dir;
        uint32_t rng = 0u, li = 0u;
        if (id < n) {
            if ((id & 31u) == 0u) ps.stored[id >> 5] = 0u;  // this lane's word of stored bits
            if (nl != 0u && spec.n_nodes != 0u) {
                const uint32_t slot = id / np, local = id - slot * np;
                rng = pcg_hash(local + (base_frame + slot) * 927163u);
                li = (uint32_t)(rand01(rng) * (float)nl);
                if (li > nl - 1u) li = nl - 1u;
                const wsrt_light_info& L = sv.lights[li];
                const float4 v0 = *reinterpret_cast<const float4*>(L.v0);
                const float4 prm = *reinterpret_cast<const float4*>(L.params);
                f3 emit_pos;
                if (prm.x == 1.0f) {  // point light (extension): uniform sphere
                    emit_pos = xyz(v0);
                    emit_dir = random_unit_vector(rng);
                } else {
                    const float4 v1 = *reinterpret_cast<const float4*>(L.v1);
                    const float4 v2 = *reinterpret_cast<const float4*>(L.v2);
                    f3 a = xyz(v0), b = xyz(v1), cc3 = xyz(v2);
                    f3 e1 = b - a, e2 = cc3 - a;
                    f3 nrm = normalize(cross(e1, e2));
                    float r1 = rand01(rng), r2 = rand01(rng), sr = sqrtf(r1);
                    float u = 1.0f - sr, v = r2 * sr, w = 1.0f - u - v;
                    emit_pos = a * u + b * v + cc3 * w;
                    emit_dir = normalize(nrm + random_unit_vector(rng));
                }
                pos = emit_pos + emit_dir * 0.001f;
                c = true;
                if (use_cone) {
                    const float4 k = cone[li];
                    c = dot(emit_dir, xyz(k)) >= k.w;
                }
            }
            if (nl != 0u) {
                // the root node is read only by the slab test of a candidate
                if (count && c) my_nodes++;
                my_rays++;
            }
        }
        // append this wave's candidates to its buffer; test 64 at a time
        const uint64_t m = __ballot(c);
        if (c) {
            const uint32_t k = ncand + (uint32_t)__popcll(m & lanemask_lt());
            cb.id[k] = id;
            cb.px[k] = pos.x;
            cb.py[k] = pos.y;
            cb.pz[k] = pos.z;
            cb.dx[k] = emit_dir.x;
            cb.dy[k] = emit_dir.y;
            cb.dz[k] = emit_dir.z;
            cb.rng[k] = rng;
            cb.li[k] = li;
        }
        ncand += (uint32_t)__popcll(m);
        if (ncand >= 64u) {
            emit_candidates(sv, spec, cb, 64u, cc, qs, ps);
            ncand -= 64u;
            // move the rest (< 64) to the front
            const uint32_t l = __lane_id();
            if (l < ncand) {
                const uint32_t a = cb.id[64 + l], r = cb.rng[64 + l], li2 = cb.li[64 + l];
                const float x = cb.px[64 + l], y = cb.py[64 + l], z = cb.pz[64 + l];
                const float dx = cb.dx[64 + l], dy = cb.dy[64 + l], dz = cb.dz[64 + l];
                cb.id[l] = a;
                cb.rng[l] = r;
                cb.li[l] = li2;
                cb.px[l] = x;
                cb.py[l] = y;
                cb.pz[l] = z;
                cb.dx[l] = dx;
                cb.dy[l] = dy;
                cb.dz[l] = dz;
            }
        }
    }
    if (ncand) emit_candidates(sv, spec, cb, ncand, cc, qs, ps);
    if (gtid == 0 && nl != 0u) atomicAdd(&ctr->kernel_rays[kRkPhotonEmit], (unsigned long long)n);
    if (count) flush_traversal_counts(ctr, kRkPhotonEmit, my_rays, my_nodes, my_tris);
    // every photon path's first ray
    const unsigned long long br = block_sum(my_rays, red);
    if (threadIdx.x == 0) ctr->photon_rays_part[blockIdx.x] += br;
}

// Register state of a photon path between its load and the end of its ray.
struct PhotonRay {
    uint32_t id;  // photon path id, also its state slot
    float wl, energy;
    uint32_t rng;
    bool caustic;
};

// The surface interaction of photon_emit.wgsl:228-329 at the closest hit h
// of the ray (pos, dir): store at a diffuse surface if the path is caustic,
// else continue through metal, glass and thin film. Returns whether the
// path continues; pos, dir, r.rng and r.caustic are then its next ray's.
__device__ __forceinline__ bool photon_interact(const SceneView& sv, const PhotonState& ps,
                                                wsrt_photon* __restrict__ photons, PhotonRay& r,
                                                f3& pos, f3& dir, const Hit& h, uint32_t& my_stored) {
    if (h.tri == kInvalid) return false;
    bool alive = false;
    bool caustic = r.caustic;
    uint32_t rng = r.rng;
    const float wl = r.wl;
    uint32_t mat_id;
    f3 wn = hit_world_normal(sv, h.tri, h.u, h.v, &mat_id);
    f3 ff = h.front ? wn : -wn;
    const wsrt_material& m = sv.materials[mat_id];
    const float4 ex = *reinterpret_cast<const float4*>(m.extra);
    const uint32_t type = (uint32_t)ex.x;
    f3 hp = pos + dir * h.t;
    if (type == 0u) {
        if (caustic) {
            float4* dst = reinterpret_cast<float4*>(photons + r.id);
            dst[0] = make_float4(hp.x, hp.y, hp.z, wl);
            dst[1] = make_float4(-dir.x, -dir.y, -dir.z, r.energy);
            // a photon of energy 0 counts as skipped, as in the dense photon
            // buffer that wsrt_photon_grid_build takes
            const bool keep = r.energy != 0.0f;
            if (keep) atomicOr(&ps.stored[r.id >> 5], 1u << (r.id & 31u));
            my_stored += keep ? 1u : 0u;
        }
    } else if (type == 1u) {
        dir = reflect(normalize(dir), ff);
        pos = hp + ff * 0.001f;
        alive = true;
    } else if (type == 2u) {
        float lum = wl / 1000.0f;
        float ir = ex.z + (ex.w / (lum * lum));
        float ratio = h.front ? 1.0f / ir : ir;
        f3 ud = normalize(dir);
        float ct = fminf(dot(-ud, ff), 1.0f);
        float sn = sqrtf(1.0f - ct * ct);
        float r0 = (1.0f - ratio) / (1.0f + ratio);
        r0 = r0 * r0;
        float refl = r0 + (1.0f - r0) * powf(1.0f - ct, 5.0f);
        // short-circuit: total internal reflection draws no random number
        if (ratio * sn > 1.0f || refl > rand01(rng)) {
            dir = reflect(ud, ff);
        } else {
            caustic = true;
            dir = refract(ud, ff, ratio);
        }
        f3 off = dot(dir, ff) > 0.0f ? ff : -ff;
        pos = hp + off * 0.001f;
        alive = true;
    } else if (type == 4u) {
        float n2 = ex.z, n3 = ex.w, d = fmaxf(ex.y, 50.0f);
        f3 ud = normalize(dir);
        float ct = fminf(dot(-ud, ff), 1.0f);
        float ratio = h.front ? n3 / 1.0f : 1.0f / n3;
        float stsq = ratio * ratio * (1.0f - ct * ct);
        if (stsq > 1.0f) {
            dir = reflect(ud, ff);
        } else {
            float F = thin_film(ct, 1.0f, n2, n3, d, wl);
            if (rand01(rng) < F) dir = reflect(ud, ff);
            else dir = refract(ud, ff, ratio);
        }
        caustic = true;
        alive = true;
        // the origin is not advanced in this branch (photon_emit.wgsl:291-326)
    }
    r.rng = rng;
    r.caustic = caustic;
    return alive;
}

// The rays of entries [begin, end) of a bounce, one per lane at a time.
template <bool kCount>
__device__ __forceinline__ void photon_bounce_range(const SceneView& sv, uint32_t depth, const Queues& qs,
                                                    const PhotonState& ps, const uint32_t* pre,
                                                    wsrt_photon* __restrict__ photons, Stack st, TriShare& sh,
                                                    uint32_t begin, uint32_t end,
                                                    uint32_t& my_rays, uint32_t& my_stored, uint32_t& my_nodes,
                                                    uint32_t& my_tris) {
    const PhotonQueue& cur = ps.q[depth & 1u];
    const PhotonQueue& nxt = ps.q[(depth + 1u) & 1u];
    for (uint32_t base = begin; base < end; base += 64u) {
        const uint32_t i = base + __lane_id();
        const bool have = i < end;
        bool alive = false;
        PhotonRay r{};
        f3 pos = mk3(0.0f, 0.0f, 0.0f), dir = mk3(0.0f, 0.0f, 1.0f);
        const uint32_t s0 = seg_of_wave_base(pre, base);  // whole wave
        if (have) {
            const uint32_t p = seg_entry_from(pre, s0, i, qs.pseg_cap);
            r.id = cur.id[p];
            const float4 pw = cur.pos[p], de = cur.dir[p];
            r.wl = pw.w;
            r.energy = de.w;
            r.rng = cur.rng[p];
            r.caustic = cur.flags[p] != 0u;
            pos = xyz(pw);
            dir = xyz(de);
        }
        // the wave traverses its rays together, triangle tests shared
        const Hit h = trace_wave<kCount>(sv, pos, dir, 0.0001f, 100.0f, have, st, sh, my_nodes, my_tris);
        if (have) {
            my_rays++;
            alive = photon_interact(sv, ps, photons, r, pos, dir, h, my_stored);
        }
        const uint32_t e = seg_append(qs, depth + 1, kQPhoton, qs.pseg_cap, alive);
        if (alive) {
            nxt.id[e] = r.id;
            nxt.pos[e] = f4(pos, r.wl);
            nxt.dir[e] = f4(dir, r.energy);
            nxt.rng[e] = r.rng;
            nxt.flags[e] = r.caustic ? 1u : 0u;
        }
    }
}

// One bounce of every queued photon path (photon_emit.wgsl:219-330): closest
// hit, then the surface interaction. Survivors are compacted into the next
// queue, so the few paths refracting through glass run in full waves. Each
// wave takes a contiguous range of the queue.
// kCount: the traversal-counting build, a kernel of its own so that its
// counters do not take registers from the plain one.
template <bool kCount>
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(5) void k_photon_bounce(
    SceneView sv, uint32_t depth, Queues qs, PhotonState ps,
    wsrt_photon* __restrict__ photons,
    Counters* __restrict__ ctr, uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    __shared__ unsigned long long red[kTraceThreads / 64];
    __shared__ TriShare share_all[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t n = seg_prefix(qs, depth, kQPhoton, 1, pre);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t nchunks = (n + 63u) >> 6;
    const uint32_t per_wave = (nchunks + nwaves - 1u) / nwaves;
    const uint32_t wave = gtid >> 6;
    const uint32_t begin = (uint32_t)min((uint64_t)n, (uint64_t)wave * per_wave * 64u);
    const uint32_t end = (uint32_t)min((uint64_t)n, (uint64_t)(wave + 1u) * per_wave * 64u);
    uint32_t my_rays = 0, my_stored = 0, my_nodes = 0, my_tris = 0;
    photon_bounce_range<kCount>(sv, depth, qs, ps, pre, photons, st, share_all[threadIdx.x >> 6], begin, end,
                                my_rays, my_stored, my_nodes, my_tris);
    if (gtid == 0) atomicAdd(&ctr->kernel_rays[kRkPhotonBounce], (unsigned long long)n);
    if (kCount) flush_traversal_counts(ctr, kRkPhotonBounce, my_rays, my_nodes, my_tris);
    // per-block statistics, summed on the host; the emit kernel counted the
    // first rays
    const unsigned long long br = block_sum(depth != 0 ? my_rays : 0u, red);
    const unsigned long long bs = block_sum(my_stored, red);
    if (threadIdx.x == 0) {
        ctr->photon_rays_part[blockIdx.x] += br;
        ctr->photons_stored_part[blockIdx.x] += bs;
    }
}

// Cell of every photon slot and the atomic histogram. Grid cell = frame
// slot * hash_size + hash; skipped photons (energy 0) get no cell.
__global__ __launch_bounds__(256) void k_grid_hash(const wsrt_photon* __restrict__ photons,
                                                   uint32_t n, uint32_t np, FrameArgs fa,
                                                   uint32_t* __restrict__ cell_id,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ valid_flag,
                                                   uint32_t* __restrict__ count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = i / np;
    const FrameConst& fc = fa.table[(*fa.frame + slot) & fa.mask];
    const float4 a = reinterpret_cast<const float4*>(photons + i)[0];
    const float e = reinterpret_cast<const float4*>(photons + i)[1].w;
    uint32_t c = kInvalid, key = 0, valid = 0;
    if (e != 0.0f) {
        c = hash_cell(xyz(a), mk3(fc.jitter[0], fc.jitter[1], fc.jitter[2]), fc.cell_size, fc.hash_size);
        key = slot * fc.hash_size + c;
        atomicAdd(&count[key], 1u);
        valid = 1;
    }
    cell_id[i] = c;
    keys[i] = key;
    valid_flag[i] = valid;
}

// Stream compaction of the valid photons in slot order (positions from an
// exclusive scan of the flags), so the sort sees only stored photons.
__global__ __launch_bounds__(256) void k_grid_compact(uint32_t n, const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ valid_flag,
                                                      const uint32_t* __restrict__ pos,
                                                      uint32_t* __restrict__ ckeys,
                                                      uint32_t* __restrict__ cvals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !valid_flag[i]) return;
    ckeys[pos[i]] = keys[i];
    cvals[pos[i]] = i;
}

// Stored photons per 32-slot word of the stored bits.
__global__ __launch_bounds__(256) void k_word_popc(const uint32_t* __restrict__ bits, uint32_t words,
                                                   uint32_t* __restrict__ cnt) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < words) cnt[w] = (uint32_t)__popc(bits[w]);
}

// Cell and histogram of the stored photons only, compacted in slot order:
// one lane per 32-slot word, positions from the exclusive scan of the words'
// counts plus the rank of the bit within its word.
__global__ __launch_bounds__(256) void k_grid_hash_stored(const wsrt_photon* __restrict__ photons,
                                                          const uint32_t* __restrict__ bits,
                                                          const uint32_t* __restrict__ wpos, uint32_t words,
                                                          uint32_t np, FrameArgs fa,
                                                          uint32_t* __restrict__ cell_id,
                                                          uint32_t* __restrict__ count,
                                                          uint32_t* __restrict__ ckeys,
                                                          uint32_t* __restrict__ cvals) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t m = bits[w];
    uint32_t k = m ? wpos[w] : 0u;
    while (m) {
        const uint32_t i = w * 32u + (uint32_t)(__ffs(m) - 1);
        m &= m - 1u;
        const uint32_t slot = i / np;
        const FrameConst& fc = fa.table[(*fa.frame + slot) & fa.mask];
        const float4 a = reinterpret_cast<const float4*>(photons + i)[0];
        const uint32_t c =
            hash_cell(xyz(a), mk3(fc.jitter[0], fc.jitter[1], fc.jitter[2]), fc.cell_size, fc.hash_size);
        const uint32_t key = slot * fc.hash_size + c;
        atomicAdd(&count[key], 1u);
        cell_id[i] = c;
        ckeys[k] = key;
        cvals[k] = i;
        k++;
    }
}

// Offset of every non-empty cell: the position of the first of its photons
// in the sorted order (keys ascending), and each photon's cell kept for the
// next build's clear.
__global__ __launch_bounds__(256) void k_grid_run_offsets(const uint32_t* __restrict__ skeys,
                                                          const uint32_t* __restrict__ n_valid,
                                                          uint32_t* __restrict__ offset,
                                                          uint32_t* __restrict__ touched) {
    const uint32_t nv = *n_valid;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nv; k += gridDim.x * blockDim.x) {
        const uint32_t key = skeys[k];
        if (k == 0 || skeys[k - 1] != key) offset[key] = k;
        touched[k] = key;
    }
}

// Zeroes the counts the last build left non-zero (the rest are zero).
__global__ __launch_bounds__(256) void k_grid_clear_touched(const uint32_t* __restrict__ touched,
                                                            const uint32_t* __restrict__ n_last,
                                                            uint32_t* __restrict__ count) {
    const uint32_t nl = *n_last;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x)
        count[touched[k]] = 0u;
}

__global__ void k_grid_copy_count(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
}

__global__ __launch_bounds__(256) void k_grid_copy(const wsrt_photon* __restrict__ photons,
                                                   const uint32_t* __restrict__ sorted_vals,
                                                   const uint32_t* __restrict__ n_valid,
                                                   uint32_t* __restrict__ sorted_slot,
                                                   wsrt_photon* __restrict__ sorted) {
    // grid-stride over the valid photons only (their count is on the device)
    const uint32_t nv = *n_valid;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nv; k += gridDim.x * blockDim.x) {
        const uint32_t slot = sorted_vals[k];
        sorted_slot[k] = slot;
        if (sorted) {  // null: slots only
            const float4* src = reinterpret_cast<const float4*>(photons + slot);
            float4* dst = reinterpret_cast<float4*>(sorted + k);
            dst[0] = src[0];
            dst[1] = src[1];
        }
    }
}
// One frame's constants into a table entry, and the device frame counter
// set to 0 (so entry 0 is the frame's).
__global__ void k_store_frame(FrameConst fc, FrameConst* __restrict__ dst, uint32_t* __restrict__ frame) {
    if (threadIdx.x == 0) {
        *dst = fc;
        *frame = 0u;
    }
}
// p[k] = v for *start <= k < n.
__global__ __launch_bounds__(256) void k_fill_after(uint32_t* __restrict__ p, const uint32_t* __restrict__ start,
                                                    uint32_t n, uint32_t v) {
    for (uint32_t k = *start + blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = v;
}
// blocks of a launch over at most n valid photons: enough to fill the chip,
// no more than the photons need
inline uint32_t valid_grid(uint32_t n) { return std::max(1u, std::min((n + 255u) / 256u, 2048u)); }

}  // namespace

void launch_photon_emit(const SceneView& sv, const FrameArgs& fa, uint32_t np, uint32_t nf, const Queues& qs,
                        const PhotonState& ps, uint32_t count_traversal, Counters* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_photon_emit, dim3(kTraceBlocks), dim3(kTraceThreads), 0, s, sv, fa, np, nf, qs, ps,
                       count_traversal, ctr);
}

void launch_photon_bounce(const SceneView& sv, uint32_t depth, uint32_t count_traversal, const PhotonState& ps,
                          const Queues& qs, wsrt_photon* photons, Counters* ctr, uint2* gstack, hipStream_t s) {
    hipLaunchKernelGGL(count_traversal ? k_photon_bounce<true> : k_photon_bounce<false>,
                       dim3(trace_grid(5, qs.trace_rounds)), dim3(kTraceThreads), 0, s, sv, depth, qs, ps, photons,
                       ctr, gstack);
}

void launch_store_frame(const FrameConst& fc, FrameConst* dst, uint32_t* frame, hipStream_t s) {
    hipLaunchKernelGGL(k_store_frame, dim3(1), dim3(64), 0, s, fc, dst, frame);
}

void launch_fill_after(uint32_t* p, const uint32_t* start, uint32_t n, uint32_t v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_after, dim3(valid_grid(n)), dim3(256), 0, s, p, start, n, v);
}

size_t grid_temp_words(uint32_t n, uint32_t cells) {
    size_t sort = radix_temp_words(n);
    size_t scan = scan_temp_words(cells > n ? cells : n);
    return 5 * (size_t)n + (sort > scan ? sort : scan) + 4;
}

int grid_build(const wsrt_photon* photons, uint32_t np, uint32_t nf, uint32_t hash_size,
               const FrameArgs& fa, GridBuffers& g, hipStream_t s) {
    // g.temp: keys[n] flags[n] pos[n] ckeys[n] cvals[n] | sort/scan work | n_valid
    const uint32_t n = np * nf, cells = hash_size * nf;
    uint32_t* keys = g.temp;
    uint32_t* flags = keys + n;
    uint32_t* pos = flags + n;
    uint32_t* ckeys = pos + n;
    uint32_t* cvals = ckeys + n;
    uint32_t* work = cvals + n;
    uint32_t* n_valid = g.temp + grid_temp_words(n, cells) - 1;
    fill_u32(g.count, cells, 0u, s);
    fill_u32(n_valid, 1, 0u, s);
    if (n == 0) {
        if (scan_exclusive(g.count, g.offset, cells, work, nullptr, s)) return -1;
        return 0;
    }
    hipLaunchKernelGGL(k_grid_hash, dim3((n + 255) / 256), dim3(256), 0, s, photons, n, np, fa, g.cell_id,
                       keys, flags, g.count);
    if (scan_exclusive(flags, pos, n, work, n_valid, s)) return -1;
    hipLaunchKernelGGL(k_grid_compact, dim3((n + 255) / 256), dim3(256), 0, s, n, keys, flags, pos, ckeys,
                       cvals);
    if (scan_exclusive(g.count, g.offset, cells, work, nullptr, s)) return -1;
    int bits = 1;
    while ((1ull << bits) < (uint64_t)cells) bits++;  // keys in [0, cells)
    bool in_tmp = false;
    if (radix_sort_pairs(ckeys, cvals, n, bits, work, s, &in_tmp, n_valid)) return -1;
    const uint32_t* sorted_vals = in_tmp ? work + n : cvals;
    hipLaunchKernelGGL(k_grid_copy, dim3(valid_grid(n)), dim3(256), 0, s, photons, sorted_vals, n_valid,
                       g.sorted_slot, g.sorted);
    hipLaunchKernelGGL(k_grid_copy_count, dim3(1), dim3(64), 0, s, n_valid, g.n_valid);
    return 0;
}

int grid_build_stored(const wsrt_photon* photons, const uint32_t* stored, uint32_t np, uint32_t nf,
                      uint32_t hash_size, const FrameArgs& fa, GridBuffers& g, hipStream_t s) {
    // g.temp: as grid_build (keys and flags unused)
    const uint32_t n = np * nf, cells = hash_size * nf;
    uint32_t* pos = g.temp + 2ull * n;
    uint32_t* ckeys = pos + n;
    uint32_t* cvals = ckeys + n;
    uint32_t* work = cvals + n;
    uint32_t* n_valid = g.temp + grid_temp_words(n, cells) - 1;
    // the counts are zero except the cells of the last build's photons
    hipLaunchKernelGGL(k_grid_clear_touched, dim3(1024), dim3(256), 0, s, g.touched, g.n_valid, g.count);
    fill_u32(n_valid, 1, 0u, s);
    if (n == 0) {
        fill_u32(g.n_valid, 1, 0u, s);
        return 0;
    }
    // pos: count of stored photons per 32-slot word, then their exclusive scan
    const uint32_t words = (n + 31u) / 32u;
    uint32_t* wcnt = pos;
    uint32_t* wpos = pos + words;
    hipLaunchKernelGGL(k_word_popc, dim3((words + 255) / 256), dim3(256), 0, s, stored, words, wcnt);
    if (scan_exclusive(wcnt, wpos, words, work, n_valid, s)) return -1;
    hipLaunchKernelGGL(k_grid_hash_stored, dim3((words + 255) / 256), dim3(256), 0, s, photons, stored, wpos, words,
                       np, fa, g.cell_id, g.count, ckeys, cvals);
    int bits = 1;
    while ((1ull << bits) < (uint64_t)cells) bits++;  // keys in [0, cells)
    bool in_tmp = false;
    if (radix_sort_pairs(ckeys, cvals, n, bits, work, s, &in_tmp, n_valid)) return -1;
    const uint32_t* sorted_keys = in_tmp ? work : ckeys;
    const uint32_t* sorted_vals = in_tmp ? work + n : cvals;
    hipLaunchKernelGGL(k_grid_run_offsets, dim3(1024), dim3(256), 0, s, sorted_keys, n_valid, g.offset, g.touched);
    hipLaunchKernelGGL(k_grid_copy, dim3(valid_grid(n)), dim3(256), 0, s, photons, sorted_vals, n_valid,
                       g.sorted_slot, g.sorted);
    hipLaunchKernelGGL(k_grid_copy_count, dim3(1), dim3(64), 0, s, n_valid, g.n_valid);
    return 0;
}

}  // namespace wsrt
