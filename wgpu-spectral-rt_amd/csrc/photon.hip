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

// Emission of photon path i (photon_emit.wgsl:159-216): light choice, point
// on the light, cosine direction, hero wavelength and flux; every path's
// stored bit is cleared (its photon slot holds a photon only once the bit
// is set). A photon is stored only after a specular interaction, so a path
// whose first ray meets no metal, glass or thin-film triangle ends at its
// first hit without storing anything. A test against the root of the
// specular BVH finds most of those here (the same cost for every lane), and
// only the others enter the depth-0 queue with their state (in queue
// order). Grid-stride over at most kGridLanes lanes, which bounds each queue
// segment.
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(5) void k_photon_emit(SceneView sv, FrameArgs fa, uint32_t np,
                                                                uint32_t nf, Queues qs, PhotonState ps,
                                                                uint32_t count_traversal,
                                                                Counters* __restrict__ ctr,
                                                                uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ unsigned long long red[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const SceneView spec = specular_view(sv);
    const uint32_t n = np * nf, nl = sv.n_lights;
    const uint32_t base_frame = *fa.frame;
    const FrameConst& f0 = fa.table[base_frame & fa.mask];
    const CauchyConst cc{f0.cdf_min, f0.cdf_range};
    const bool count = count_traversal != 0u;
    uint32_t my_rays = 0, my_nodes = 0, my_tris = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = gtid & ~63u; base < n; base += stride) {
        const uint32_t id = base + __lane_id();
        bool go = false;
        float4 sp, sd;
        uint32_t srng = 0u;
        if (id < n) {
            if ((id & 31u) == 0u) ps.stored[id >> 5] = 0u;  // this lane's word of stored bits
            if (nl != 0u) {
                const uint32_t slot = id / np, local = id - slot * np;
                uint32_t rng = pcg_hash(local + (base_frame + slot) * 927163u);
                uint32_t li = (uint32_t)(rand01(rng) * (float)nl);
                if (li > nl - 1u) li = nl - 1u;
                const wsrt_light_info& L = sv.lights[li];
                const float4 v0 = *reinterpret_cast<const float4*>(L.v0);
                const float4 prm = *reinterpret_cast<const float4*>(L.params);
                f3 emit_pos, emit_dir;
                float area;
                if (prm.x == 1.0f) {  // point light (extension): uniform sphere
                    emit_pos = xyz(v0);
                    emit_dir = random_unit_vector(rng);
                    area = 4.0f;
                } else {
                    const float4 v1 = *reinterpret_cast<const float4*>(L.v1);
                    const float4 v2 = *reinterpret_cast<const float4*>(L.v2);
                    f3 a = xyz(v0), b = xyz(v1), c = xyz(v2);
                    f3 e1 = b - a, e2 = c - a;
                    f3 nrm = normalize(cross(e1, e2));
                    area = length(cross(e1, e2)) * 0.5f;
                    float r1 = rand01(rng), r2 = rand01(rng), sr = sqrtf(r1);
                    float u = 1.0f - sr, v = r2 * sr, w = 1.0f - u - v;
                    emit_pos = a * u + b * v + c * w;
                    emit_dir = normalize(nrm + random_unit_vector(rng));
                }
                const f3 pos = emit_pos + emit_dir * 0.001f;
                // a ray that meets no child box of the specular BVH's root
                // cannot reach a specular triangle; the others are traced in
                // full, and only they need a wavelength and an energy (drawn
                // after the direction, in the reference's order)
                const bool meets = meets_root_child(spec, pos, emit_dir, 0.0001f, 100.0f);
                if (count && spec.n_nodes) my_nodes++;
                my_rays++;
                if (meets) {
                    const float wl = cauchy_sample(rand01(rng), cc);
                    const float pdf = cauchy_pdf(wl, cc);
                    const float radiance = blackbody(wl, prm.y);
                    const float flux = (radiance / pdf) * area * kPi;
                    const float energy = flux * prm.z;
                    go = true;
                    sp = f4(pos, wl);
                    sd = f4(emit_dir, energy);
                    srng = rng;
                }
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
__device__ __forceinline__ void photon_bounce_range(const SceneView& sv, uint32_t depth, const Queues& qs,
                                                    const PhotonState& ps, const uint32_t* pre,
                                                    wsrt_photon* __restrict__ photons, Stack st,
                                                    uint32_t begin, uint32_t end, bool count,
                                                    uint32_t& my_rays, uint32_t& my_stored, uint32_t& my_nodes,
                                                    uint32_t& my_tris) {
    const PhotonQueue& cur = ps.q[depth & 1u];
    const PhotonQueue& nxt = ps.q[(depth + 1u) & 1u];
    for (uint32_t base = begin; base < end; base += 64u) {
        const uint32_t i = base + __lane_id();
        bool alive = false;
        PhotonRay r{};
        f3 pos, dir;
        if (i < end) {
            const uint32_t p = seg_entry(pre, kSegs, i, qs.pseg_cap);
            r.id = cur.id[p];
            const float4 pw = cur.pos[p], de = cur.dir[p];
            r.wl = pw.w;
            r.energy = de.w;
            r.rng = cur.rng[p];
            r.caustic = cur.flags[p] != 0u;
            pos = xyz(pw);
            dir = xyz(de);
            const Hit h = trace_ray<false>(sv, pos, dir, 0.0001f, 100.0f, st, count, my_nodes, my_tris);
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
__global__ __launch_bounds__(kTraceThreads) WSRT_WAVES_PER_SIMD(4) void k_photon_bounce(
    SceneView sv, uint32_t depth, Queues qs, uint32_t count_traversal, PhotonState ps,
    wsrt_photon* __restrict__ photons,
    Counters* __restrict__ ctr, uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ uint32_t pre[kSegs + 1];
    __shared__ unsigned long long red[kTraceThreads / 64];
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
    photon_bounce_range(sv, depth, qs, ps, pre, photons, st, begin, end, count_traversal != 0u,
                        my_rays, my_stored, my_nodes, my_tris);
    if (gtid == 0) atomicAdd(&ctr->kernel_rays[kRkPhotonBounce], (unsigned long long)n);
    if (count_traversal) flush_traversal_counts(ctr, kRkPhotonBounce, my_rays, my_nodes, my_tris);
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
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *n_valid) return;
    const uint32_t slot = sorted_vals[k];
    sorted_slot[k] = slot;
    const float4* src = reinterpret_cast<const float4*>(photons + slot);
    float4* dst = reinterpret_cast<float4*>(sorted + k);
    dst[0] = src[0];
    dst[1] = src[1];
}

}  // namespace

void launch_photon_emit(const SceneView& sv, const FrameArgs& fa, uint32_t np, uint32_t nf, const Queues& qs,
                        const PhotonState& ps, uint32_t count_traversal, Counters* ctr, uint2* gstack,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_photon_emit, dim3(kTraceBlocks), dim3(kTraceThreads), 0, s, sv, fa, np, nf, qs, ps,
                       count_traversal, ctr, gstack);
}

void launch_photon_bounce(const SceneView& sv, uint32_t depth, uint32_t count_traversal, const PhotonState& ps,
                          const Queues& qs, wsrt_photon* photons, Counters* ctr, uint2* gstack, hipStream_t s) {
    hipLaunchKernelGGL(k_photon_bounce, dim3(trace_grid(4)), dim3(kTraceThreads), 0, s, sv, depth, qs,
                       count_traversal, ps, photons, ctr, gstack);
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
    hipLaunchKernelGGL(k_grid_copy, dim3((n + 255) / 256), dim3(256), 0, s, photons, sorted_vals, n_valid,
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
    hipLaunchKernelGGL(k_grid_copy, dim3((n + 255) / 256), dim3(256), 0, s, photons, sorted_vals, n_valid,
                       g.sorted_slot, g.sorted);
    hipLaunchKernelGGL(k_grid_copy_count, dim3(1), dim3(64), 0, s, n_valid, g.n_valid);
    return 0;
}

}  // namespace wsrt
