// photon.hip - the caustic photon pass (photon_emit.wgsl:154-331) and the
// photon grid (build_grid.wgsl:25-73) rebuilt as a counting sort.
//
// Photon path i writes its stored photon (or energy 0) to slot i, so the
// photon buffer is deterministic; the reference appends with atomicAdd in
// arbitrary order, which holds the same multiset. The grid is then:
//   hash    cell_id[i] and an atomic histogram count[cell]
//   scan    offset = exclusive prefix sum of count (wave-level scans)
//   sort    stable radix sort of (cell, slot), so cells list slots ascending
//   copy    photons in cell order, so a gather reads one contiguous range.
#include "kernels.hpp"
#include "shading.hpp"

namespace wsrt {

namespace {

__global__ __launch_bounds__(kTraceThreads) void k_photon_trace(SceneView sv, FrameArgs fa,
                                                                uint32_t np, uint32_t nf,
                                                                uint32_t count_traversal,
                                                                wsrt_photon* __restrict__ photons,
                                                                Counters* __restrict__ ctr,
                                                                uint2* __restrict__ gstack) {
    const uint32_t base_frame = *fa.frame;
    const FrameConst& f0 = fa.table[base_frame & fa.mask];
    const uint32_t photon_max_depth = f0.photon_max_depth;
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t stride = gridDim.x * blockDim.x;
    const CauchyConst cc{f0.cdf_min, f0.cdf_range};
    const uint32_t n_photons = np * nf;
    uint32_t my_rays = 0, my_stored = 0, my_nodes = 0, my_tris = 0;
    const bool count = count_traversal != 0u;
    for (uint32_t id = gtid; id < n_photons; id += stride) {
        float4 out0 = splat4(0.0f), out1 = splat4(0.0f);
        const uint32_t slot = id / np, local = id - slot * np;
        uint32_t rng = pcg_hash(local + (base_frame + slot) * 927163u);
        const uint32_t nl = sv.n_lights;
        if (nl != 0u) {
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
            float wl = cauchy_sample(rand01(rng), cc);
            float pdf = cauchy_pdf(wl, cc);
            float radiance = blackbody(wl, prm.y);
            float flux = (radiance / pdf) * area * kPi;
            float energy = flux * prm.z;
            f3 pos = emit_pos + emit_dir * 0.001f;
            f3 dir = emit_dir;
            bool caustic = false;
            for (uint32_t depth = 0; depth < photon_max_depth; ++depth) {
                Hit h = trace_ray<false>(sv, pos, dir, 0.0001f, 100.0f, st, count, my_nodes, my_tris);
                my_rays++;
                if (h.tri == kInvalid) break;
                uint32_t mat_id;
                f3 wn = hit_world_normal(sv, h.tri, h.u, h.v, &mat_id);
                f3 ff = h.front ? wn : -wn;
                const wsrt_material& m = sv.materials[mat_id];
                const float4 ex = *reinterpret_cast<const float4*>(m.extra);
                const uint32_t type = (uint32_t)ex.x;
                f3 hp = pos + dir * h.t;
                if (type == 0u) {
                    if (caustic) {
                        out0 = make_float4(hp.x, hp.y, hp.z, wl);
                        out1 = make_float4(-dir.x, -dir.y, -dir.z, energy);
                        my_stored++;
                    }
                    break;
                } else if (type == 1u) {
                    dir = reflect(normalize(dir), ff);
                    pos = hp + ff * 0.001f;
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
                    // the origin is not advanced in this branch (photon_emit.wgsl:291-326)
                } else {
                    break;
                }
            }
        }
        float4* dst = reinterpret_cast<float4*>(photons + id);
        dst[0] = out0;
        dst[1] = out1;
    }
    // wave-aggregated statistics
    if (count) flush_traversal_counts(ctr, 0, my_rays, my_nodes, my_tris);
    for (int off = 32; off > 0; off >>= 1) {
        my_rays += __shfl_xor(my_rays, off, 64);
        my_stored += __shfl_xor(my_stored, off, 64);
    }
    if (__lane_id() == 0) {
        atomicAdd(&ctr->photon_rays, (unsigned long long)my_rays);
        atomicAdd(&ctr->photons_stored, (unsigned long long)my_stored);
        atomicAdd(&ctr->stored, my_stored);
    }
}

// cell id per slot (hash_size marks skipped photons so they sort last) and
// the atomic histogram.
__global__ __launch_bounds__(256) void k_grid_hash(const wsrt_photon* __restrict__ photons,
                                                   uint32_t n, uint32_t np, uint32_t nf, FrameArgs fa,
                                                   uint32_t* __restrict__ cell_id,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ count,
                                                   uint32_t* __restrict__ n_valid) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    if (i < n) {
        const uint32_t slot = i / np;
        const FrameConst& fc = fa.table[(*fa.frame + slot) & fa.mask];
        const float4 a = reinterpret_cast<const float4*>(photons + i)[0];
        const float e = reinterpret_cast<const float4*>(photons + i)[1].w;
        uint32_t c = kInvalid, key = nf * fc.hash_size;
        if (e != 0.0f) {
            c = hash_cell(xyz(a), mk3(fc.jitter[0], fc.jitter[1], fc.jitter[2]), fc.cell_size,
                          fc.hash_size);
            key = slot * fc.hash_size + c;
            atomicAdd(&count[key], 1u);
            valid = true;
        }
        cell_id[i] = c;
        keys[i] = key;
        vals[i] = i;
    }
    uint64_t m = __ballot(valid);
    if (__lane_id() == 0 && m) atomicAdd(n_valid, (uint32_t)__popcll(m));
}

__global__ __launch_bounds__(256) void k_grid_copy(const wsrt_photon* __restrict__ photons,
                                                   const uint32_t* __restrict__ sorted_slot,
                                                   const uint32_t* __restrict__ n_valid,
                                                   wsrt_photon* __restrict__ sorted) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *n_valid) return;
    const float4* src = reinterpret_cast<const float4*>(photons + sorted_slot[k]);
    float4* dst = reinterpret_cast<float4*>(sorted + k);
    dst[0] = src[0];
    dst[1] = src[1];
}

}  // namespace

void launch_photon_trace(const SceneView& sv, const FrameArgs& fa, uint32_t np, uint32_t nf,
                         uint32_t count_traversal, wsrt_photon* photons, Counters* ctr, uint2* gstack,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_photon_trace, dim3(kTraceBlocks), dim3(kTraceThreads), 0, s, sv, fa, np, nf,
                       count_traversal, photons, ctr, gstack);
}

size_t grid_temp_words(uint32_t n, uint32_t cells) {
    size_t sort = radix_temp_words(n);
    size_t scan = scan_temp_words(cells);
    return 2 * (size_t)n + (sort > scan ? sort : scan) + 4;
}

int grid_build(const wsrt_photon* photons, uint32_t np, uint32_t nf, uint32_t hash_size,
               const FrameArgs& fa, GridBuffers& g, hipStream_t s) {
    // g.temp layout: keys[n], vals[n], sort/scan temp, n_valid
    const uint32_t n = np * nf, cells = hash_size * nf;
    uint32_t* keys = g.temp;
    uint32_t* vals = g.temp + n;
    uint32_t* work = g.temp + 2 * (size_t)n;
    uint32_t* n_valid = g.temp + grid_temp_words(n, cells) - 1;
    fill_u32(g.count, cells, 0u, s);
    fill_u32(n_valid, 1, 0u, s);
    if (n == 0) {
        if (scan_exclusive(g.count, g.offset, cells, work, nullptr, s)) return -1;
        return 0;
    }
    hipLaunchKernelGGL(k_grid_hash, dim3((n + 255) / 256), dim3(256), 0, s, photons, n, np, nf, fa,
                       g.cell_id, keys, vals, g.count, n_valid);
    if (scan_exclusive(g.count, g.offset, cells, work, nullptr, s)) return -1;
    int bits = 1;
    while ((1ull << bits) <= (uint64_t)cells) bits++;  // keys in [0, cells]
    bool in_tmp = false;
    if (radix_sort_pairs(keys, vals, n, bits, work, s, &in_tmp)) return -1;
    const uint32_t* sorted_slot = in_tmp ? work + n : vals;
    if (hipMemcpyAsync(g.sorted_slot, sorted_slot, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_grid_copy, dim3((n + 255) / 256), dim3(256), 0, s, photons, g.sorted_slot,
                       n_valid, g.sorted);
    if (hipMemcpyAsync(g.n_valid, n_valid, sizeof(uint32_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return -1;
    return 0;
}

}  // namespace wsrt
