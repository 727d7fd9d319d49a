// path_core.hpp - the per-path steps of the camera path tracer, shared by
// the wavefront kernels (path.hip) and the persistent multi-scene batch
// kernel (batch.hip), so both run the same arithmetic in the same order.
#pragma once

#include "common.hpp"
#include "dev_math.hpp"
#include "shading.hpp"
#include "traverse.hpp"

namespace wsrt {

constexpr uint32_t kFlagDiffuse = 1u, kFlagCaustic = 2u;

// State of one camera path between vertices.
struct PathRegs {
    float4 o4, d4;  // origin, direction (w unused)
    float4 tp, L, wl;
    uint32_t rng, flags;
};

// Camera ray and hero wavelengths of pixel (x, y) (shader.wgsl:709-767).
__device__ __forceinline__ void camera_path(const FrameConst& fc, uint32_t x, uint32_t y, PathRegs& r) {
    uint32_t rng = pcg_hash(x + y * fc.width + fc.frame * 927163u);
    float jx = rand01(rng), jy = rand01(rng);
    float ux = ((float)x + jx) / (float)fc.width, uy = ((float)y + jy) / (float)fc.height;
    float v[4] = {ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, 1.0f, 1.0f};
    float tw[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = fc.M[0 + i] * v[0];
        s = s + fc.M[4 + i] * v[1];
        s = s + fc.M[8 + i] * v[2];
        s = s + fc.M[12 + i] * v[3];
        tw[i] = s;
    }
    f3 origin = mk3(fc.origin[0], fc.origin[1], fc.origin[2]);
    f3 dir = normalize(mk3(tw[0] / tw[3], tw[1] / tw[3], tw[2] / tw[3]) - origin);
    float u = rand01(rng);
    float rot = rand01(rng);
    const CauchyConst cc{fc.cdf_min, fc.cdf_range};
    float hero = cauchy_sample(u, cc);
    const float q = kLambdaRange / 4.0f;
    float4 w = make_float4(hero, hero + q * 1.0f, hero + q * 2.0f, hero + q * 3.0f);
    w.x = kLambdaMin + fmod_wgsl(w.x - kLambdaMin, kLambdaRange);
    w.y = kLambdaMin + fmod_wgsl(w.y - kLambdaMin, kLambdaRange);
    w.z = kLambdaMin + fmod_wgsl(w.z - kLambdaMin, kLambdaRange);
    w.w = kLambdaMin + fmod_wgsl(w.w - kLambdaMin, kLambdaRange);
    if (rot < 0.25f) {
    } else if (rot < 0.5f) {
        w = make_float4(w.y, w.z, w.w, w.x);
    } else if (rot < 0.75f) {
        w = make_float4(w.z, w.w, w.x, w.y);
    } else {
        w = make_float4(w.w, w.x, w.y, w.z);
    }
    r.o4 = f4(origin, 0.0f);
    r.d4 = f4(dir, 0.0f);
    r.tp = splat4(1.0f);
    r.L = splat4(0.0f);
    r.wl = w;
    r.rng = rng;
    r.flags = 0u;
}

// What a shaded vertex asks for next.
struct VertexOut {
    bool alive;     // continue with r.o4 / r.d4
    bool shadow;    // trace (hp, sdir, 0.001, stmax); add contrib if unoccluded
    bool deferred;  // guided Lambert vertex: scattering happens later
    f3 hp, sdir, ff;
    float stmax;
    float4 contrib, galb;
};

// One hit of a camera path (shader.wgsl:568-701): emission, photon gather
// (fs non-null), next-event estimation, BSDF sampling, Russian roulette.
// r.d4.w carries the cosine pdf of the last Lambert scattering (0 after a
// camera ray or a specular bounce) when the WSRT_FIX_EMISSION_MIS bit is set.
__device__ __forceinline__ VertexOut shade_vertex(const SceneView& sv, const GridView& grid,
                                                  const FrameConst* fs, uint32_t cell_base,
                                                  uint32_t rr_after, uint32_t quirks, uint32_t bounce, bool guided,
                                                  float t, uint32_t tri, bool front, float hu, float hv,
                                                  PathRegs& r, float light_mix = 1.0f) {
    VertexOut out;
    out.alive = out.shadow = out.deferred = false;
    out.hp = out.sdir = out.ff = mk3(0, 0, 0);
    out.stmax = 0.0f;
    out.contrib = out.galb = splat4(0.0f);
    uint32_t mat_id;
    f3 wn = hit_world_normal(sv, tri, hu, hv, &mat_id);
    f3 ff = front ? wn : -wn;
    const wsrt_material& m = sv.materials[mat_id];
    const float4 col = *reinterpret_cast<const float4*>(m.color);
    const float4 ex = *reinterpret_cast<const float4*>(m.extra);
    const uint32_t type = (uint32_t)ex.x;
    const f3 o = xyz(r.o4), d = xyz(r.d4);
    const float4 wl = r.wl;
    if (type == 3u) {  // emitter (shader.wgsl:586-631); the MIS weight is not applied
        if (!(r.flags & kFlagCaustic)) {
            float4 em = mul4s(blackbody4(wl, col.x), col.y);
            if (front || ex.y > 0.0f) {
                if ((quirks & WSRT_FIX_EMISSION_MIS) && r.d4.w > 0.0f) {
                    // balance heuristic against the light sample's pdf
                    const float cl = fmaxf(dot(ff, -d), 0.0001f);
                    // the light technique's density: uniform light sampling
                    // with probability light_mix, plus the guided light
                    // proposal's share (1 - light_mix) p_guide, which the
                    // guided scatter left in o.w (0 for plain paths)
                    const float lp = (t * t) / (world_tri_area(sv, tri) * cl * (float)sv.n_lights);
                    const float lq = light_mix * lp + r.o4.w;
                    r.L = add4(r.L, mul4s(mul4(em, r.tp), r.d4.w / (r.d4.w + lq)));
                } else {
                    r.L = add4(r.L, mul4(em, r.tp));
                }
            }
        }
        return out;
    }
    if (type == 0u) {
        r.flags = (r.flags | kFlagDiffuse) & ~kFlagCaustic;
    } else if (type == 2u) {
        if (r.flags & kFlagDiffuse) r.flags |= kFlagCaustic;
    }
    const f3 hp = o + d * t;
    out.hp = hp;
    if (type == 0u) {
        float4 spec = rgb_to_spectrum4(mk3(col.x, col.y, col.z), wl);
        if (fs) {
            float4 g = gather_photons(grid, *fs, cell_base, hp, ff, wl);
            r.L = add4(r.L, mul4(mul4(r.tp, div4s(spec, kPi)), g));
        }
        if (guided) {
            // next-event estimation and scattering happen later, in the
            // guided scatter kernel, which has the vertex's mixture
            out.deferred = true;
            out.ff = ff;
            out.galb = spec;
            r.o4 = f4(hp, 0.0f);
            return out;
        }
        LightSample ls = sample_light(sv, hp, wl, r.rng);
        if (ls.pdf > 0.0f) {
            float ct = fmaxf(dot(ff, ls.dir), 0.0f);
            float4 be = div4s(spec, kPi);
            float bp = ct / kPi;
            float w = ls.delta ? 1.0f / ls.pdf : (ls.pdf / (ls.pdf + bp)) / ls.pdf;
            out.contrib = mul4s(mul4s(mul4(mul4(r.tp, be), ls.radiance), ct), w);
            out.shadow = true;
            out.sdir = ls.dir;
            out.stmax = ls.dist - 0.001f;
        }
    }
    Scatter sc = sample_bsdf(d, ff, m, front, wl, r.tp, r.rng, quirks);
    if (!sc.absorbed) {
        r.tp = mul4(r.tp, sc.att);
        r.o4 = f4(hp, 0.0f);
        const float prev_pdf = ((quirks & WSRT_FIX_EMISSION_MIS) && type == 0u)
                                   ? fmaxf(dot(ff, sc.dir), 0.0001f) / kPi
                                   : 0.0f;
        r.d4 = f4(sc.dir, prev_pdf);
        out.alive = true;
        // Russian roulette (shader.wgsl:696-700); a guided renderer plays
        // none (see k_guide_scatter): its paths' throughput is low exactly
        // where the guide sends them toward light
        if (!guided && bounce > rr_after) {
            float pr = fmaxf(fmaxf(r.tp.x, r.tp.y), fmaxf(r.tp.z, r.tp.w));
            if (quirks & WSRT_FIX_RR_CLAMP) pr = fminf(pr, 1.0f);
            if (rand01(r.rng) > pr) out.alive = false;
            else r.tp = div4s(r.tp, pr);
        }
    }
    return out;
}

// Rays a path loop traced, and their traversal work while counting.
struct RayTally {
    uint32_t extend_rays = 0, extend_nodes = 0, extend_tris = 0;
    uint32_t shadow_rays = 0, shadow_nodes = 0, shadow_tris = 0;
};

// Bounces first .. max_depth-1 of one path, each in turn: the closest hit
// (a miss ends the path, adding the environment, shader.wgsl:566), the
// vertex, and its shadow ray. The same steps, in the same order, as one
// path's share of the wavefront kernels extend, shade and shadow.
__device__ __forceinline__ void trace_bounces(const SceneView& sv, const GridView& grid, const FrameConst* fs,
                                              uint32_t cell_base, uint32_t rr_after, uint32_t quirks, uint32_t first,
                                              uint32_t max_depth, PathRegs& r, Stack st, bool count,
                                              RayTally& tally) {
    for (uint32_t b = first; b < max_depth; ++b) {
        const Hit h = trace_ray<false>(sv, xyz(r.o4), xyz(r.d4), 0.0001f, 100.0f, st, count, tally.extend_nodes,
                                       tally.extend_tris);
        tally.extend_rays++;
        if (h.tri == kInvalid) {
            if (sv.env) r.L = add4(r.L, mul4(r.tp, env_radiance(sv, xyz(r.d4), r.wl)));
            break;
        }
        const VertexOut vo =
            shade_vertex(sv, grid, fs, cell_base, rr_after, quirks, b, false, h.t, h.tri, h.front, h.u, h.v, r);
        if (vo.shadow) {
            const Hit s = trace_ray<true>(sv, vo.hp, vo.sdir, 0.001f, vo.stmax, st, count, tally.shadow_nodes,
                                          tally.shadow_tris);
            tally.shadow_rays++;
            if (s.tri == kInvalid) r.L = add4(r.L, vo.contrib);
        }
        if (!vo.alive) break;
    }
}

// trace_bounces for a wave whose lanes each carry a path (`live` marks the
// lanes with one), the closest hits traversed by the whole wave with their
// triangle tests shared (trace_wave); the shadow rays stay per lane. Every
// lane of the wave calls it, in wave-uniform control flow; each path's
// result is trace_bounces', bit for bit. Not counting.
__device__ __forceinline__ void trace_bounces_wave(const SceneView& sv, const GridView& grid, const FrameConst* fs,
                                                   uint32_t cell_base, uint32_t rr_after, uint32_t quirks,
                                                   uint32_t first, uint32_t max_depth, PathRegs& r, Stack st,
                                                   TriShare& sh, bool live, RayTally& tally) {
    for (uint32_t b = first; b < max_depth; ++b) {
        if (!__ballot(live)) break;
        uint32_t nn = 0, nt = 0;
        const Hit h = trace_wave<false>(sv, xyz(r.o4), xyz(r.d4), 0.0001f, 100.0f, live, st, sh, nn, nt);
        if (!live) continue;
        tally.extend_rays++;
        if (h.tri == kInvalid) {
            if (sv.env) r.L = add4(r.L, mul4(r.tp, env_radiance(sv, xyz(r.d4), r.wl)));
            live = false;
            continue;
        }
        const VertexOut vo =
            shade_vertex(sv, grid, fs, cell_base, rr_after, quirks, b, false, h.t, h.tri, h.front, h.u, h.v, r);
        if (vo.shadow) {
            const Hit s = trace_ray<true>(sv, vo.hp, vo.sdir, 0.001f, vo.stmax, st, false, tally.shadow_nodes,
                                          tally.shadow_tris);
            tally.shadow_rays++;
            if (s.tri == kInvalid) r.L = add4(r.L, vo.contrib);
        }
        if (!vo.alive) live = false;
    }
}

// One sample into the running accumulation (shader.wgsl:771-808): hero
// wavelengths to XYZ over the sum of their pdfs, to linear sRGB, added as
// (rgb, 1); frame 0 restarts the accumulation.
__device__ __forceinline__ float4 sample_color(const FrameConst& fc, float4 wl, float4 L) {
    f3 xz = mk3(0, 0, 0);
    xz = xz + cie_xyz(wl.x) * L.x;
    xz = xz + cie_xyz(wl.y) * L.y;
    xz = xz + cie_xyz(wl.z) * L.z;
    xz = xz + cie_xyz(wl.w) * L.w;
    // sum of the hero wavelengths' Cauchy pdfs (cauchy_pdf), with hardware
    // reciprocals
    float q = 0.0f;
    const float wv[4] = {wl.x, wl.y, wl.z, wl.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float d = (wv[i] - 555.0f) * (1.0f / 40.0f);
        q += __builtin_amdgcn_rcpf(1.0f + d * d);
    }
    const float sp = q * __builtin_amdgcn_rcpf(fc.cdf_range * kPi * 40.0f);
    xz = xz * __builtin_amdgcn_rcpf(sp);
    f3 c = xyz_to_srgb(xz);
    return make_float4(c.x, c.y, c.z, 1.0f);
}
__device__ __forceinline__ float4 accumulate_sample(float4 na, const FrameConst& fc, float4 wl, float4 L) {
    float4 cur = fc.frame > 0u ? na : splat4(0.0f);
    return add4(cur, sample_color(fc, wl, L));
}

// Mean, clamp at 0, Reinhard on luma, RGBA8 without gamma (shader.wgsl:809-822).
__device__ __forceinline__ uchar4 tone_map(float4 na) {
    f3 fcol = mk3(fmaxf(na.x / na.w * 1.0f, 0.0f), fmaxf(na.y / na.w * 1.0f, 0.0f),
                  fmaxf(na.z / na.w * 1.0f, 0.0f));
    float luma = dot(fcol, mk3(0.2126f, 0.7152f, 0.0722f));
    float ml = luma / (luma + 1.0f);
    if (luma > 0.00001f) fcol = fcol * (ml / luma);
    uchar4 o;
    o.x = (uint8_t)__float2int_rn(fminf(fmaxf(fcol.x, 0.0f), 1.0f) * 255.0f);
    o.y = (uint8_t)__float2int_rn(fminf(fmaxf(fcol.y, 0.0f), 1.0f) * 255.0f);
    o.z = (uint8_t)__float2int_rn(fminf(fmaxf(fcol.z, 0.0f), 1.0f) * 255.0f);
    o.w = 255;
    return o;
}

}  // namespace wsrt
