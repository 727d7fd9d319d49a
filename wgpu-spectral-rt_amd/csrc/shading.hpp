// shading.hpp - per-hit shading pieces shared by the photon and camera
// kernels: interpolated world normal, light sampling, BSDF sampling and the
// photon gather (shader.wgsl:171-545, photon_emit.wgsl:134-149).
#pragma once

#include "common.hpp"
#include "dev_math.hpp"
#include "traverse.hpp"

namespace wsrt {

// Frame constants are read from a device table indexed by a device-side
// frame counter, so a captured frame graph reads the current frame's values.
__device__ __forceinline__ FrameConst frame_const(const FrameArgs& fa) {
    return fa.table[*fa.frame & fa.mask];
}

// shader.wgsl:530-545 and :573-579: interpolated object normal, world normal
// through the transposed world-to-object matrix.
// World-space area of world triangle tri: its object-space edges through
// the linear part of the instance transform.
__device__ __forceinline__ float world_tri_area(const SceneView& sv, uint32_t tri) {
    const Instance& in = sv.instances[sv.tri_instance[tri]];
    const wsrt_mesh_info mi = sv.meshes[in.mesh];
    const uint32_t io = mi.index_offset + (tri - in.first_tri) * 3u;
    const float4 p0 = *reinterpret_cast<const float4*>(sv.verts[sv.indices[io] + mi.vertex_offset].pos);
    const float4 p1 = *reinterpret_cast<const float4*>(sv.verts[sv.indices[io + 1] + mi.vertex_offset].pos);
    const float4 p2 = *reinterpret_cast<const float4*>(sv.verts[sv.indices[io + 2] + mi.vertex_offset].pos);
    const float* M = in.m;
    const f3 c0 = mk3(M[0], M[1], M[2]), c1 = mk3(M[4], M[5], M[6]), c2 = mk3(M[8], M[9], M[10]);
    const f3 a = xyz(p1) - xyz(p0), b = xyz(p2) - xyz(p0);
    const f3 wa = c0 * a.x + c1 * a.y + c2 * a.z, wb = c0 * b.x + c1 * b.y + c2 * b.z;
    return length(cross(wa, wb)) * 0.5f;
}

__device__ __forceinline__ f3 hit_world_normal(const SceneView& sv, uint32_t tri, float u, float v,
                                               uint32_t* material) {
    const uint32_t ii = sv.tri_instance[tri];
    const Instance& in = sv.instances[ii];
    *material = in.material;
    const uint32_t prim = tri - in.first_tri;
    const wsrt_mesh_info mi = sv.meshes[in.mesh];
    const uint32_t io = mi.index_offset + prim * 3u;
    const uint32_t i0 = sv.indices[io] + mi.vertex_offset;
    const uint32_t i1 = sv.indices[io + 1] + mi.vertex_offset;
    const uint32_t i2 = sv.indices[io + 2] + mi.vertex_offset;
    const float4 n0 = *reinterpret_cast<const float4*>(sv.verts[i0].normal);
    const float4 n1 = *reinterpret_cast<const float4*>(sv.verts[i1].normal);
    const float4 n2 = *reinterpret_cast<const float4*>(sv.verts[i2].normal);
    float w = 1.0f - u - v;
    f3 ln = mk3(n0.x * w + n1.x * u + n2.x * v, n0.y * w + n1.y * u + n2.y * v,
                n0.z * w + n1.z * u + n2.z * v);
    ln = normalize(ln);
    const float* W = in.w2o;
    f3 wn = mk3(dot(ln, mk3(W[0], W[1], W[2])), dot(ln, mk3(W[3], W[4], W[5])),
                dot(ln, mk3(W[6], W[7], W[8])));
    return normalize(wn);
}

struct LightSample {
    f3 dir;
    float dist;
    float pdf;
    float4 radiance;
    bool delta;
};

// shader.wgsl:477-528; light kind 1 (point light) is an extension that
// draws the same three numbers.
__device__ __forceinline__ LightSample sample_light(const SceneView& sv, f3 hp, float4 wl,
                                                   uint32_t& rng) {
    LightSample ls;
    ls.pdf = 0.0f;
    ls.delta = false;
    ls.dist = 0.0f;
    ls.dir = mk3(0, 0, 0);
    ls.radiance = splat4(0.0f);
    const uint32_t nl = sv.n_lights;
    if (nl == 0u) return ls;
    uint32_t li = (uint32_t)(rand01(rng) * (float)nl);
    if (li > nl - 1u) li = nl - 1u;
    const wsrt_light_info& L = sv.lights[li];
    const float4 v0 = *reinterpret_cast<const float4*>(L.v0);
    const float4 prm = *reinterpret_cast<const float4*>(L.params);
    if (prm.x == 1.0f) {
        (void)rand01(rng);
        (void)rand01(rng);
        f3 dl = xyz(v0) - hp;
        float dsq = dot(dl, dl);
        float dist = sqrtf(dsq);
        ls.dir = dl * (1.0f / dist);
        ls.dist = dist;
        ls.pdf = dsq / (float)nl;
        ls.delta = true;
        ls.radiance = mul4s(blackbody4(wl, prm.y), prm.z);
        return ls;
    }
    const float4 v1 = *reinterpret_cast<const float4*>(L.v1);
    const float4 v2 = *reinterpret_cast<const float4*>(L.v2);
    float r1 = rand01(rng), r2 = rand01(rng), sr = sqrtf(r1);
    float u = 1.0f - sr, v = r2 * sr, w = 1.0f - u - v;
    f3 a = xyz(v0), b = xyz(v1), c = xyz(v2);
    f3 p = a * u + b * v + c * w;
    f3 e1 = b - a, e2 = c - a;
    f3 nrm = normalize(cross(e1, e2));
    float area = length(cross(e1, e2)) * 0.5f;
    f3 dl = p - hp;
    float dsq = dot(dl, dl);
    float dist = sqrtf(dsq);
    f3 l = mk3(dl.x / dist, dl.y / dist, dl.z / dist);
    float cl = fmaxf(dot(nrm, -l), 0.0f);
    if (cl <= 0.0001f) return ls;
    ls.dir = l;
    ls.dist = dist;
    ls.pdf = dsq / (area * (float)nl * cl);
    ls.radiance = mul4s(blackbody4(wl, prm.y), prm.z);
    return ls;
}

// shader.wgsl:171-188, WGSL sign(0) = 0 (normals with z == 0 give a NaN
// basis), or sign(0) = +1 with sign_fix (WSRT_FIX_ONB_SIGN).
__device__ __forceinline__ void onb(f3 n, f3& right, f3& up, bool sign_fix) {
    float sg = n.z > 0.0f ? 1.0f : (n.z < 0.0f ? -1.0f : (sign_fix ? 1.0f : 0.0f));
    float a = -1.0f / (sg + n.z);
    float b = n.x * n.y * a;
    right = mk3(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x);
    up = mk3(b, sg + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ float g1_schlick(float ndv, float r) {
    float k = (r * r) / 2.0f;
    return ndv / (ndv * (1.0f - k) + k);
}

struct Scatter {
    f3 dir;
    float4 att;
    bool absorbed;
};

// shader.wgsl:347-471
__device__ __forceinline__ Scatter sample_bsdf(f3 rdir, f3 ff, const wsrt_material& m, bool front,
                                               float4 wl, float4 tp, uint32_t& rng, uint32_t quirks) {
    Scatter rec;
    rec.absorbed = false;
    rec.dir = mk3(0, 0, 0);
    rec.att = splat4(0.0f);
    const float4 col = *reinterpret_cast<const float4*>(m.color);
    const float4 ex = *reinterpret_cast<const float4*>(m.extra);
    const uint32_t type = (uint32_t)ex.x;
    f3 v = normalize(-rdir);
    float4 spec = rgb_to_spectrum4(mk3(col.x, col.y, col.z), wl);
    if (type == 1u || type == 5u) {
        // 5 (extension): the same GGX lobe with the Fresnel factor of a thin
        // film (thickness color.r nm, index extra.z) on a substrate of index
        // extra.w, per hero wavelength, at the half-vector angle
        float rough = fmaxf(ex.y, 0.001f);
        float a2 = rough * rough;
        float r1 = rand01(rng), r2 = rand01(rng);
        float phi = 2.0f * kPi * r1;
        float ct = sqrtf((1.0f - r2) / (1.0f + (a2 - 1.0f) * r2));
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        f3 ht = mk3(cosf(phi) * st, sinf(phi) * st, ct);
        f3 rt, up;
        onb(ff, rt, up, (quirks & WSRT_FIX_ONB_SIGN) != 0u);
        f3 h = normalize(rt * ht.x + up * ht.y + ff * ht.z);
        f3 l = reflect(-v, h);
        float ndl = dot(ff, l), ndv = dot(ff, v);
        if (ndl > 0.0f && ndv > 0.0f) {
            float ndh = fmaxf(dot(ff, h), 0.0f), vdh = fmaxf(dot(v, h), 0.0f);
            float4 F;
            if (type == 5u) {
                const float d = fmaxf(col.x, 50.0f);
                F = make_float4(thin_film(vdh, 1.0f, ex.z, ex.w, d, wl.x), thin_film(vdh, 1.0f, ex.z, ex.w, d, wl.y),
                                thin_film(vdh, 1.0f, ex.z, ex.w, d, wl.z), thin_film(vdh, 1.0f, ex.z, ex.w, d, wl.w));
            } else {
                float p5 = powf(fminf(fmaxf(1.0f - vdh, 0.0f), 1.0f), 5.0f);
                F = make_float4(spec.x + (1.0f - spec.x) * p5, spec.y + (1.0f - spec.y) * p5,
                                spec.z + (1.0f - spec.z) * p5, spec.w + (1.0f - spec.w) * p5);
            }
            float G = g1_schlick(ndl, rough) * g1_schlick(ndv, rough);
            float den = fmaxf(ndv * ndh, 0.0001f);
            rec.dir = l;
            rec.att = div4s(mul4s(mul4s(F, G), vdh), den);
        } else {
            rec.absorbed = true;
        }
    } else if (type == 2u) {
        float A = ex.z, B = ex.w;
        float lum = wl.x / 1000.0f;
        float ir = A + (B / (lum * lum));
        float ratio = front ? 1.0f / ir : ir;
        f3 ud = normalize(rdir);
        float ct = fminf(dot(-ud, ff), 1.0f);
        float si2 = fmaxf(0.0f, 1.0f - ct * ct);
        float st2 = ratio * ratio * si2;
        if (st2 >= 1.0f) {
            rec.dir = reflect(ud, ff);
            rec.att = splat4(1.0f);
        } else {
            float ctt = sqrtf(1.0f - st2);
            float cf = front ? ct : ctt;
            if (schlick(cf, ratio) > rand01(rng)) {
                rec.dir = reflect(ud, ff);
                rec.att = splat4(1.0f);
            } else {
                rec.dir = refract(ud, ff, ratio);
                float scale = ratio * ratio;
                float4 base = mul4s(spec, scale);
                if (B > 0.0f) {
                    bool collapsed = tp.y < 0.0001f && tp.z < 0.0001f && tp.w < 0.0001f;
                    float mult = collapsed ? 1.0f : 4.0f;
                    rec.att = make_float4(base.x * mult, 0.0f, 0.0f, 0.0f);
                } else {
                    rec.att = base;
                }
            }
        }
    } else if (type == 4u) {
        float n2 = ex.z, n3 = ex.w, d = fmaxf(ex.y, 50.0f);
        f3 ud = normalize(rdir);
        float c1 = fminf(dot(-ud, ff), 1.0f);
        float4 F = make_float4(thin_film(c1, 1.0f, n2, n3, d, wl.x), thin_film(c1, 1.0f, n2, n3, d, wl.y),
                               thin_film(c1, 1.0f, n2, n3, d, wl.z), thin_film(c1, 1.0f, n2, n3, d, wl.w));
        float p = F.x;
        if (rand01(rng) < p) {
            rec.dir = reflect(ud, ff);
            rec.att = div4s(F, p);
        } else {
            rec.dir = ud;
            float4 T = make_float4(1.0f - F.x, 1.0f - F.y, 1.0f - F.z, 1.0f - F.w);
            rec.att = div4s(T, 1.0f - p);
        }
    } else {
        f3 sd = ff + random_unit_vector(rng);
        if (length(sd) < 0.001f) sd = ff;
        rec.dir = normalize(sd);
        rec.att = spec;
    }
    return rec;
}

// Environment light (extension): equirectangular RGB radiance, row 0 at +Y,
// nearest texel; u = atan2(z, x) / 2 pi + 1/2, v = acos(y) / pi. The RGB
// value is turned into a spectrum by the reference's reflectance fit
// (shader.wgsl:106-126) applied to rgb / max(rgb), scaled back by max(rgb).
__device__ __forceinline__ float4 env_radiance(const SceneView& sv, f3 d, float4 wl) {
    const float u = atan2f(d.z, d.x) * (0.5f / kPi) + 0.5f;
    const float v = acosf(fminf(fmaxf(d.y, -1.0f), 1.0f)) * (1.0f / kPi);
    const uint32_t x = min((uint32_t)fmaxf(u * (float)sv.env_w, 0.0f), sv.env_w - 1u);
    const uint32_t y = min((uint32_t)fmaxf(v * (float)sv.env_h, 0.0f), sv.env_h - 1u);
    const float* t = sv.env + 3ull * ((size_t)y * sv.env_w + x);
    const float r = t[0], g = t[1], b = t[2];
    const float m = fmaxf(fmaxf(r, g), b);
    if (!(m > 0.0f)) return splat4(0.0f);
    const float im = 1.0f / m;
    return mul4s(rgb_to_spectrum4(mk3(r * im, g * im, b * im), wl), m);
}

// shader.wgsl:297-342 over the counting-sort grid: the first gather_cap
// photons of the cell in slot order, read from one contiguous range.
__device__ __forceinline__ float4 gather_photons(const GridView& g, const FrameConst& fc,
                                                 uint32_t cell_base, f3 hp, f3 n, float4 wl) {
    float4 total = splat4(0.0f);
    f3 j = mk3(fc.jitter[0], fc.jitter[1], fc.jitter[2]);
    uint32_t h = cell_base + hash_cell(hp, j, fc.cell_size, fc.hash_size);
    uint32_t cnt = g.count[h];
    if (cnt > fc.gather_cap) cnt = fc.gather_cap;
    const uint32_t off = cnt ? g.offset[h] : 0u;  // offsets are written for non-empty cells only
    const float4* ph = reinterpret_cast<const float4*>(g.sorted);
    const float r = fc.radius, r2 = fc.radius_sq, sr = fc.spectral_radius;
    for (uint32_t k = 0; k < cnt; ++k) {
        float4 a = ph[2ull * (off + k)], b = ph[2ull * (off + k) + 1];
        f3 dd = hp - xyz(a);
        float dsq = dot(dd, dd);
        if (dsq < r2) {
            if (dot(n, xyz(b)) > 0.0f) {
                float dist = sqrtf(dsq);
                float sw = 1.0f - (dist / r);
                float es = b.w * sw;
                float lp = a.w;
                float4 spw = make_float4(fmaxf(1.0f - fabsf(wl.x - lp) / sr, 0.0f),
                                         fmaxf(1.0f - fabsf(wl.y - lp) / sr, 0.0f),
                                         fmaxf(1.0f - fabsf(wl.z - lp) / sr, 0.0f),
                                         fmaxf(1.0f - fabsf(wl.w - lp) / sr, 0.0f));
                total = add4(total, mul4s(spw, es));
            }
        }
    }
    return mul4s(total, fc.density);
}

}  // namespace wsrt
