// dev_math.hpp - device-side math of the render loop: f32 vectors, the
// reference's PCG hash RNG, spectral conversions and the Fresnel/BSDF pieces.
// Compiled with -ffp-contract=off: every expression rounds as written, in the
// WGSL order of operations of src/renderer/shaders/shader.wgsl.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wsrt {

constexpr float kPi = 3.14159265359f;  // shader.wgsl:6
constexpr float kLambdaMin = 380.0f;   // shader.wgsl:81-83
constexpr float kLambdaMax = 780.0f;
constexpr float kLambdaRange = 400.0f;

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ f3 reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
__device__ __forceinline__ f3 refract(f3 i, f3 n, float eta) {
    float ni = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - ni * ni);
    if (k < 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    return i * eta - n * (eta * ni + sqrtf(k));
}
__device__ __forceinline__ f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 mul4(float4 a, float4 b) {
    return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}
__device__ __forceinline__ float4 mul4s(float4 a, float s) {
    return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float4 div4s(float4 a, float s) {
    return make_float4(a.x / s, a.y / s, a.z / s, a.w / s);
}
__device__ __forceinline__ float4 splat4(float s) { return make_float4(s, s, s, s); }

// ---- RNG: PCG-RXS-M-XS hash (shader.wgsl:151-160) -------------------------
__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// f32(u32) / 4294967295.0 where the f32 literal is 2^32: exact power-of-two
// scaling of the round-to-nearest conversion, so rand() may return 1.0.
__device__ __forceinline__ float rand01(uint32_t& s) {
    s = pcg_hash(s);
    return (float)s * 2.3283064365386963e-10f;
}
__device__ __forceinline__ f3 random_unit_vector(uint32_t& s) {  // shader.wgsl:162-169
    float u1 = rand01(s);
    float u2 = rand01(s);
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    // cos / sin of phi = 2 pi u2 with the hardware functions, which take
    // their argument in revolutions (u2 in [0, 1)): a few ulp from the
    // correctly rounded forms, inside the parity bars
    return mk3(r * __builtin_amdgcn_cosf(u2), r * __builtin_amdgcn_sinf(u2), z);
}

// ---- spectral (shader.wgsl:85-139) -----------------------------------------
// pow(x, 2.0) of a possibly negative x is evaluated as x * x.
__device__ __forceinline__ float sq(float x) { return x * x; }

// The smooth spectral fits below (CIE lobes, rgb -> spectrum, Planck) run
// with the hardware exponential and multiplications by reciprocal constants:
// a few ulp from the IEEE forms the oracle uses, far inside the 1e-4
// relative parity bar (WGSL's own exp and division are not correctly rounded
// either), and several times fewer instructions.
__device__ __forceinline__ float gauss_lobe(float l, float mu, float inv_sigma) {
    return __expf(-sq((l - mu) * inv_sigma));
}
__device__ __forceinline__ f3 cie_xyz(float l) {
    float x = gauss_lobe(l, 599.8f, 1.0f / 37.9f) * 1.056f + gauss_lobe(l, 442.0f, 1.0f / 16.0f) * 0.362f -
              gauss_lobe(l, 501.1f, 1.0f / 26.7f) * 0.065f;
    float y = gauss_lobe(l, 568.8f, 1.0f / 46.9f) * 0.821f + gauss_lobe(l, 530.9f, 1.0f / 16.3f) * 0.286f;
    float z = gauss_lobe(l, 437.0f, 1.0f / 11.8f) * 1.217f + gauss_lobe(l, 459.0f, 1.0f / 26.0f) * 0.681f;
    return mk3(fmaxf(0.0f, x), fmaxf(0.0f, y), fmaxf(0.0f, z));
}
__device__ __forceinline__ f3 xyz_to_srgb(f3 c) {
    return mk3(3.2404542f * c.x + -1.5371385f * c.y + -0.4985314f * c.z,
               -0.9692660f * c.x + 1.8760108f * c.y + 0.0415560f * c.z,
               0.0556434f * c.x + -0.2040259f * c.y + 1.0572252f * c.z);
}
__device__ __forceinline__ float rgb_to_spectrum(f3 rgb, float l) {
    float white = fminf(rgb.x, fminf(rgb.y, rgb.z));
    float cr = fmaxf(rgb.x - white, 0.0f), cg = fmaxf(rgb.y - white, 0.0f),
          cb = fmaxf(rgb.z - white, 0.0f);
    float r = cr * gauss_lobe(l, 630.0f, 1.0f / 20.0f);
    float g = cg * gauss_lobe(l, 525.0f, 1.0f / 20.0f);
    float b = cb * gauss_lobe(l, 460.0f, 1.0f / 20.0f);
    return fminf(fmaxf(white + r + g + b, 0.0f), 0.99f);
}
__device__ __forceinline__ float4 rgb_to_spectrum4(f3 rgb, float4 wl) {
    return make_float4(rgb_to_spectrum(rgb, wl.x), rgb_to_spectrum(rgb, wl.y),
                       rgb_to_spectrum(rgb, wl.z), rgb_to_spectrum(rgb, wl.w));
}
__device__ __forceinline__ float blackbody(float lambda_nm, float temp_k) {
    float lambda_m = lambda_nm * 1e-9f;
    const float h = 6.62607015e-34f, c = 299792458.0f, k = 1.380649e-23f;
    float c1 = 2.0f * h * c * c;
    float c2 = (h * c) / (k * temp_k);
    const float l2 = lambda_m * lambda_m;
    const float inv_l = __builtin_amdgcn_rcpf(lambda_m);
    float radiance = c1 * __builtin_amdgcn_rcpf(l2 * l2 * lambda_m * (__expf(c2 * inv_l) - 1.0f));
    return radiance * 1e-13f;
}
__device__ __forceinline__ float4 blackbody4(float4 wl, float t) {
    return make_float4(blackbody(wl.x, t), blackbody(wl.y, t), blackbody(wl.z, t),
                       blackbody(wl.w, t));
}
// Airy reflectance of a film (shader.wgsl:233-262 / photon_emit.wgsl:107-132).
__device__ __forceinline__ float thin_film(float cos1, float n1, float n2, float n3, float d,
                                           float l) {
    float s1 = fmaxf(0.0f, 1.0f - cos1 * cos1);
    float s2 = (n1 / n2) * (n1 / n2) * s1;
    if (s2 >= 1.0f) return 1.0f;
    float cos2 = sqrtf(1.0f - s2);
    float s3 = (n1 / n3) * (n1 / n3) * s1;
    if (s3 >= 1.0f) return 1.0f;
    float cos3 = sqrtf(1.0f - s3);
    float r12 = (n1 * cos1 - n2 * cos2) / (n1 * cos1 + n2 * cos2);
    float r23 = (n2 * cos2 - n3 * cos3) / (n2 * cos2 + n3 * cos3);
    float r12s = r12 * r12, r23s = r23 * r23;
    float delta = (4.0f * kPi * n2 * d * cos2) / l;
    float cd = cosf(delta);
    float num = r12s + r23s + 2.0f * r12 * r23 * cd;
    float den = 1.0f + r12s * r23s + 2.0f * r12 * r23 * cd;
    return fminf(fmaxf(num / den, 0.0f), 1.0f);
}
__device__ __forceinline__ float schlick(float cosine, float ref_idx) {  // shader.wgsl:264-268
    float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * powf(1.0f - cosine, 5.0f);
}

// Hero-wavelength truncated Cauchy (peak 555, gamma 40) on [380, 780].
struct CauchyConst {
    float cdf_min, cdf_range;
};
__device__ __forceinline__ CauchyConst cauchy_const() {
    float mn = (1.0f / kPi) * atanf((kLambdaMin - 555.0f) / 40.0f) + 0.5f;
    float mx = (1.0f / kPi) * atanf((kLambdaMax - 555.0f) / 40.0f) + 0.5f;
    return CauchyConst{mn, mx - mn};
}
__device__ __forceinline__ float cauchy_sample(float u01, CauchyConst c) {
    float u = c.cdf_min + u01 * c.cdf_range;
    float l = 555.0f + 40.0f * tanf(kPi * (u - 0.5f));
    return fminf(fmaxf(l, kLambdaMin), kLambdaMax);
}
__device__ __forceinline__ float cauchy_pdf(float l, CauchyConst c) {
    float pdf_norm = 1.0f / (c.cdf_range * kPi * 40.0f);
    float d = (l - 555.0f) / 40.0f;
    return pdf_norm / (1.0f + d * d);
}
__device__ __forceinline__ float fmod_wgsl(float x, float y) { return x - y * truncf(x / y); }

// ---- photon grid hash (build_grid.wgsl:25-52) ------------------------------
__device__ __forceinline__ f3 grid_jitter(uint32_t frame, float cell) {
    uint32_t seed = pcg_hash(frame * 114514u);
    float rx = (float)(seed & 0xFFu) / 255.0f;
    float ry = (float)((seed >> 8u) & 0xFFu) / 255.0f;
    float rz = (float)((seed >> 16u) & 0xFFu) / 255.0f;
    return mk3((rx - 0.5f) * cell, (ry - 0.5f) * cell, (rz - 0.5f) * cell);
}
__device__ __forceinline__ uint32_t hash_cell(f3 p, f3 j, float cell, uint32_t hash_size) {
    int32_t gx = (int32_t)floorf((p.x + j.x) / cell);
    int32_t gy = (int32_t)floorf((p.y + j.y) / cell);
    int32_t gz = (int32_t)floorf((p.z + j.z) / cell);
    uint32_t h = ((uint32_t)gx * 73856093u) ^ ((uint32_t)gy * 19349663u) ^ ((uint32_t)gz * 83492791u);
    return h & (hash_size - 1u);
}

// ---- wave-level helpers (64-wide wavefronts) --------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
// Wave-aggregated append: one atomic per wave; returns this lane's slot in
// the array `counter` counts, or 0xFFFFFFFF for lanes with pred false. All
// lanes of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t m = __ballot(pred);
    if (m == 0) return 0xFFFFFFFFu;
    uint32_t leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + (uint32_t)__popcll(m & lanemask_lt()) : 0xFFFFFFFFu;
}

}  // namespace wsrt
