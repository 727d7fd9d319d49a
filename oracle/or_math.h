/*
 * or_math.h - f32 vector helpers and the reference's spectral/RNG/BSDF
 * building blocks, restated for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle is the checker of the HIP path and is
 * never linked into the product library. Built with -ffp-contract=off so every
 * operation rounds as written (WGSL order of operations).
 */
#ifndef OR_MATH_H
#define OR_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#define OR_PI 3.14159265359f /* shader.wgsl:6 */
#define OR_LAMBDA_MIN 380.0f /* shader.wgsl:81-83 */
#define OR_LAMBDA_MAX 780.0f
#define OR_LAMBDA_RANGE 400.0f

/* Storage type of the f32 fields of the ABI structs (wsrt.h). The
 * double-precision build of the oracle rewrites every other float to
 * double and keeps the lines marked KEEP_F32. */
typedef float or_f32; /* KEEP_F32 */

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } v4;

static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v4 V4(float x, float y, float z, float w) { v4 r = {x, y, z, w}; return r; }
static inline v4 V4s(float s) { return V4(s, s, s, s); }
static inline v3 add3(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3s(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 neg3(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length3(v3 a) { return sqrtf(dot3(a, a)); }
/* normalize(v) = v * (1 / sqrt(dot(v, v))) - the form the HIP path uses too. */
static inline v3 normalize3(v3 a) {
    float inv = 1.0f / sqrtf(dot3(a, a));
    return mul3s(a, inv);
}
/* WGSL reflect(e1, e2) = e1 - 2 * dot(e2, e1) * e2 */
static inline v3 reflect3(v3 i, v3 n) { return sub3(i, mul3s(n, 2.0f * dot3(n, i))); }
/* WGSL refract(e1, e2, e3): k = 1 - e3^2 (1 - dot(e2,e1)^2); k < 0 -> 0 */
static inline v3 refract3(v3 i, v3 n, float eta) {
    float ni = dot3(n, i);
    float k = 1.0f - eta * eta * (1.0f - ni * ni);
    if (k < 0.0f) return V3(0.0f, 0.0f, 0.0f);
    return sub3(mul3s(i, eta), mul3s(n, eta * ni + sqrtf(k)));
}
static inline v4 add4(v4 a, v4 b) { return V4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4 mul4(v4 a, v4 b) { return V4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static inline v4 mul4s(v4 a, float s) { return V4(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline v4 div4s(v4 a, float s) { return V4(a.x / s, a.y / s, a.z / s, a.w / s); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float maxf(float a, float b) { return a > b ? a : b; }
static inline float minf(float a, float b) { return a < b ? a : b; }

/* ---- RNG: shader.wgsl:144-160, photon_emit.wgsl:64-81 ------------------- */
static inline uint32_t or_pcg(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
/* f32(u32) / 4294967295.0: the f32 literal is 2^32, so rand may equal 1.0. */
static inline float or_rand(uint32_t* s) {
    *s = or_pcg(*s);
    return (float)(*s) / 4294967296.0f;
}

/* ---- spectral helpers: shader.wgsl:85-139 ------------------------------ */
/* pow(x, 2.0) of a possibly negative x is taken as x * x. */
static inline float or_sq(float x) { return x * x; }

static inline v3 or_cie_xyz(float l) {
    float x = expf(-or_sq((l - 599.8f) / 37.9f)) * 1.056f +
              expf(-or_sq((l - 442.0f) / 16.0f)) * 0.362f -
              expf(-or_sq((l - 501.1f) / 26.7f)) * 0.065f;
    float y = expf(-or_sq((l - 568.8f) / 46.9f)) * 0.821f +
              expf(-or_sq((l - 530.9f) / 16.3f)) * 0.286f;
    float z = expf(-or_sq((l - 437.0f) / 11.8f)) * 1.217f +
              expf(-or_sq((l - 459.0f) / 26.0f)) * 0.681f;
    return V3(maxf(0.0f, x), maxf(0.0f, y), maxf(0.0f, z));
}

/* XYZ -> linear sRGB (shader.wgsl:100-104, column-major). */
static inline v3 or_xyz_to_srgb(v3 c) {
    return V3(3.2404542f * c.x + -1.5371385f * c.y + -0.4985314f * c.z,
              -0.9692660f * c.x + 1.8760108f * c.y + 0.0415560f * c.z,
              0.0556434f * c.x + -0.2040259f * c.y + 1.0572252f * c.z);
}

static inline float or_rgb_to_spectrum(v3 rgb, float l) {
    float white = minf(rgb.x, minf(rgb.y, rgb.z));
    float cr = maxf(rgb.x - white, 0.0f), cg = maxf(rgb.y - white, 0.0f),
          cb = maxf(rgb.z - white, 0.0f);
    float r = cr * expf(-or_sq((l - 630.0f) / 20.0f));
    float g = cg * expf(-or_sq((l - 525.0f) / 20.0f));
    float b = cb * expf(-or_sq((l - 460.0f) / 20.0f));
    return clampf(white + r + g + b, 0.0f, 0.99f);
}
static inline v4 or_rgb_to_spectrum4(v3 rgb, v4 wl) {
    return V4(or_rgb_to_spectrum(rgb, wl.x), or_rgb_to_spectrum(rgb, wl.y),
              or_rgb_to_spectrum(rgb, wl.z), or_rgb_to_spectrum(rgb, wl.w));
}

static inline float or_blackbody(float lambda_nm, float temp_k) {
    float lambda_m = lambda_nm * 1e-9f;
    const float h = 6.62607015e-34f, c = 299792458.0f, k = 1.380649e-23f;
    float c1 = 2.0f * h * c * c;
    float c2 = (h * c) / (k * temp_k);
    float radiance = c1 / (powf(lambda_m, 5.0f) * (expf(c2 / lambda_m) - 1.0f));
    return radiance * 1e-13f;
}
static inline v4 or_blackbody4(v4 wl, float t) {
    return V4(or_blackbody(wl.x, t), or_blackbody(wl.y, t), or_blackbody(wl.z, t),
              or_blackbody(wl.w, t));
}

/* Airy reflectance of a thin film, shader.wgsl:233-262 (per wavelength). */
static inline float or_thin_film(float cos1, float n1, float n2, float n3, float d, float l) {
    float s1 = maxf(0.0f, 1.0f - cos1 * cos1);
    float s2 = (n1 / n2) * (n1 / n2) * s1;
    if (s2 >= 1.0f) return 1.0f;
    float cos2 = sqrtf(1.0f - s2);
    float s3 = (n1 / n3) * (n1 / n3) * s1;
    if (s3 >= 1.0f) return 1.0f;
    float cos3 = sqrtf(1.0f - s3);
    float r12 = (n1 * cos1 - n2 * cos2) / (n1 * cos1 + n2 * cos2);
    float r23 = (n2 * cos2 - n3 * cos3) / (n2 * cos2 + n3 * cos3);
    float r12s = r12 * r12, r23s = r23 * r23;
    float delta = (4.0f * OR_PI * n2 * d * cos2) / l;
    float cd = cosf(delta);
    float num = r12s + r23s + 2.0f * r12 * r23 * cd;
    float den = 1.0f + r12s * r23s + 2.0f * r12 * r23 * cd;
    return clampf(num / den, 0.0f, 1.0f);
}

/* Schlick, shader.wgsl:264-268 */
static inline float or_reflectance(float cosine, float ref_idx) {
    float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * powf(1.0f - cosine, 5.0f);
}

/* Hero-wavelength Cauchy constants, shader.wgsl:734-741 */
static inline void or_cauchy_cdf(float* cdf_min, float* cdf_range) {
    const float peak = 555.0f, gamma = 40.0f;
    float mn = (1.0f / OR_PI) * atanf((OR_LAMBDA_MIN - peak) / gamma) + 0.5f;
    float mx = (1.0f / OR_PI) * atanf((OR_LAMBDA_MAX - peak) / gamma) + 0.5f;
    *cdf_min = mn;
    *cdf_range = mx - mn;
}
static inline float or_cauchy_sample(float u01, float cdf_min, float cdf_range) {
    float u = cdf_min + u01 * cdf_range;
    float l = 555.0f + 40.0f * tanf(OR_PI * (u - 0.5f));
    return clampf(l, OR_LAMBDA_MIN, OR_LAMBDA_MAX);
}
static inline float or_cauchy_pdf(float l, float cdf_range) {
    float pdf_norm = 1.0f / (cdf_range * OR_PI * 40.0f);
    float d = (l - 555.0f) / 40.0f;
    return pdf_norm / (1.0f + d * d);
}
/* WGSL float % : x - y * trunc(x / y) */
static inline float or_fmod_wgsl(float x, float y) { return x - y * truncf(x / y); }

/* Grid jitter, build_grid.wgsl:25-38 / shader.wgsl:282-295 */
static inline v3 or_grid_jitter(uint32_t frame, float cell) {
    uint32_t state = (frame * 114514u) * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    uint32_t seed = (word >> 22u) ^ word;
    float rx = (float)(seed & 0xFFu) / 255.0f;
    float ry = (float)((seed >> 8u) & 0xFFu) / 255.0f;
    float rz = (float)((seed >> 16u) & 0xFFu) / 255.0f;
    return V3((rx - 0.5f) * cell, (ry - 0.5f) * cell, (rz - 0.5f) * cell);
}
/* build_grid.wgsl:41-52 (hash_size a power of two). */
static inline uint32_t or_hash_cell(v3 p, v3 jitter, float cell, uint32_t hash_size) {
    int32_t gx = (int32_t)floorf((p.x + jitter.x) / cell);
    int32_t gy = (int32_t)floorf((p.y + jitter.y) / cell);
    int32_t gz = (int32_t)floorf((p.z + jitter.z) / cell);
    uint32_t h = ((uint32_t)gx * 73856093u) ^ ((uint32_t)gy * 19349663u) ^
                 ((uint32_t)gz * 83492791u);
    return h & (hash_size - 1u);
}

#endif /* OR_MATH_H */
