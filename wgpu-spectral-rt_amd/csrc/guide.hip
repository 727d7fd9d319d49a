// guide.hip - learned path guiding on CDNA4: the network's forward pass
// fused with one-sample MIS scattering at Lambert vertices, and its online
// training (forward + backward, weight gradients, Adam), all as BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
//
// One wave evaluates the network for 64 samples: four 16-sample column
// tiles, features as rows. The weights of a block live in LDS as
// pre-packed A operands (one 16-byte read per lane per 16x32 block), each
// layer's result feeds the next layer's B operand in registers (chain order,
// see guide.hpp), and only the 40 outputs go through LDS, where each lane
// picks up the mixture of its own sample.
#include <algorithm>

#include "guide.hpp"
#include "kernels.hpp"
#include "shading.hpp"

namespace wsrt {
namespace {

using namespace guide;

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kStage = 41;            // floats per sample row of the output staging
constexpr int kScatterWaves = 8;      // waves per block of the standalone inference kernel
// waves per block of the scattering kernel: 3 per SIMD, which its register
// budget (<= 168 VGPRs) and LDS (weights + FP16 output staging) allow
constexpr int kScatterBlockWaves = 12;
constexpr int kTrainWaves = 4;        // waves per block of the training kernel
// 16-sample tiles per wave in the training forward and backward passes. A
// training step covers 16384 samples; one tile per wave gives it 1024 waves,
// one per SIMD, and the two passes run 27 % faster than with 2 and 4 tiles.
constexpr int kTrainTiles = 1;
constexpr int kBwdTiles = 1;
// path flags 4 << b: the Lambert vertex of bounce b is recorded for training

// Round to nearest even: gfx950's v_cvt_pk_bf16_f32 (a NaN stays a quiet
// NaN), paired into one instruction per two values where the compiler can.
__device__ __forceinline__ short bf16_bits(float f) {
    return __builtin_bit_cast(short, (__bf16)f);
}
__device__ __forceinline__ float bf16_float(short h) {
    return __uint_as_float((uint32_t)(uint16_t)h << 16);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copies n16 16-byte words from global to LDS (whole block).
__device__ __forceinline__ void copy_to_lds(uint4* dst, const uint4* src, int n16) {
    for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
}

// Activation store of one training step: [row][Ns] BF16, column = sample.
struct Act {
    short* base;
    uint32_t ns;
    __device__ __forceinline__ void put(uint32_t row_off, uint32_t row, uint32_t col, float v) const {
        base[(size_t)(row_off + row) * ns + col] = bf16_bits(v);
    }
};
// row offsets (in rows of Ns) of the activation tensors
constexpr uint32_t kH0 = 0, kH1 = 192;  // H2..H4 follow H1 at 64-row steps
constexpr uint32_t kD1 = 448, kD2 = 512, kD3 = 576, kD4 = 640, kD5 = 704, kActRows = 752;

// Forward pass of one wave over 16 NT samples. s_own: the 7 scalars of the
// lane's own sample (sample index = lane). Leaves the 40 outputs of sample
// k at stage[k * kStage + o]. With kTrain, also stores the layer inputs to
// act (columns col0 .. col0 + 16 NT - 1) and returns the ReLU masks.
// T0: the first 16-sample tile (samples 16 T0 .. 16 (T0 + NT) - 1 of the
// wave's lanes); StageT: the staging type of the outputs.
template <bool kTrain, int NT, int T0 = 0, typename StageT = float>
__device__ void mlp_forward(const bf16x8* __restrict__ Wf, const float* __restrict__ bias,
                            const float (&s_own)[kScalars], StageT* __restrict__ stage,
                            uint32_t (&mask)[4][2], const Act& act, uint32_t col0) {
    const int l = __lane_id(), g = l >> 4, n = l & 15;
    float x[NT][kScalars];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int a = 0; a < kScalars; ++a) x[t][a] = __shfl(s_own[a], 16 * (T0 + t) + n, 64);

    f32x4 C[4][NT];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int t = 0; t < NT; ++t) C[mb][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // layer 1: the encoding is computed straight into the B operand
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) {
        const int q = 4 * kb + g;
        bf16x8 B[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (q < 3 * kScalars) {
                const int a = q / 3, bq = q % 3;
                float xa = x[t][0];
#pragma unroll
                for (int k = 1; k < kScalars; ++k)
                    if (a == k) xa = x[t][k];
                // band 4 bq: sin and cos of 2 pi frac(2^(4 bq - 1) x) by the
                // hardware (arguments in revolutions); the quad's next three
                // bands by the double-angle formulas, full-rate arithmetic in
                // place of six quarter-rate transcendentals (the error grows
                // 2x per doubling, to ~1e-6, far below BF16's 2^-9)
                const float tt = ldexpf(xa, 4 * bq - 1);
                const float u = tt - floorf(tt);
                float sn = __builtin_amdgcn_sinf(u), cs = __builtin_amdgcn_cosf(u);
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    B[t][2 * bb] = bf16_bits(sn);
                    B[t][2 * bb + 1] = bf16_bits(cs);
                    const float s2 = 2.0f * sn * cs;
                    cs = 1.0f - 2.0f * sn * sn;
                    sn = s2;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) B[t][j] = 0;
            }
            if (kTrain) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    act.base[(size_t)(kH0 + 32 * kb + 8 * g + j) * act.ns + col0 + 16 * t + n] = B[t][j];
            }
        }
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const bf16x8 A = Wf[kF1 / 8 + (mb * 6 + kb) * 64 + l];
#pragma unroll
            for (int t = 0; t < NT; ++t) C[mb][t] = mfma(A, B[t], C[mb][t]);
        }
    }
    // hidden layers 2..4 and the output layer
#pragma unroll
    for (int layer = 1; layer <= 4; ++layer) {
        const float* b = bias + 64 * (layer - 1);
        bf16x8 B[2][NT];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mb + 4 * g + r;
                    const float z = C[mb][t][r] + b[row];
                    const float h = fmaxf(z, 0.0f);
                    if (kTrain) {
                        const int bit = mb * 16 + t * 4 + r;
                        if (z > 0.0f) mask[layer - 1][bit >> 5] |= 1u << (bit & 31);
                        act.put(kH0 + 192 + 64 * (layer - 1), row, col0 + 16 * t + n, h);
                    }
                    B[mb >> 1][t][(mb & 1) * 4 + r] = bf16_bits(h);
                }
            }
        const int off = layer == 1 ? kF2 : layer == 2 ? kF3 : layer == 3 ? kF4 : kF5;
        const int MB = layer == 4 ? 3 : 4;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int t = 0; t < NT; ++t) C[mb][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                if (mb >= MB) continue;
                const bf16x8 A = Wf[off / 8 + (mb * 2 + kb) * 64 + l];
#pragma unroll
                for (int t = 0; t < NT; ++t) C[mb][t] = mfma(A, B[kb][t], C[mb][t]);
            }
    }
    const float* b5 = bias + 256;
#pragma unroll
    for (int mb = 0; mb < 3; ++mb)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * mb + 4 * g + r;
                if (row < kOut) stage[(16 * (T0 + t) + n) * kStage + row] = (StageT)(C[mb][t][r] + b5[row]);
            }
}

// The mixture of one sample from its 40 outputs.
struct Mixture {
    f3 mu[kLobes];
    float kappa[kLobes], pi[kLobes], norm[kLobes], len[kLobes], s[kLobes];
};
__device__ __forceinline__ void mixture_from(const float* o, Mixture& m) {
    float mx = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, kLogitScale * o[5 * k + 4]);
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const f3 v = mk3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
        const float len = fmaxf(sqrtf(dot(v, v)), 1e-6f);
        m.len[k] = len;
        m.mu[k] = v * (1.0f / len);
        m.s[k] = kKappaScale * o[5 * k + 3];  // log kappa before the clamp
        const float kap = expf(fminf(fmaxf(m.s[k], kLogKappaMin), kLogKappaMax));
        m.kappa[k] = kap;
        m.norm[k] = kap / (2.0f * kPi * -expm1f(-2.0f * kap));
        m.pi[k] = expf(kLogitScale * o[5 * k + 4] - mx);
        z += m.pi[k];
    }
    const float iz = 1.0f / z;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) m.pi[k] *= iz;
}
// The same mixture for the tracer's sampling and densities, with the
// hardware's approximate transcendentals (v_exp_f32, v_log_f32, v_sin_f32,
// v_cos_f32, v_rcp_f32, v_sqrt_f32, v_rsq_f32; about 1 ulp) in place of the
// correctly rounded library forms the training kernels keep. A sampled
// direction and the density its weight uses come from the same Mixture, so
// the one-sample MIS estimate stays unbiased; only the density's rounding
// differs from the exact formula's.
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ void mixture_from_fast(const float* o, Mixture& m) {
    float mx = -3.0e38f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, kLogitScale * o[5 * k + 4]);
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const f3 v = mk3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
        const float len = fmaxf(__builtin_amdgcn_sqrtf(dot(v, v)), 1e-6f);
        m.len[k] = len;
        m.mu[k] = v * fast_rcp(len);
        m.s[k] = kKappaScale * o[5 * k + 3];
        const float kap = fast_exp(fminf(fmaxf(m.s[k], kLogKappaMin), kLogKappaMax));
        m.kappa[k] = kap;
        // 1 - e^{-2 kappa} >= 0.095 for kappa >= e^-3 (kLogKappaMin)
        m.norm[k] = kap * fast_rcp(2.0f * kPi * (1.0f - fast_exp(-2.0f * kap)));
        m.pi[k] = fast_exp(kLogitScale * o[5 * k + 4] - mx);
        z += m.pi[k];
    }
    const float iz = fast_rcp(z);
#pragma unroll
    for (int k = 0; k < kLobes; ++k) m.pi[k] *= iz;
}
__device__ __forceinline__ float mixture_pdf_fast(const Mixture& m, f3 w) {
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) p += m.pi[k] * m.norm[k] * fast_exp(m.kappa[k] * (dot(m.mu[k], w) - 1.0f));
    return p;
}
// A direction from the vMF lobe (mu, kappa): cos theta by inverting the CDF,
// 1 + log(e^{-2k} + x1 (1 - e^{-2k})) / k, azimuth 2 pi x2, in the frame of
// mu's branch-free orthonormal basis.
__device__ __forceinline__ f3 vmf_sample_fast(f3 mu, float kap, float x1, float x2) {
    const float e = fast_exp(-2.0f * kap);
    // log(y) = log2(y) ln 2; y >= e^{-2 kappa} > 0
    float w = 1.0f + __builtin_amdgcn_logf(e + x1 * (1.0f - e)) * 0.693147181f * fast_rcp(kap);
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    const float sw = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float sgn = copysignf(1.0f, mu.z);
    const float a = -fast_rcp(sgn + mu.z);
    const float b = mu.x * mu.y * a;
    const f3 t1 = mk3(1.0f + sgn * mu.x * mu.x * a, sgn * b, -sgn * mu.x);
    const f3 t2 = mk3(b, sgn + mu.y * mu.y * a, -mu.y);
    // v_sin/v_cos take revolutions: x2 in [0, 1) is the azimuth / 2 pi
    const f3 d = mu * w + t1 * (sw * __builtin_amdgcn_cosf(x2)) + t2 * (sw * __builtin_amdgcn_sinf(x2));
    return d * __builtin_amdgcn_rsqf(dot(d, d));
}

// Network inputs of a hit: position over the scene box, normal, hero wavelength.
__device__ __forceinline__ void scalars_of(const GuideView& gv, f3 hp, f3 n, float wl, float (&s)[kScalars]) {
    s[0] = fminf(fmaxf((hp.x - gv.bmin[0]) * gv.binv[0], 0.0f), 1.0f);
    s[1] = fminf(fmaxf((hp.y - gv.bmin[1]) * gv.binv[1], 0.0f), 1.0f);
    s[2] = fminf(fmaxf((hp.z - gv.bmin[2]) * gv.binv[2], 0.0f), 1.0f);
    s[3] = 0.5f * (n.x + 1.0f);
    s[4] = 0.5f * (n.y + 1.0f);
    s[5] = 0.5f * (n.z + 1.0f);
    s[6] = fminf(fmaxf((wl - 380.0f) * (1.0f / 400.0f), 0.0f), 1.0f);
}

__device__ __forceinline__ void load_biases(float* dst, const float* params) {
    for (int i = threadIdx.x; i < 4 * kHid + kOutPad; i += blockDim.x) {
        const int layer = i / kHid;
        const int off = layer == 0 ? kB1 : layer == 1 ? kB2 : layer == 2 ? kB3 : layer == 3 ? kB4 : kB5;
        dst[i] = params[off + (i - kHid * layer)];
    }
}

// A direction from the mixture: lobe k with probability pi_k, then its vMF.
__device__ __forceinline__ f3 mixture_sample_fast(const Mixture& m, uint32_t& rng) {
    // the lobe by walking the CDF, its mean and concentration selected in
    // the same walk as values (a lobe index used afterwards makes the
    // compiler put the mixture in scratch memory and load it indexed)
    const float xl = rand01(rng);
    float c = m.pi[0];
    float mx = m.mu[0].x, my = m.mu[0].y, mz = m.mu[0].z, kap = m.kappa[0];
#pragma unroll
    for (int j = 1; j < kLobes; ++j) {
        const bool take = xl >= c;
        c = take ? c + m.pi[j] : c;
        mx = take ? m.mu[j].x : mx;
        my = take ? m.mu[j].y : my;
        mz = take ? m.mu[j].z : mz;
        kap = take ? m.kappa[j] : kap;
    }
    const float x1 = rand01(rng), x2 = rand01(rng);
    return vmf_sample_fast(mk3(mx, my, mz), kap, x1, x2);
}

// Next-event estimation at a guided Lambert vertex with a learned light
// proposal. With probability light_mix the light point comes from the plain
// tracer's uniform light sampling (shader.wgsl:477-528); otherwise a
// direction is drawn from the vertex's mixture and the nearest area-light
// triangle along it is the light point (none, or its back: no sample). The
// sample is weighted by the density of the two proposals together, q =
// light_mix lp + (1 - light_mix) p_guide (solid angle; lp the uniform
// sampler's), so the estimate of the reference's NEE integral -- L f cos
// weighted by lp / (lp + cos/pi), shader.wgsl:650-671 -- is unbiased for any
// mixture; with mis (WSRT_FIX_EMISSION_MIS) the weight is the balance
// heuristic against the vertex's own scattering density p_mix instead.
struct GuidedNee {
    bool shadow;
    f3 dir;
    float dist;
    float4 contrib;
};
__device__ __forceinline__ GuidedNee guided_nee(const SceneView& sv, const Mixture& m, float guide_mix,
                                                float light_mix, bool mis, f3 hp, f3 ff, float4 wl, float4 tp,
                                                float4 spec, uint32_t& rng) {
    GuidedNee g;
    g.shadow = false;
    g.dir = mk3(0, 0, 0);
    g.dist = 0.0f;
    g.contrib = splat4(0.0f);
    const uint32_t nl = sv.n_lights;
    if (nl == 0u) return g;
    f3 dir = mk3(0, 0, 1);
    float dist = 0.0f, lp = 0.0f;
    float4 rad = splat4(0.0f);
    bool delta = false;
    if (light_mix >= 1.0f || rand01(rng) < light_mix) {
        const LightSample ls = sample_light(sv, hp, wl, rng);
        if (ls.pdf <= 0.0f) return g;
        dir = ls.dir;
        dist = ls.dist;
        lp = ls.pdf;
        rad = ls.radiance;
        delta = ls.delta;
    } else {
        dir = mixture_sample_fast(m, rng);
        // the nearest light triangle along dir (every light is an area
        // light here: light_mix < 1 only then)
        float best = 3.0e38f;
        uint32_t bi = nl;
        f3 bn = mk3(0, 0, 0);
        float barea = 0.0f;
        for (uint32_t li = 0; li < nl; ++li) {
            const wsrt_light_info& L = sv.lights[li];
            const f3 a = xyz(*reinterpret_cast<const float4*>(L.v0));
            const f3 e1 = xyz(*reinterpret_cast<const float4*>(L.v1)) - a;
            const f3 e2 = xyz(*reinterpret_cast<const float4*>(L.v2)) - a;
            float t, u, v;
            bool fr;
            if (tri_intersect(hp, dir, a, e1, e2, t, u, v, fr) && t > 0.001f && t < best) {
                best = t;
                bi = li;
                const f3 c = cross(e1, e2);
                bn = normalize(c);
                barea = length(c) * 0.5f;
            }
        }
        if (bi == nl) return g;
        const float cl = fmaxf(dot(bn, -dir), 0.0f);
        if (cl <= 0.0001f) return g;  // its back: the uniform sampler gives it no density either
        dist = best;
        lp = (best * best) / (barea * (float)nl * cl);
        const float4 prm = *reinterpret_cast<const float4*>(sv.lights[bi].params);
        rad = mul4s(blackbody4(wl, prm.y), prm.z);
    }
    const float ct = fmaxf(dot(ff, dir), 0.0f);
    if (ct <= 0.0f) return g;
    const float bp = ct / kPi;
    float w;
    if (delta) {
        w = 1.0f / lp;  // a point light (light_mix is 1 then)
    } else {
        const float pg = mixture_pdf_fast(m, dir);
        const float q = light_mix * lp + (1.0f - light_mix) * pg;
        w = mis ? 1.0f / (q + guide_mix * pg + (1.0f - guide_mix) * bp) : (lp / (lp + bp)) / q;
    }
    g.shadow = true;
    g.dir = dir;
    g.dist = dist;
    g.contrib = mul4s(mul4s(mul4(mul4(tp, div4s(spec, kPi)), rad), ct), w);
    return g;
}

// Lambert vertices of a bounce: network forward, next-event estimation with
// the learned light proposal, one-sample MIS between the guide (probability
// mix) and the reference's cosine sampling, throughput update f cos / p_mix
// (no Russian roulette), and the training records of the first four Lambert
// vertices of the paths a training step reads.
__global__ __launch_bounds__(64 * kScatterBlockWaves) void k_guide_scatter(SceneView sv, PathState ps, Queues qs,
                                                                      GuideView gv, FrameArgs fa, uint32_t bounce,
                                                                      uint32_t* __restrict__ next_queue,
                                                                      Counters* __restrict__ ctr) {
    __shared__ bf16x8 Wf[kFwdPack / 8];
    __shared__ float bias[4 * kHid + kOutPad];
    // outputs staged as FP16: sampling and the pdf both read these values,
    // so the mixture a vertex samples is the one its weight uses
    __shared__ _Float16 stage_all[kScatterBlockWaves][64 * kStage];
    __shared__ uint32_t pre[kSegs + 1];
    copy_to_lds(reinterpret_cast<uint4*>(Wf), reinterpret_cast<const uint4*>(gv.wf), kFwdPack / 8);
    load_biases(bias, gv.params);
    const uint32_t n = seg_prefix(qs, bounce, kQGuide, 1, pre);  // syncs the block
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctr->guide_samples, (unsigned long long)n);
    const uint32_t base_frame = *fa.frame;
    const bool mis = (fa.table[base_frame & fa.mask].quirks & WSRT_FIX_EMISSION_MIS) != 0u;
    const float light_mix = fa.table[base_frame & fa.mask].guide_light_mix;
    const int wave = threadIdx.x >> 6, l = __lane_id();
    _Float16* stage = stage_all[wave];
    const uint32_t wstride = gridDim.x * kScatterBlockWaves * 64;
    for (uint32_t base = (blockIdx.x * kScatterBlockWaves + wave) * 64; base < n; base += wstride) {
        const uint32_t i = base + l;
        const bool have = i < n;
        uint32_t p = 0;
        f3 hp = mk3(0, 0, 0), ff = mk3(0, 0, 1);
        float4 wl = splat4(500.0f);
        float s[kScalars];
        if (have) {
            p = ps.guide_q[seg_entry(pre, kSegs, i, qs.seg_cap)];
            hp = xyz(ps.o[p]);
            ff = xyz(ps.gn[p]);
            wl = ps.wl[p];
        }
        scalars_of(gv, hp, ff, wl.x, s);
        uint32_t mask[4][2];
        const Act none{nullptr, 0};
        // two passes of 32 samples: half the accumulator registers of one
        // 64-sample pass
        mlp_forward<false, 2, 0>(Wf, bias, s, stage, mask, none, 0);
        __builtin_amdgcn_sched_barrier(0);  // the passes do not interleave
        mlp_forward<false, 2, 2>(Wf, bias, s, stage, mask, none, 0);
        wave_sync();
        float o[kOut];
#pragma unroll
        for (int k = 0; k < kOut; ++k) o[k] = (float)stage[l * kStage + k];
        wave_sync();
        bool alive = false, shadow = false;
        GuidedNee nee;
        if (have) {
            Mixture m;
            mixture_from_fast(o, m);
            uint32_t rng = ps.rng[p];
            uint32_t flags = ps.flags[p];
            float4 tp = ps.tp[p];
            const float4 spec = ps.galb[p];
            const float4 Lcur = ps.L[p];
            nee = guided_nee(sv, m, gv.mix, light_mix, mis, hp, ff, wl, tp, spec, rng);
            shadow = nee.shadow;
            f3 w;
            if (rand01(rng) < gv.mix) {
                w = mixture_sample_fast(m, rng);
            } else {
                f3 sd = ff + random_unit_vector(rng);
                if (length(sd) < 0.001f) sd = ff;
                w = normalize(sd);
            }
            const float cw = dot(ff, w);
            if (cw > 0.0f) {
                const float pg = mixture_pdf_fast(m, w);
                const float pm = gv.mix * pg + (1.0f - gv.mix) * (cw / kPi);
                tp = mul4(mul4s(spec, (cw / kPi) / pm), tp);
                if (bounce < kRecBounces && gv.rec.in0) {
                    // kept only if a training step of its frame reads it
                    const uint32_t slot = fdiv(p, gv.rec.div_npx);
                    const uint32_t j = bounce * gv.rec.npx + (p - slot * gv.rec.npx);
                    const uint32_t fs = base_frame + slot;
                    const uint32_t phase = fs - fdiv(fs, gv.rec.div_stride) * gv.rec.stride;
                    const uint32_t k = j >= phase ? fdiv(j - phase, gv.rec.div_stride) : gv.rec.ns;
                    if (k < gv.rec.ns && k * gv.rec.stride + phase == j) {
                        const size_t ri = (size_t)slot * gv.rec.ns + k;
                        gv.rec.in0[ri] = make_float4(s[0], s[1], s[2], s[3]);
                        gv.rec.in1[ri] = make_float4(s[4], s[5], s[6], pm);
                        gv.rec.dir[ri] = make_float4(w.x, w.y, w.z, cw);
                        gv.rec.tp[ri] = tp;
                        gv.rec.Lb[ri] = Lcur;  // before this vertex's light sample
                        gv.rec.nee[ri] = shadow ? nee.contrib : splat4(0.0f);
                        flags |= kFlagRecorded << bounce;
                    }
                }
                // No Russian roulette on guided paths (the plain tracer keeps
                // the reference's, shader.wgsl:696-700, after bounce
                // rr_after): a guided step multiplies the throughput by
                // f cos / p_mix, which is small exactly where the guide sends
                // the path toward light (p_mix large), so a survival
                // probability of max(throughput) kills the useful paths and
                // weights the few survivors up. On the occluded room it took C
                // from 17.4 without roulette to 123 with it (each path then
                // runs to max_depth: 7.9 against 5.6 ms per spp).
                alive = true;
                if (alive) {
                    // for the balance weight of emission the direction
                    // reaches (WSRT_FIX_EMISSION_MIS): d.w its scattering
                    // density, o.w the light proposal's share of the light
                    // technique's density there
                    ps.d[p] = f4(w, mis ? pm : 0.0f);
                    if (mis) ps.o[p] = f4(hp, (1.0f - light_mix) * pg);
                    ps.tp[p] = tp;
                    ps.rng[p] = rng;
                    ps.flags[p] = flags;
                }
            }
            if (!alive) {
                ps.rng[p] = rng;
                ps.flags[p] = flags;
            }
        }
        const uint32_t se = seg_append(qs, bounce, kQShadow, qs.seg_cap, shadow);
        if (shadow) {
            float4* r = ps.shadow + 3ull * se;
            r[0] = f4(hp, nee.dist - 0.001f);
            r[1] = make_float4(nee.dir.x, nee.dir.y, nee.dir.z, __uint_as_float(p));
            r[2] = nee.contrib;
        }
        if (ps.tag) {
            if (alive) ps.tag[p] = (uint8_t)(bounce + 1);
        } else {
            const uint32_t ne = seg_append(qs, bounce + 1, kQActive, qs.seg_cap, alive);
            if (alive) next_queue[ne] = p;
        }
    }
}

// Training inputs of one step from the records of frame slot `slot`:
// sample k is record k of the slot, the vertex j = k * stride + phase
// (< kRecBounces npx, since stride = floor(kRecBounces npx / ns)) of bounce
// j / npx on pixel j % npx. The weight of a sample is
// L_in cos / p_mix, L_in the incident radiance along the sampled direction
// (radiance gathered after the vertex over the throughput after it, mean
// over the hero wavelengths that carry throughput); records a path did not
// make have weight 0.
__global__ __launch_bounds__(256) void k_guide_prepare(PathState ps, GuideView gv, FrameArgs fa, uint32_t npx,
                                                       uint32_t slot, uint32_t ns, uint32_t stride,
                                                       float4* __restrict__ in0, float4* __restrict__ in1,
                                                       float4* __restrict__ dir) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const uint32_t phase = (*fa.frame + slot) % stride;
    const uint64_t j = (uint64_t)k * stride + phase;
    const uint32_t bounce = (uint32_t)(j / npx);
    const uint32_t p = slot * npx + (uint32_t)(j % npx);
    const size_t ri = (size_t)slot * ns + k;
    float wgt = 0.0f;
    float4 a = splat4(0.0f), b = splat4(0.0f), d = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    if (bounce < kRecBounces && (ps.flags[p] & (kFlagRecorded << bounce))) {
        a = gv.rec.in0[ri];
        b = gv.rec.in1[ri];
        d = gv.rec.dir[ri];
        const float4 T = gv.rec.tp[ri], L = ps.L[p];
        float4 Lb = gv.rec.Lb[ri];
        // the vertex's own light sample is not incident radiance along the
        // sampled direction
        if (ps.flags[p] & (kFlagNeeVisible << bounce)) Lb = add4(Lb, gv.rec.nee[ri]);
        float sum = 0.0f;
        int cnt = 0;
        if (T.x > 0.0f) { sum += (L.x - Lb.x) / T.x; cnt++; }
        if (T.y > 0.0f) { sum += (L.y - Lb.y) / T.y; cnt++; }
        if (T.z > 0.0f) { sum += (L.z - Lb.z) / T.z; cnt++; }
        if (T.w > 0.0f) { sum += (L.w - Lb.w) / T.w; cnt++; }
        const float lin = cnt ? fmaxf(sum / (float)cnt, 0.0f) : 0.0f;
        wgt = lin * d.w / b.w;
        if (!(wgt < 3.0e38f)) wgt = 0.0f;  // inf / nan
    }
    in0[k] = a;
    in1[k] = make_float4(b.x, b.y, b.z, wgt);
    dir[k] = d;
}

// Forward pass of one training step, 16 kTrainTiles samples per wave, and
// the output gradient of the loss -w log p_guide(dir) averaged over the ns
// samples. Stores every layer's input and the output gradient to act.
__global__ __launch_bounds__(64 * kTrainWaves) void k_guide_train_fwd(GuideView gv, const float4* __restrict__ in0,
                                                                      const float4* __restrict__ in1,
                                                                      const float4* __restrict__ dir, uint32_t ns,
                                                                      float norm, Act act,
                                                                      uint32_t* __restrict__ step, uint32_t bump) {
    __shared__ bf16x8 Wf[kFwdPack / 8];
    __shared__ float bias[4 * kHid + kOutPad];
    __shared__ float stage_all[kTrainWaves][64 * kStage];
    copy_to_lds(reinterpret_cast<uint4*>(Wf), reinterpret_cast<const uint4*>(gv.wf), kFwdPack / 8);
    load_biases(bias, gv.params);
    if (bump && blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1u;  // Adam step of this update
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id();
    float* stage = stage_all[wave];
    const float inv_ns = norm;
    constexpr uint32_t kWs = 16 * kTrainTiles;  // samples per wave
    const uint32_t wstride = gridDim.x * kTrainWaves * kWs;
    for (uint32_t col0 = (blockIdx.x * kTrainWaves + wave) * kWs; col0 < ns; col0 += wstride) {
        const uint32_t k = col0 + (l & (kWs - 1));  // lanes >= kWs repeat a sample and store nothing
        const float4 a = in0[k], b = in1[k], dv = dir[k];
        const float s[kScalars] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z};
        uint32_t mask[4][2];
        mlp_forward<true, kTrainTiles>(Wf, bias, s, stage, mask, act, col0);
        wave_sync();
        if (l < (int)kWs) {
            float o[kOut];
#pragma unroll
            for (int j = 0; j < kOut; ++j) o[j] = stage[l * kStage + j];
            Mixture m;
            mixture_from(o, m);
            const f3 w = mk3(dv.x, dv.y, dv.z);
            float v[kLobes], cth[kLobes], p = 0.0f;
#pragma unroll
            for (int j = 0; j < kLobes; ++j) {
                cth[j] = dot(m.mu[j], w);
                v[j] = m.norm[j] * expf(m.kappa[j] * (cth[j] - 1.0f));
                p += m.pi[j] * v[j];
            }
            const float wt = b.w * inv_ns;
            const float ip = wt != 0.0f ? 1.0f / fmaxf(p, 1e-30f) : 0.0f;
            const uint32_t col = col0 + l;
#pragma unroll
            for (int j = 0; j < kLobes; ++j) {
                const float r = m.pi[j] * v[j] * ip;  // responsibility (0 when the weight is 0)
                const float kap = m.kappa[j];
                const float dlk = 1.0f / kap - 2.0f / expm1f(2.0f * kap) + (cth[j] - 1.0f);
                const bool in_range = m.s[j] > kLogKappaMin && m.s[j] < kLogKappaMax;
                const float gs = in_range ? -wt * r * kap * dlk * kKappaScale : 0.0f;
                const float gm = -wt * r * kap / m.len[j];
                const f3 perp = w - m.mu[j] * cth[j];
                act.put(kD5, 5 * j + 0, col, gm * perp.x);
                act.put(kD5, 5 * j + 1, col, gm * perp.y);
                act.put(kD5, 5 * j + 2, col, gm * perp.z);
                act.put(kD5, 5 * j + 3, col, gs);
                act.put(kD5, 5 * j + 4, col, kLogitScale * wt * (m.pi[j] - r));
            }
#pragma unroll
            for (int row = kOut; row < kOutPad; ++row) act.put(kD5, row, col, 0.0f);
        }
        wave_sync();
    }
}

// Backward pass of one training step: delta_l = (W_{l+1}^T delta_{l+1})
// masked by ReLU' (from the stored activations, h > 0 <=> z > 0), for
// layers 4..1, 64 samples per wave.
__global__ __launch_bounds__(64 * kTrainWaves) void k_guide_train_bwd(GuideView gv, uint32_t ns, Act act) {
    __shared__ bf16x8 Wt[kBwdPack / 8];
    copy_to_lds(reinterpret_cast<uint4*>(Wt), reinterpret_cast<const uint4*>(gv.wt), kBwdPack / 8);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id(), g = l >> 4, n = l & 15;
    constexpr uint32_t kWs = 16 * kBwdTiles;  // samples per wave
    const uint32_t wstride = gridDim.x * kTrainWaves * kWs;
    for (uint32_t col0 = (blockIdx.x * kTrainWaves + wave) * kWs; col0 < ns; col0 += wstride) {
        // output gradient, natural row order, as the B operand (rows >= 48 zero)
        bf16x8 B[2][kBwdTiles];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int t = 0; t < kBwdTiles; ++t)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int row = 32 * kb + 8 * g + j;
                    B[kb][t][j] = row < kOutPad ? act.base[(size_t)(kD5 + row) * ns + col0 + 16 * t + n] : (short)0;
                }
#pragma unroll 1
        for (int layer = 4; layer >= 1; --layer) {
            const int off = layer == 4 ? kT5 : layer == 3 ? kT4 : layer == 2 ? kT3 : kT2;
            f32x4 C[4][kBwdTiles];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int t = 0; t < kBwdTiles; ++t) C[mb][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int mb = 0; mb < 4; ++mb) {
                    const bf16x8 A = Wt[off / 8 + (mb * 2 + kb) * 64 + l];
#pragma unroll
                    for (int t = 0; t < kBwdTiles; ++t) C[mb][t] = mfma(A, B[kb][t], C[mb][t]);
                }
            const uint32_t drow = kD1 + 64 * (layer - 1);
            const uint32_t hrow = kH1 + 64 * (layer - 1);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int t = 0; t < kBwdTiles; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * mb + 4 * g + r;
                        const size_t col = col0 + 16 * t + n;
                        const bool on = act.base[(size_t)(hrow + row) * ns + col] > 0;  // bf16 of h >= 0
                        const float dval = on ? C[mb][t][r] : 0.0f;
                        act.base[(size_t)(drow + row) * ns + col] = bf16_bits(dval);
                        B[mb >> 1][t][(mb & 1) * 4 + r] = bf16_bits(dval);
                    }
        }
    }
}

// Weight and bias gradients of one step: dW_l = sum over samples of
// delta_l (x) input_{l-1}, with samples as the MFMA k dimension. Task
// (chunk, part): part 0/1 = layer 1 input columns 0..95 / 96..191, parts
// 2..5 = layers 2..5. Writes the chunk's partial sums (every parameter).
__global__ __launch_bounds__(64) void k_guide_wgrad(Act act, uint32_t chunk, float* __restrict__ partial) {
    const uint32_t c = blockIdx.x / 6, part = blockIdx.x % 6;
    const int l = __lane_id(), g = l >> 4, i = l & 15;
    const int layer = part < 2 ? 1 : (int)part;
    const uint32_t arow = layer == 1 ? kD1 : layer == 2 ? kD2 : layer == 3 ? kD3 : layer == 4 ? kD4 : kD5;
    const uint32_t brow = layer == 1 ? kH0 + (part == 1 ? 96u : 0u) : kH0 + 192 + 64 * (layer - 2);
    const int MB = layer == 5 ? 3 : 4;
    const int NB = layer == 1 ? 6 : 4;
    const int in_dim = layer == 1 ? kIn : kHid;
    const int wofs = layer == 1 ? kW1 : layer == 2 ? kW2 : layer == 3 ? kW3 : layer == 4 ? kW4 : kW5;
    const int bofs = layer == 1 ? kB1 : layer == 2 ? kB2 : layer == 3 ? kB3 : layer == 4 ? kB4 : kB5;
    const int ncol0 = part == 1 ? 96 : 0;
    f32x4 acc[4][6];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 6; ++nb) acc[mb][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool do_bias = part != 1;
    const uint32_t k_begin = c * chunk, k_end = k_begin + chunk;
    for (uint32_t k0 = k_begin; k0 < k_end; k0 += 32) {
        bf16x8 A[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            if (mb >= MB) continue;
            A[mb] = *reinterpret_cast<const bf16x8*>(act.base + (size_t)(arow + 16 * mb + i) * act.ns + k0 + 8 * g);
            if (do_bias) {
#pragma unroll
                for (int j = 0; j < 8; ++j) bsum[mb] += bf16_float(A[mb][j]);
            }
        }
#pragma unroll
        for (int nb = 0; nb < 6; ++nb) {
            if (nb >= NB) continue;
            const bf16x8 Bv =
                *reinterpret_cast<const bf16x8*>(act.base + (size_t)(brow + 16 * nb + i) * act.ns + k0 + 8 * g);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                if (mb >= MB) continue;
                acc[mb][nb] = mfma(A[mb], Bv, acc[mb][nb]);
            }
        }
    }
    float* out = partial + (size_t)c * kParams;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        if (mb >= MB) continue;
#pragma unroll
        for (int nb = 0; nb < 6; ++nb) {
            if (nb >= NB) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = 16 * mb + 4 * g + r;
                out[wofs + o * in_dim + ncol0 + 16 * nb + i] = acc[mb][nb][r];
            }
        }
        if (do_bias) {
            float v = bsum[mb];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (g == 0) out[bofs + 16 * mb + i] = v;
        }
    }
}

// Whether parameter i is padding (kept at zero).
__device__ __forceinline__ bool is_padding(int i) {
    if (i >= kW1 && i < kB1) return (i - kW1) % kIn >= kFeat;
    if (i >= kW5 && i < kB5) return (i - kW5) / kHid >= kOut;
    if (i >= kB5) return i - kB5 >= kOut;
    return false;
}

// physical k position of logical row q (0..31) in the chain order
__device__ __forceinline__ int chain_pos(int q) {
    return q < 16 ? 8 * (q >> 2) + (q & 3) : 8 * ((q - 16) >> 2) + 4 + ((q - 16) & 3);
}

// Sums the partial gradients, applies Adam (update != 0) and writes the
// packed BF16 operands of every weight.
// Exponential average of the parameters the tracer samples with (Polyak
// averaging): the Adam iterate itself moves by lr-sized steps on noisy
// gradients, and one network taken at a given step can be markedly worse
// than the ones a few steps either side of it (on the occluded room, C of
// 128-171 over 8 trainings with the iterate, 127-136 with the average at
// 0.99, mean 128.3 at 0.995 and 128.2 at 0.999 over 4). kAvgBeta 0.995
// averages over about 200 Adam steps (12 frames of 16 steps); early steps use
// (1 + t) / (10 + t) instead, so the average follows the first updates.
constexpr float kAvgBeta = 0.995f;

__global__ __launch_bounds__(256) void k_guide_adam(const float* __restrict__ partial, uint32_t nchunks,
                                                    float* __restrict__ params, float* __restrict__ m1,
                                                    float* __restrict__ m2, const uint32_t* __restrict__ step,
                                                    float lr, uint32_t update, short* __restrict__ wf,
                                                    short* __restrict__ wt, float* __restrict__ avg,
                                                    short* __restrict__ wf_avg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float w = params[i];
    float wa = 0.0f;
    if (update && !is_padding(i)) {
        float gsum = 0.0f;
#pragma unroll 8
        for (uint32_t c = 0; c < nchunks; ++c) gsum += partial[(size_t)c * kParams + i];
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float t = (float)step[0];
        const float m = b1 * m1[i] + (1.0f - b1) * gsum;
        const float v = b2 * m2[i] + (1.0f - b2) * gsum * gsum;
        m1[i] = m;
        m2[i] = v;
        const float mh = m / (1.0f - powf(b1, t)), vh = v / (1.0f - powf(b2, t));
        w -= lr * mh / (sqrtf(vh) + eps);
        params[i] = w;
        if (avg) {
            const float beta = fminf(kAvgBeta, (1.0f + t) / (10.0f + t));
            wa = beta * avg[i] + (1.0f - beta) * w;
            avg[i] = wa;
        }
    }
    int layer, o, c, in_dim;
    if (i < kB1) { layer = 1; in_dim = kIn; o = (i - kW1) / kIn; c = (i - kW1) % kIn; }
    else if (i >= kW2 && i < kB2) { layer = 2; in_dim = kHid; o = (i - kW2) / kHid; c = (i - kW2) % kHid; }
    else if (i >= kW3 && i < kB3) { layer = 3; in_dim = kHid; o = (i - kW3) / kHid; c = (i - kW3) % kHid; }
    else if (i >= kW4 && i < kB4) { layer = 4; in_dim = kHid; o = (i - kW4) / kHid; c = (i - kW4) % kHid; }
    else if (i >= kW5 && i < kB5) { layer = 5; in_dim = kHid; o = (i - kW5) / kHid; c = (i - kW5) % kHid; }
    else return;  // biases are read as fp32
    (void)in_dim;
    const short h = bf16_bits(w);
    const bool pack_avg = update && avg && wf_avg && !is_padding(i);
    const short ha = pack_avg ? bf16_bits(wa) : (short)0;
    // forward operand A = W_l: row o, k = c (identity order for layer 1)
    {
        const int kb = c >> 5, q = c & 31;
        const int pphys = layer == 1 ? q : chain_pos(q);
        const int KB = layer == 1 ? 6 : 2;
        const int base = layer == 1 ? kF1 : layer == 2 ? kF2 : layer == 3 ? kF3 : layer == 4 ? kF4 : kF5;
        const int lane = 16 * (pphys >> 3) + (o & 15);
        const int at = base + (((o >> 4) * KB + kb) * 64 + lane) * 8 + (pphys & 7);
        wf[at] = h;
        if (pack_avg) wf_avg[at] = ha;
    }
    // backward operand A = W_l^T: row c, k = o (layers 2..5)
    if (layer >= 2 && wt) {
        const int kb = o >> 5, q = o & 31;
        const int pphys = layer == 5 ? q : chain_pos(q);
        const int base = layer == 5 ? kT5 : layer == 4 ? kT4 : layer == 3 ? kT3 : kT2;
        const int lane = 16 * (pphys >> 3) + (c & 15);
        wt[base + (((c >> 4) * 2 + kb) * 64 + lane) * 8 + (pphys & 7)] = h;
    }
}

// Network outputs (40 per sample) for given input scalars.
__global__ __launch_bounds__(64 * kScatterWaves) void k_guide_infer(GuideView gv, const float4* __restrict__ in0,
                                                                    const float4* __restrict__ in1, uint32_t n,
                                                                    float* __restrict__ out) {
    __shared__ bf16x8 Wf[kFwdPack / 8];
    __shared__ float bias[4 * kHid + kOutPad];
    __shared__ float stage_all[kScatterWaves][64 * kStage];
    copy_to_lds(reinterpret_cast<uint4*>(Wf), reinterpret_cast<const uint4*>(gv.wf), kFwdPack / 8);
    load_biases(bias, gv.params);
    __syncthreads();
    const int wave = threadIdx.x >> 6, l = __lane_id();
    float* stage = stage_all[wave];
    const uint32_t wstride = gridDim.x * kScatterWaves * 64;
    for (uint32_t base = (blockIdx.x * kScatterWaves + wave) * 64; base < n; base += wstride) {
        const uint32_t i = base + l;
        float4 a = splat4(0.0f), b = splat4(0.0f);
        if (i < n) {
            a = in0[i];
            b = in1[i];
        }
        const float s[kScalars] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z};
        uint32_t mask[4][2];
        const Act none{nullptr, 0};
        mlp_forward<false, 4>(Wf, bias, s, stage, mask, none, 0);
        wave_sync();
        if (i < n)
            for (int k = 0; k < kOut; ++k) out[(size_t)i * kOut + k] = stage[l * kStage + k];
        wave_sync();
    }
}

}  // namespace

size_t guide_act_words(uint32_t ns) { return (size_t)kActRows * ns / 2; }  // in uint32 words

void launch_guide_scatter(const SceneView& sv, const PathState& ps, const Queues& qs, const GuideView& gv,
                          const FrameArgs& fa, uint32_t bounce, uint32_t* next_queue, Counters* ctr,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_guide_scatter, dim3(256), dim3(64 * kScatterBlockWaves), 0, s, sv, ps, qs, gv, fa, bounce,
                       next_queue, ctr);
}

void launch_guide_pack(const GuideView& gv, float* params, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, nullptr, 0u, params, nullptr,
                       nullptr, nullptr, 0.0f, 0u, const_cast<short*>(reinterpret_cast<const short*>(gv.wf)),
                       const_cast<short*>(reinterpret_cast<const short*>(gv.wt)), nullptr, nullptr);
}

void launch_guide_train_step(const GuideView& gv, const GuideTrain& gt, const float4* in0, const float4* in1,
                             const float4* dir, uint32_t ns, float norm, bool update, hipStream_t s) {
    const Act act{reinterpret_cast<short*>(gt.act), ns};
    const uint32_t per_block = 16 * kTrainTiles * kTrainWaves;
    const uint32_t blocks = std::min<uint32_t>(512u, (ns + per_block - 1) / per_block);
    hipLaunchKernelGGL(k_guide_train_fwd, dim3(blocks), dim3(64 * kTrainWaves), 0, s, gv, in0, in1, dir, ns, norm,
                       act, gt.step, update ? 1u : 0u);
    const uint32_t bper = 16 * kBwdTiles * kTrainWaves;
    const uint32_t bblocks = std::min<uint32_t>(1024u, (ns + bper - 1) / bper);
    hipLaunchKernelGGL(k_guide_train_bwd, dim3(bblocks), dim3(64 * kTrainWaves), 0, s, gv, ns, act);
    const uint32_t nchunks = ns / gt.chunk;
    hipLaunchKernelGGL(k_guide_wgrad, dim3(6 * nchunks), dim3(64), 0, s, act, gt.chunk, gt.partial);
    if (update)
        hipLaunchKernelGGL(k_guide_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, gt.partial, nchunks,
                           gt.params, gt.m1, gt.m2, gt.step, gt.lr, 1u,
                           const_cast<short*>(reinterpret_cast<const short*>(gv.wf)),
                           const_cast<short*>(reinterpret_cast<const short*>(gv.wt)), gt.avg,
                           reinterpret_cast<short*>(gt.wf_avg));
}

void launch_guide_prepare(const PathState& ps, const GuideView& gv, const FrameArgs& fa, uint32_t npx,
                          uint32_t slot, uint32_t ns, uint32_t stride, float4* in0, float4* in1, float4* dir,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_guide_prepare, dim3((ns + 255) / 256), dim3(256), 0, s, ps, gv, fa, npx, slot, ns,
                       stride, in0, in1, dir);
}

void launch_guide_infer(const GuideView& gv, const float4* in0, const float4* in1, uint32_t n, float* out,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_guide_infer, dim3(256), dim3(64 * kScatterWaves), 0, s, gv, in0, in1, n, out);
}

}  // namespace wsrt
