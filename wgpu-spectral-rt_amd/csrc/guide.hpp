// guide.hpp - the path-guiding network: layout of its parameters, of the
// packed BF16 operands of its MFMA kernels, the input encoding and the
// von Mises-Fisher mixture it outputs.
//
// Network: 7 input scalars in [0, 1] (hit position over the scene box, the
// forward-facing shading normal mapped from [-1, 1], the hero wavelength
// over 380..780 nm), each encoded by 12 frequency bands as sin and cos of
// 2^b * pi * x (168 features, padded to 192), four hidden layers of 64 ReLU
// units, and 40 outputs: 8 lobes of (mu_x, mu_y, mu_z, s, a), with
// kappa = exp(clamp(8 s, -3, 7)) and logit = 4 a. The guide density is
// sum_k softmax(logit)_k * vMF(w; mu_k / |mu_k|, kappa_k).
//
// Feature order (the column order of W1): column c = 8 q + j with slot
// q = 3 a + bq (scalar a, band quad bq = 0..2) and j = 2 (band % 4) + (cos ? 1 : 0),
// band = 4 bq + j / 2. Slots 21..23 (columns 168..191) are zero padding.
//
// MFMA operands (v_mfma_f32_16x16x32_bf16, one wave): A is 16 x 32 with lane
// l holding A[l % 16][8 (l / 16) + j], B is 32 x 16 with lane l holding
// B[8 (l / 16) + j][l % 16], and the 16 x 16 result has lane l holding
// D[4 (l / 16) + r][l % 16], j = 0..7, r = 0..3. Activations are kept with
// features as rows and samples as columns, so a layer's result feeds the next
// layer's B operand without moving data when the next weight matrix's columns
// are stored in the "chain" order: physical k position 8 g + j of a 32-wide
// k block holds logical row 4 g + j (j < 4) or 16 + 4 g + (j - 4) (j >= 4)
// of the two 16-row result blocks that make up the k block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace wsrt {
namespace guide {

constexpr int kScalars = 7, kBands = 12, kFeat = 168, kIn = 192, kHid = 64, kOut = 40, kOutPad = 48;
constexpr int kLobes = 8;

// fp32 parameters (master copy, Adam moments and gradients share the layout);
// matrices are row-major [out][in]
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHid * kIn;
constexpr int kW2 = kB1 + kHid;
constexpr int kB2 = kW2 + kHid * kHid;
constexpr int kW3 = kB2 + kHid;
constexpr int kB3 = kW3 + kHid * kHid;
constexpr int kW4 = kB3 + kHid;
constexpr int kB4 = kW4 + kHid * kHid;
constexpr int kW5 = kB4 + kHid;
constexpr int kB5 = kW5 + kOutPad * kHid;
constexpr int kParams = kB5 + kOutPad;  // 27952

// packed BF16 operand blocks of 64 lanes x 8 values
constexpr int kBlk = 512;
// forward: A = W_l (rows = outputs, k = inputs), [mb][kb] blocks
constexpr int kF1 = 0;                    // 4 x 6 blocks, identity k order
constexpr int kF2 = kF1 + 24 * kBlk;      // 4 x 2 blocks, chain k order
constexpr int kF3 = kF2 + 8 * kBlk;
constexpr int kF4 = kF3 + 8 * kBlk;
constexpr int kF5 = kF4 + 8 * kBlk;       // 3 x 2 blocks, chain k order
constexpr int kFwdPack = kF5 + 6 * kBlk;  // 27648 values
// backward: A = W_l^T (rows = inputs of layer l, k = outputs of layer l)
constexpr int kT5 = 0;                    // 4 x 2 blocks, natural k order, k >= 48 zero
constexpr int kT4 = kT5 + 8 * kBlk;       // 4 x 2 blocks, chain k order
constexpr int kT3 = kT4 + 8 * kBlk;
constexpr int kT2 = kT3 + 8 * kBlk;
constexpr int kBwdPack = kT2 + 8 * kBlk;  // 16384 values

// logical index of physical k position p (0..31) in the chain order
__host__ __device__ inline int chain_row(int p) {
    const int g = p >> 3, j = p & 7;
    return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4);
}

// Feature of encoded column c for the scalars s[7].
struct FeatureSlot {
    int scalar, band, is_cos;  // scalar < 0: padding
};
__host__ __device__ inline FeatureSlot feature_slot(int c) {
    FeatureSlot f;
    const int q = c >> 3, j = c & 7;
    if (q >= 3 * kScalars) {
        f.scalar = -1;
        f.band = 0;
        f.is_cos = 0;
        return f;
    }
    f.scalar = q / 3;
    f.band = 4 * (q % 3) + (j >> 1);
    f.is_cos = j & 1;
    return f;
}

// Output scaling: kappa = exp(clamp(kKappaScale * s, kLogKappaMin, kLogKappaMax))
// and lobe logit = kLogitScale * raw, so that Adam steps of size lr move
// log kappa and the logits kKappaScale and kLogitScale times faster.
constexpr float kLogKappaMin = -3.0f, kLogKappaMax = 7.0f;
constexpr float kKappaScale = 8.0f, kLogitScale = 4.0f;

// Training records of the Lambert vertices at bounces 0 .. kRecBounces-1 of
// each path (structure of arrays, index bounce * capacity + path id).
constexpr uint32_t kRecBounces = kGuideRecordBounces;
// Unsigned 32-bit division by a divisor fixed at set-up: q = (hi32(m n) + n)
// >> l in 64 bits, with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1
// (Granlund and Montgomery 1994, round-up multiplier); exact for every n.
struct FastDiv {
    uint32_t m, l;
};
inline FastDiv fast_div(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) ++l;
    return FastDiv{(uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1), l};
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
    return (uint32_t)(((uint64_t)__umulhi(f.m, n) + n) >> f.l);
}

// Training records: only the vertices a training step will read are kept.
// Vertex (bounce b, pixel x) of frame slot s is record k of the slot when
// j = b * npx + x satisfies j = k * stride + phase with k < ns and phase =
// (frame of the slot) mod stride; it is stored at s * ns + k.
struct Records {
    float4* in0;   // scalars 0..3
    float4* in1;   // scalars 4..6, pdf of the mixture the direction came from
    float4* dir;   // sampled direction, cos to the normal
    float4* tp;    // throughput after the scatter (before roulette)
    float4* Lb;    // radiance accumulated before the scatter
    float4* nee;   // the vertex's light-sample contribution (added if unoccluded)
    uint32_t npx;     // pixels of a frame
    uint32_t ns;      // records (training samples) per frame slot
    uint32_t stride;  // vertices between consecutive records
    FastDiv div_npx, div_stride;
};

}  // namespace guide
}  // namespace wsrt
