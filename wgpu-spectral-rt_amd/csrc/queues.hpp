// queues.hpp - work queues of the wavefront kernels without hot global
// atomics.
//
// A device-scope atomic on one address from every wave of a 2M-item launch
// serializes in the memory system (~10 ns each on MI355X, where the eight
// XCDs do not share an L2), which made queue appends the dominant cost. A
// queue is therefore split into kSegs segments: producer block b appends to
// segment b % kSegs (one wave-aggregated atomic on a counter that sits on
// its own 128-byte line), and a consumer turns its item index into (segment,
// offset) through a prefix sum of the segment counts held in LDS. A
// producer of L <= kGridLanes lanes (blocks of whole waves, L / kSegs lanes
// per segment) gives each lane at most ceil(n / L) of its n items, grid-stride
// or as a wave's contiguous range, so segment s receives at most
// (L / kSegs) * (n / L + 1) = n / kSegs + L / kSegs items: a segment of
// seg_cap = ceil(capacity / kSegs) + kGridLanes / kSegs entries never
// overflows, whatever grid the producer runs. Items enter a queue in an order that depends on scheduling;
// every consumer treats items independently, so results do not depend on it.
#pragma once

#include "common.hpp"
#include "dev_math.hpp"

namespace wsrt {

constexpr int kSegs = 64;
constexpr int kSegStride = 32;             // uint32 counters 128 bytes apart
constexpr uint32_t kGridLanes = 524288;     // most lanes of a queue-producing launch

enum QueueKind : uint32_t {
    kQActive = 0,                      // camera paths entering a bounce
    kQShade0 = 1,                      // hits of a bounce, one queue per material type
    kQShadow = kQShade0 + kMatBins,    // shadow rays of a bounce
    kQPhoton = kQShadow + 1,           // photon paths entering a bounce
    kQGuide = kQPhoton + 1,            // Lambert vertices waiting for guided scattering
    kQKinds
};

struct Queues {
    uint32_t* counts;  // [depth][kind][segment] counters, kSegStride apart
    uint32_t seg_cap;  // camera queues: entries per segment
    uint32_t pseg_cap; // photon queues: entries per segment
    // camera bounce + 1 from which one persistent launch finishes the
    // remaining paths (0: not yet); null where the tail launch is not used
    uint32_t* tail = nullptr;
    // once fewer camera paths than this enter a bounce (from bounce
    // kTailFirstBounce on), the wavefront kernels of that and later bounces
    // leave them to the tail launch: a launch per bounce for a few thousand
    // paths costs its launch and latency, not its work
    uint32_t tail_paths = 0;
    // traversal launches (extend, shadow, tail, photon bounce) use this many
    // rounds of resident waves (wsrt_config.trace_rounds; results do not
    // depend on it)
    uint32_t trace_rounds = 1;
};

constexpr uint32_t kTailFirstBounce = 2;

// Whether the tail launch has taken over (uniform across a launch).
__device__ __forceinline__ bool tail_started(const Queues& q, uint32_t bounce) {
    return q.tail && bounce >= kTailFirstBounce && *q.tail != 0u;
}

__host__ __device__ inline uint32_t seg_capacity(uint64_t capacity) {
    return (uint32_t)((capacity + kSegs - 1) / kSegs + kGridLanes / kSegs);
}
__host__ __device__ inline size_t queue_count_words(uint32_t depths) {
    return (size_t)depths * kQKinds * kSegs * kSegStride;
}

__device__ __forceinline__ uint32_t* qcount(const Queues& q, uint32_t depth, uint32_t kind,
                                            uint32_t seg) {
    return q.counts + (((size_t)depth * kQKinds + kind) * kSegs + seg) * kSegStride;
}

// Appends from every lane with pred set; returns the lane's entry index in
// the queue's array (segment * cap + slot) or kInvalid. All lanes of the
// wave must call it.
__device__ __forceinline__ uint32_t seg_append(const Queues& q, uint32_t depth, uint32_t kind,
                                               uint32_t cap, bool pred) {
    const uint32_t seg = blockIdx.x % kSegs;
    const uint32_t slot = wave_append(qcount(q, depth, kind, seg), pred);
    return pred ? seg * cap + slot : kInvalid;
}

// Exclusive prefix of the segment counts of kinds [k0, k0 + nk) into pre
// (LDS, nk * kSegs + 1 entries, nk <= 8); returns the total. Block-wide:
// all threads must call it.
__device__ __forceinline__ uint32_t seg_prefix(const Queues& q, uint32_t depth, uint32_t k0,
                                               uint32_t nk, uint32_t* pre) {
    const uint32_t ne = nk * kSegs;
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x, per = ne / 64;
        uint32_t v[8];
        uint32_t s = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            v[k] = 0;
            if (k < per) {
                const uint32_t e = lane * per + k;
                v[k] = *qcount(q, depth, k0 + e / kSegs, e % kSegs);
                s += v[k];
            }
        }
        uint32_t inc = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t t = __shfl_up(inc, off, 64);
            if ((int)lane >= off) inc += t;
        }
        uint32_t ex = inc - s;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k)
            if (k < per) {
                pre[lane * per + k] = ex;
                ex += v[k];
            }
        if (lane == 63) pre[ne] = inc;
    }
    __syncthreads();
    return pre[ne];
}

// Entry of item i (< total): segment j with pre[j] <= i < pre[j + 1].
__device__ __forceinline__ uint32_t seg_entry(const uint32_t* pre, uint32_t ne, uint32_t i,
                                              uint32_t cap, uint32_t* seg_out = nullptr) {
    uint32_t lo = 0, hi = ne;  // invariant: pre[lo] <= i < pre[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= i) lo = mid;
        else hi = mid;
    }
    if (seg_out) *seg_out = lo;
    return lo * cap + (i - pre[lo]);
}

// Segment holding item base of a queue of kSegs segments (pre: kSegs + 1
// prefix entries): lane l reads pre[l], and the segments starting at or
// before base form a prefix of the wave's ballot. All 64 lanes must call it
// with the same base.
__device__ __forceinline__ uint32_t seg_of_wave_base(const uint32_t* pre, uint32_t base) {
    const uint64_t m = __ballot(pre[__lane_id()] <= base);
    return (uint32_t)__popcll(m) - 1u;
}
// Entry of item i (base <= i < total) from the segment s0 of base: steps
// over the segments that end at or before i, usually none or one.
__device__ __forceinline__ uint32_t seg_entry_from(const uint32_t* pre, uint32_t s0, uint32_t i, uint32_t cap) {
    uint32_t s = s0;
    while (s + 1u < (uint32_t)kSegs && pre[s + 1u] <= i) ++s;
    return s * cap + (i - pre[s]);
}

// Block sum of one value per thread (blockDim a multiple of 64, <= 1024),
// valid in thread 0.
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long* lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (uint32_t k = 0; k < nw; ++k) t += lds[k];
    __syncthreads();
    return t;
}

}  // namespace wsrt
