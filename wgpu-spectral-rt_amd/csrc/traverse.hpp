// traverse.hpp - software ray queries over the compressed BVH8 (CDNA4 has no
// ray-tracing units). Replaces rayQueryProceed of shader.wgsl:561-566,654-659
// and photon_emit.wgsl:222-227.
//
// Semantics (kept identical in the CPU oracle):
//  * closest hit on (tmin, tmax), ties broken by the lower world triangle id;
//  * triangle test Moller-Trumbore, no culling; front face <=> det > 0, i.e.
//    dot(dir, cross(v1 - v0, v2 - v0)) < 0 (the DXR/Vulkan default winding);
//  * any hit terminates on the first accepted triangle.
// Node boxes are tested with fused slab values plus a bound of their f32
// error (folded into the far planes), so culling never rejects a box the
// exact box would accept; the
// visiting order (leaf children first, then internal children in the order
// of the ray's direction octant over the build-time octant slots) is what
// the oracle's BVH8 traversal counts.
#pragma once

#include "common.hpp"
#include "dev_math.hpp"

namespace wsrt {

constexpr float kBoxPad = 1.00001f;
constexpr int kShortStack = 8;  // entries per lane in LDS; deeper ones spill to HBM

struct Hit {
    float t;
    uint32_t tri;  // kInvalid on miss
    float u, v;
    bool front;
};

__device__ __forceinline__ float safe_inv(float d) {
    if (fabsf(d) < 1e-20f) d = d < 0.0f ? -1e-20f : 1e-20f;
    return 1.0f / d;
}

// Cross and dot products of the triangle test with fused multiply-adds:
// each component is one product plus one fused product (cross) or one
// product and two fused ones (dot), each rounded once. The oracle's triangle
// test uses the same fused forms in the same order.
__device__ __forceinline__ f3 cross_fma(f3 a, f3 b) {
    return mk3(__fmaf_rn(a.y, b.z, -(a.z * b.y)), __fmaf_rn(a.z, b.x, -(a.x * b.z)),
               __fmaf_rn(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float dot_fma(f3 a, f3 b) { return __fmaf_rn(a.x, b.x, __fmaf_rn(a.y, b.y, a.z * b.z)); }

// Moller-Trumbore with the barycentric tests on the unscaled numerators
// (signs normalised by det's), so only a triangle that passes them divides.
__device__ __forceinline__ bool tri_intersect(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float& t, float& u,
                                              float& v, bool& front) {
    f3 pvec = cross_fma(d, e2);
    float det = dot_fma(e1, pvec);
    if (det == 0.0f) return false;
    const float sg = det > 0.0f ? 1.0f : -1.0f, adet = det * sg;
    f3 tvec = o - v0;
    float U = dot_fma(tvec, pvec);
    if (U * sg < 0.0f || U * sg > adet) return false;
    f3 qvec = cross_fma(tvec, e1);
    float V = dot_fma(d, qvec);
    if (V * sg < 0.0f || (U + V) * sg > adet) return false;
    float inv = 1.0f / det;
    t = dot_fma(e2, qvec) * inv;
    u = U * inv;
    v = V * inv;
    front = det > 0.0f;
    return true;
}

// Traversal stack: entries (node, tnear bits). The first kShortStack entries
// of a lane live in LDS at lds[k * lds_stride]; deeper entries spill to the
// lane's private HBM region gstack[0 .. stack_size - kShortStack). The LDS
// part is addressed as LDS (address space 3), so accesses compile to ds
// instructions rather than flat ones that would wait on both counters.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x2 lds_entry;
struct Stack {
    lds_entry* lds;
    uint32_t lds_stride;
    uint2* glob;
    __device__ __forceinline__ void put(int k, uint2 e) {
        if (k < kShortStack) {
            u32x2 v;
            v.x = e.x;
            v.y = e.y;
            lds[k * lds_stride] = v;
        } else {
            glob[k - kShortStack] = e;
        }
    }
    __device__ __forceinline__ uint2 get(int k) {
        if (k < kShortStack) {
            const u32x2 v = lds[k * lds_stride];
            return make_uint2(v.x, v.y);
        }
        return glob[k - kShortStack];
    }
};
__device__ __forceinline__ lds_entry* lds_slot(uint2* shared_base) {
    return (lds_entry*)(shared_base);
}

__device__ __forceinline__ uint32_t byte_of(uint32_t lo, uint32_t hi, int c) {
    return ((c < 4 ? lo : hi) >> (8 * (c & 3))) & 0xFFu;
}

// Traversal state of one ray between node visits.
struct TravState {
    f3 o, d;
    float inv[3];
    float tmin;
    int oct;          // direction octant: bit a set when d[a] < 0
    int sp;           // stack entries
    uint32_t next;    // node to visit next when have_next
    bool have_next;
    Hit best;
    uint32_t nvis, ntest;
};

__device__ __forceinline__ void trav_init(const SceneView& sv, TravState& s, f3 o, f3 d, float tmin, float tmax) {
    s.best.t = tmax;
    s.best.tri = kInvalid;
    s.best.u = s.best.v = 0.0f;
    s.best.front = false;
    s.o = o;
    s.d = d;
    s.inv[0] = safe_inv(d.x);
    s.inv[1] = safe_inv(d.y);
    s.inv[2] = safe_inv(d.z);
    s.tmin = tmin;
    s.oct = (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0);
    s.sp = 0;
    // the node to visit next is kept in a register: the root first, then the
    // first internal child in visiting order of each node (it would be
    // pushed and popped at once, and with best.t unchanged in between its
    // pop-time cull test could not fail)
    s.next = 0u;
    // a ray with a NaN in its origin or direction (a NaN shading normal
    // gives one) meets no triangle: its Moller-Trumbore values are NaN and
    // fail every test, while its slab values (NaN, dropped by fmin/fmax)
    // would open every box. It is not traversed.
    const float sum = o.x + o.y + o.z + d.x + d.y + d.z;
    s.have_next = sv.n_nodes != 0u && sum == sum;
    s.nvis = s.ntest = 0;
}

// Tests triangle `tri` against the lane's ray and keeps the closest hit
// (ties to the lower id); returns true when an any-hit query has its hit.
template <bool kAny, bool kCount>
__device__ __forceinline__ bool test_tri(const SceneView& sv, TravState& s, uint32_t tri) {
    const float4* tr = sv.tris + 3ull * tri;
    const float4 a = tr[0], b = tr[1], cc = tr[2];
    // the three loads complete together (one round trip to memory per
    // triangle): without this the vertex load is issued only after the
    // determinant test, a second dependent round trip
    asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z), "v"(cc.x), "v"(cc.y),
                 "v"(cc.z));
    if (kCount) s.ntest++;
    float tt, u, v;
    bool fr;
    if (!tri_intersect(s.o, s.d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), tt, u, v, fr))
        return false;
    if (!(tt > s.tmin)) return false;
    const uint32_t id = __float_as_uint(a.w);
    if (tt < s.best.t || (tt == s.best.t && s.best.tri != kInvalid && id < s.best.tri)) {
        s.best.t = tt;
        s.best.tri = id;
        s.best.u = u;
        s.best.v = v;
        s.best.front = fr;
        return kAny;
    }
    return false;
}

// The node a traversal step visits: what its box tests found and the words
// its children are located from.
struct NodeVisit {
    uint32_t meta_lo, meta_hi;  // per-slot meta bytes: slots 0-3, 4-7
    uint32_t child_base, tri_base;
    uint32_t hitmask;  // bit c: child c's box is hit
    uint32_t imask;    // bit c: child c is internal
    float tnear[8];
};

// Loads node s.next and tests its 8 child boxes against the lane's ray.
__device__ __forceinline__ void visit_node(const SceneView& sv, TravState& s, NodeVisit& nv) {
    s.have_next = false;
    const uint4* np = reinterpret_cast<const uint4*>(sv.nodes) + 5ull * s.next;
    uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
    // per axis: slab scale/offset, and the byte plane of each child's near
    // and far bound (qlo or qhi by the sign of the direction)
    float as[3], ao[3], an[3], fs[3], fo[3];
    const float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    const float oo[3] = {s.o.x, s.o.y, s.o.z};
    // quantized bytes: qlo x/y/z = (w2.xy, w2.zw, w3.xy), qhi x/y/z = (w3.zw, w4.xy, w4.zw)
    const uint32_t lo_l[3] = {w2.x, w2.z, w3.x}, lo_h[3] = {w2.y, w2.w, w3.y};
    const uint32_t hi_l[3] = {w3.z, w4.x, w4.z}, hi_h[3] = {w3.w, w4.y, w4.w};
    uint32_t nl[3], nh[3], fl[3], fh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool neg = (s.oct >> a) & 1;
        const int ex = (int)(int8_t)((w0.w >> (8 * a)) & 0xFFu);
        as[a] = ldexpf(s.inv[a], ex);  // 2^ex * inv, exact
        ao[a] = (org[a] - oo[a]) * s.inv[a];
        // bound of the f32 error of this axis' slab values, taken off its
        // near planes and added to its far planes (with the padding folded
        // into the far planes' constants): each axis keeps its own bound, so
        // a near-zero direction component (inv ~ 1e20) widens only its own
        // slab, which is then all or nothing, and not the others
        const float e = (fabsf(as[a]) * 255.0f + fabsf(ao[a])) * 4.0e-7f;
        an[a] = ao[a] - e;
        fs[a] = as[a] * kBoxPad;
        fo[a] = __fmaf_rn(ao[a], kBoxPad, e * (1.0f + kBoxPad));
        nl[a] = neg ? hi_l[a] : lo_l[a];
        nh[a] = neg ? hi_h[a] : lo_h[a];
        fl[a] = neg ? lo_l[a] : hi_l[a];
        fh[a] = neg ? lo_h[a] : hi_h[a];
    }
    // a child is hit when its near distance (a lower bound) is at most its
    // padded far distance (an upper bound) and the padded closest hit
    const float tcap = s.best.t * kBoxPad;
    uint32_t hitmask = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float tn = s.tmin, tf = tcap;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            tn = fmaxf(tn, __fmaf_rn((float)byte_of(nl[a], nh[a], c), as[a], an[a]));
            tf = fminf(tf, __fmaf_rn((float)byte_of(fl[a], fh[a], c), fs[a], fo[a]));
        }
        nv.tnear[c] = tn;
        // an empty slot (meta 0) is a leaf of no triangles; its box is
        // written inverted, so it is not hit in practice either
        if (tn <= tf) hitmask |= 1u << c;
    }
    nv.hitmask = hitmask;
    // slot classes: bit c of the node's imask byte is set when child c is
    // internal (meta bit 7), stored by the builder
    nv.imask = w0.w >> 24;
    nv.child_base = w1.x;
    nv.tri_base = w1.y;
    nv.meta_lo = w1.z;
    nv.meta_hi = w1.w;
}

// Leaf children: a leaf holds at most 3 triangles, and the builder stores a
// node's leaf triangles contiguously from tri_base (at most 24), so the hit
// leaves' triangles form one mask over offsets 0..23.
__device__ __forceinline__ uint32_t leaf_tri_mask(const NodeVisit& nv, uint32_t leaves) {
    uint32_t tmask = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t meta = byte_of(nv.meta_lo, nv.meta_hi, c);
        const uint32_t bits = ((1u << (meta >> 5)) - 1u) << (meta & 0x1Fu);
        tmask |= bits & (0u - ((leaves >> c) & 1u));
    }
    return tmask;
}

// Pushes the visited node's hit internal children, in the order of the ray's
// octant over the slots: slot c is pushed at rank (c ^ oct) from the top, so
// slot oct is popped first; positions come from a mask permuted by oct. A
// child is culled when its near distance (a lower bound) is beyond the
// closest hit so far (padded). The lowest rank goes to the register
// (s.next), the others onto the stack with lower ranks nearer the top.
__device__ __forceinline__ void push_children(TravState& s, Stack st, const NodeVisit& nv) {
    const float cull = s.best.t * kBoxPad;
    uint32_t push = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) push |= (!(nv.tnear[c] > cull) ? 1u : 0u) << c;
    push &= nv.hitmask & nv.imask;
    // ranks: bit c of push at position c ^ oct (swap bits, pairs, nibbles)
    const int oct = s.oct;
    uint32_t pm = push;
    if (oct & 1) pm = ((pm & 0x55u) << 1) | ((pm >> 1) & 0x55u);
    if (oct & 2) pm = ((pm & 0x33u) << 2) | ((pm >> 2) & 0x33u);
    if (oct & 4) pm = ((pm & 0x0Fu) << 4) | ((pm >> 4) & 0x0Fu);
    if (pm) {
        const int first = (__ffs(pm) - 1) ^ oct;
        const uint32_t rest = pm & (pm - 1u);
        const uint32_t stacked = push & ~(1u << first);
        s.next = nv.child_base + (byte_of(nv.meta_lo, nv.meta_hi, first) & 0x7Fu);
        if (s.sp + __popc(rest) <= kShortStack) {
            // every entry lands in the lane's LDS part (the common case)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (!((stacked >> c) & 1u)) continue;
                const int pos = s.sp + __popc(rest >> ((c ^ oct) + 1));
                u32x2 e;
                e.x = nv.child_base + (byte_of(nv.meta_lo, nv.meta_hi, c) & 0x7Fu);
                e.y = __float_as_uint(nv.tnear[c]);
                st.lds[pos * st.lds_stride] = e;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (!((stacked >> c) & 1u)) continue;
                const int pos = s.sp + __popc(rest >> ((c ^ oct) + 1));
                st.put(pos, make_uint2(nv.child_base + (byte_of(nv.meta_lo, nv.meta_hi, c) & 0x7Fu),
                                       __float_as_uint(nv.tnear[c])));
            }
        }
        s.sp += __popc(rest);
        s.have_next = true;
    }
}

// One step of a traversal: pop the next node (skipping culled entries is a
// step of its own) and visit it: box tests of its 8 children, triangle tests
// of its leaf children, the hit internal children onto the stack. Returns
// false once the traversal has ended (stack empty, or kAny and a hit).
template <bool kAny, bool kCount>
__device__ __forceinline__ bool trav_step(const SceneView& sv, TravState& s, Stack st) {
    if (!s.have_next) {
        if (s.sp == 0) return false;
        const uint2 e = st.get(--s.sp);
        if (!kAny && __uint_as_float(e.y) > s.best.t * kBoxPad) return true;
        s.next = e.x;
    }
    NodeVisit nv;
    visit_node(sv, s, nv);
    if (kCount) s.nvis++;
    const uint32_t leaves = nv.hitmask & ~nv.imask;
    if constexpr (kAny) {
        // any hit: leaf by leaf in slot order, ending at the first hit
        uint32_t lv = leaves;
        while (lv) {
            const int c = __ffs(lv) - 1;
            lv &= lv - 1;
            const uint32_t meta = byte_of(nv.meta_lo, nv.meta_hi, c);
            const uint32_t cnt = meta >> 5, off = meta & 0x1Fu;
            for (uint32_t k = 0; k < cnt; ++k) {
                if (test_tri<kAny, kCount>(sv, s, nv.tri_base + off + k)) return false;
            }
        }
    } else {
        // closest hit: a lane tests its hit leaves' triangles in one flat
        // loop, so a wave runs as many iterations as its lane with the most
        // triangles, not, leaf by leaf, the most triangles any lane has in
        // its next leaf. The closest hit (ties to the lower id) does not
        // depend on the order of the tests.
        uint32_t tmask = leaves ? leaf_tri_mask(nv, leaves) : 0u;  // skipped when no lane has a leaf hit
        while (tmask) {
            const int k = __ffs(tmask) - 1;
            tmask &= tmask - 1;
            (void)test_tri<kAny, kCount>(sv, s, nv.tri_base + (uint32_t)k);
        }
    }
    push_children(s, st, nv);
    return true;
}

// Closest-hit triangle tests shared by the whole wave. The tests a lane's
// node visit asks for (its hit leaves' triangles) are spread over every lane
// of the wave, 64 at a time, instead of each lane running its own: with 0.6 to
// 0.7 triangles per node visit on average but up to 24 for one lane, a wave
// otherwise runs 3.4 (caustic scene) to 4.5 (occluded room) test rounds per
// step where one would hold them all.
//  * Every lane's tests get consecutive positions in a wave-wide list, from
//    an exclusive prefix sum of the per-lane counts (a DPP scan).
//  * Per chunk of 64 positions, lanes write (owner lane, offset) entries to
//    LDS; lane j tests entry j on its owner's ray (fetched by lane shuffles)
//    and folds a hit into the owner's key, (ordered t bits << 32 | triangle
//    id), with an LDS atomic minimum; the lane whose key won writes t, u, v
//    and the facing.
//  * Keys order as (t, id), -0 and +0 alike (ordered_bits). The owner's
//    initial key is its closest hit so far, or, with none, the largest key
//    whose t is below tmax, so the final key is the closest hit with ties to
//    the lower id: the result of the lane's own tests in any order, bit for
//    bit.
// Any-hit rays (shadow) keep trav_step: there the lane's leaf-by-leaf loop
// ends at its first hit, and testing all of a node's triangles instead made
// the occluded room's shadow kernel 7 % slower.
// Every lane of the wave calls it, in wave-uniform control flow; lanes with
// no tests help the others.
struct TriShare {
    unsigned long long key[64];
    float t[64], u[64], v[64];
    uint32_t front[64];
    uint16_t item[64];  // one chunk: owner lane | (offset from its tri_base << 6)
};

// A float's bits mapped so that their unsigned order is the float order
// (negative values flipped, the others offset), -0 taken as +0.
__device__ __forceinline__ uint32_t ordered_bits(float x) {
    const uint32_t b = __float_as_uint(x + 0.0f);
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool kCount>
__device__ __forceinline__ void wave_test_tris(const SceneView& sv, TravState& s, uint32_t tmask,
                                               uint32_t tri_base, TriShare& sh) {
    const uint32_t lane = __lane_id();
    const uint32_t cnt = (uint32_t)__popc(tmask);
    if (kCount) s.ntest += cnt;
    // inclusive prefix sum of cnt over the wave: within rows of 16 lanes
    // (row shifts by 1, 2, 4, 8), then row 15's sum into rows 1 and 3 and
    // row 31's into rows 2 and 3 (row broadcasts)
    uint32_t incl = cnt;
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x111, 0xF, 0xF, true);
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x112, 0xF, 0xF, true);
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x114, 0xF, 0xF, true);
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x118, 0xF, 0xF, true);
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x142, 0xA, 0xF, false);
    incl += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)incl, 0x143, 0xC, 0xF, false);
    uint32_t next = incl - cnt;                                        // the lane's first list position
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);  // the wave's tests
    unsigned long long key0 = 0;
    if (cnt) {
        // (with no hit and a NaN tmax nothing is below key 0, as no t is
        // below a NaN)
        key0 = s.best.tri != kInvalid ? ((unsigned long long)ordered_bits(s.best.t) << 32) | s.best.tri
               : s.best.t == s.best.t ? ((unsigned long long)(ordered_bits(s.best.t) - 1u) << 32) | 0xFFFFFFFFull
                                      : 0ull;
        sh.key[lane] = key0;
    }
    for (uint32_t c0 = 0; c0 < total; c0 += 64u) {
        while (tmask && next < c0 + 64u) {
            const uint32_t k = (uint32_t)__ffs(tmask) - 1u;
            tmask &= tmask - 1u;
            sh.item[next - c0] = (uint16_t)(lane | (k << 6));
            ++next;
        }
        wave_sync_lds();
        const uint32_t it = sh.item[lane];
        const int ow = (int)(it & 63u);
        const f3 o = mk3(__shfl(s.o.x, ow), __shfl(s.o.y, ow), __shfl(s.o.z, ow));
        const f3 d = mk3(__shfl(s.d.x, ow), __shfl(s.d.y, ow), __shfl(s.d.z, ow));
        const float tmin = __shfl(s.tmin, ow);
        const uint32_t base = (uint32_t)__shfl((int)tri_base, ow);
        unsigned long long key = ~0ull;
        float tt, u, v;
        bool fr;
        if (c0 + lane < total) {
            const float4* tr = sv.tris + 3ull * (base + (it >> 6));
            const float4 a = tr[0], b = tr[1], cc = tr[2];
            asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z), "v"(cc.x),
                         "v"(cc.y), "v"(cc.z));  // one round trip for the three loads, as in test_tri
            if (tri_intersect(o, d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), tt, u, v, fr) &&
                tt > tmin) {
                key = ((unsigned long long)ordered_bits(tt) << 32) | __float_as_uint(a.w);
                atomicMin(&sh.key[ow], key);
            }
        }
        wave_sync_lds();
        if (key != ~0ull && sh.key[ow] == key) {
            sh.t[ow] = tt;
            sh.u[ow] = u;
            sh.v[ow] = v;
            sh.front[ow] = fr ? 1u : 0u;
        }
        wave_sync_lds();
    }
    if (cnt) {
        const unsigned long long k = sh.key[lane];
        if (k < key0) {
            s.best.t = sh.t[lane];
            s.best.tri = (uint32_t)k;
            s.best.u = sh.u[lane];
            s.best.v = sh.v[lane];
            s.best.front = sh.front[lane] != 0u;
        }
    }
}

// trav_step for closest hits with the triangle tests shared by the wave
// (wave_test_tris). Every lane of the wave calls it, in wave-uniform control
// flow; `active` marks the lanes holding a ray. Returns, for an active lane,
// what trav_step<false, kCount> returns, with the same hit.
template <bool kCount>
__device__ __forceinline__ bool trav_step_wave(const SceneView& sv, TravState& s, Stack st, bool active,
                                               TriShare& sh) {
    bool more = active, visit = active;
    if (active && !s.have_next) {
        if (s.sp == 0) {
            more = visit = false;
        } else {
            const uint2 e = st.get(--s.sp);
            if (__uint_as_float(e.y) > s.best.t * kBoxPad)
                visit = false;
            else
                s.next = e.x;
        }
    }
    NodeVisit nv;
    uint32_t tmask = 0, tri_base = 0;
    if (visit) {
        visit_node(sv, s, nv);
        if (kCount) s.nvis++;
        const uint32_t leaves = nv.hitmask & ~nv.imask;
        if (leaves) tmask = leaf_tri_mask(nv, leaves);
        tri_base = nv.tri_base;
    }
    if (__ballot(tmask != 0u)) wave_test_tris<kCount>(sv, s, tmask, tri_base, sh);
    if (visit) push_children(s, st, nv);
    return more;
}

template <bool kAny, bool kCount>
__device__ Hit traverse(const SceneView& sv, f3 o, f3 d, float tmin, float tmax, Stack st,
                        uint32_t* n_nodes_visited = nullptr, uint32_t* n_tris_tested = nullptr) {
    TravState s;
    trav_init(sv, s, o, d, tmin, tmax);
    while (trav_step<kAny, kCount>(sv, s, st)) {
    }
    if (kCount) {
        *n_nodes_visited = s.nvis;
        *n_tris_tested = s.ntest;
    }
    return s.best;
}

// Whether a ray (o, d) on (tmin, tmax) meets any child box of the root of
// sv's BVH: the same slab test as traverse() on one node. A ray that meets
// none reaches no triangle; one that meets a box may still miss them all.
__device__ __forceinline__ bool meets_root_child(const SceneView& sv, f3 o, f3 d, float tmin, float tmax) {
    if (sv.n_nodes == 0) return false;
    const float inv[3] = {safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    const float oo[3] = {o.x, o.y, o.z};
    const uint4* np = reinterpret_cast<const uint4*>(sv.nodes);
    const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
    const float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    const uint32_t lo_l[3] = {w2.x, w2.z, w3.x}, lo_h[3] = {w2.y, w2.w, w3.y};
    const uint32_t hi_l[3] = {w3.z, w4.x, w4.z}, hi_h[3] = {w3.w, w4.y, w4.w};
    float as[3], an[3], fs[3], fo[3];
    uint32_t nl[3], nh[3], fl[3], fh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool neg = inv[a] < 0.0f;
        const int ex = (int)(int8_t)((w0.w >> (8 * a)) & 0xFFu);
        as[a] = ldexpf(inv[a], ex);
        const float ao = (org[a] - oo[a]) * inv[a];
        const float e = (fabsf(as[a]) * 255.0f + fabsf(ao)) * 4.0e-7f;  // as in trav_step
        an[a] = ao - e;
        fs[a] = as[a] * kBoxPad;
        fo[a] = __fmaf_rn(ao, kBoxPad, e * (1.0f + kBoxPad));
        nl[a] = neg ? hi_l[a] : lo_l[a];
        nh[a] = neg ? hi_h[a] : lo_h[a];
        fl[a] = neg ? lo_l[a] : hi_l[a];
        fh[a] = neg ? lo_h[a] : hi_h[a];
    }
    const float tcap = tmax * kBoxPad;
    bool any = false;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float tn = tmin, tf = tcap;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            tn = fmaxf(tn, __fmaf_rn((float)byte_of(nl[a], nh[a], c), as[a], an[a]));
            tf = fminf(tf, __fmaf_rn((float)byte_of(fl[a], fh[a], c), fs[a], fo[a]));
        }
        any |= byte_of(w1.z, w1.w, c) != 0u && tn <= tf;
    }
    return any;
}

// Closest or any hit, counting node visits and triangle tests into the
// caller's per-lane totals when `count` is set (a wave-uniform flag).
template <bool kAny>
__device__ __forceinline__ Hit trace_ray(const SceneView& sv, f3 o, f3 d, float tmin, float tmax,
                                         Stack st, bool count, uint32_t& nodes, uint32_t& tris) {
    if (count) {
        uint32_t nv = 0, nt = 0;
        Hit h = traverse<kAny, true>(sv, o, d, tmin, tmax, st, &nv, &nt);
        nodes += nv;
        tris += nt;
        return h;
    }
    return traverse<kAny, false>(sv, o, d, tmin, tmax, st);
}

// Closest hit of one ray per lane, traversed step by step by the whole wave
// with the triangle tests shared (trav_step_wave); lanes with have == false
// only help. Every lane of the wave calls it, in wave-uniform control flow.
template <bool kCount>
__device__ __forceinline__ Hit trace_wave(const SceneView& sv, f3 o, f3 d, float tmin, float tmax, bool have,
                                          Stack st, TriShare& sh, uint32_t& nodes, uint32_t& tris) {
    TravState s;
    trav_init(sv, s, o, d, tmin, tmax);
    bool act = have;
    while (__ballot(act)) act = trav_step_wave<kCount>(sv, s, st, act, sh);
    if (kCount && have) {
        nodes += s.nvis;
        tris += s.ntest;
    }
    return s.best;
}

// The node visits a wave spends in lockstep on one round of rays: the most
// visits of any lane times the lanes that traversed. Every lane of the wave
// calls it, in wave-uniform control flow.
__device__ __forceinline__ uint32_t wave_lockstep_nodes(uint32_t nodes, bool traced) {
    uint32_t m = traced ? nodes : 0u;
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    return m * (uint32_t)__popcll(__ballot(traced));
}

// Adds per-lane traversal totals of a wave to the frame counters.
// max_nodes: the most node visits of one of the lane's rays (0: not kept).
__device__ __forceinline__ void flush_traversal_counts(Counters* ctr, int kind, uint32_t rays,
                                                       uint32_t nodes, uint32_t tris, uint32_t max_nodes = 0) {
    unsigned long long r = rays, n = nodes, t = tris;
    uint32_t m = max_nodes;
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        n += __shfl_xor(n, off, 64);
        t += __shfl_xor(t, off, 64);
        m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    }
    if (__lane_id() == 0 && r) {
        atomicAdd(&ctr->trav_rays[kind], r);
        atomicAdd(&ctr->trav_nodes[kind], n);
        atomicAdd(&ctr->trav_tris[kind], t);
        if (m) atomicMax(&ctr->ray_max_nodes[kind], (unsigned long long)m);
    }
}

}  // namespace wsrt
