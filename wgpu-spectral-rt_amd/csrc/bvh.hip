// bvh.hip - GPU construction of the compressed BVH8 that replaces the
// driver-built BLAS/TLAS of src/scene/builder.rs:131-254.
//
//  1. world triangles: every instance's mesh transformed by its matrix (f32,
//     glam's mul_vec4 summation order), triangle AABBs and centroids;
//  2. 30-bit Morton codes of the centroids over the centroid bounds, and
//     sort keys that add a size-class bit above them (large triangles last);
//  3. stable radix sort of (code, triangle);
//  4. binary radix tree (Karras 2012), duplicates split by position;
//  5. bottom-up boxes and triangle counts (one atomic flag per node);
//  6. top-down collapse, one level per launch: each BVH8 node opens its
//     child of largest surface area (whatever its triangle count) until it
//     has 8, subtrees of <= 3 triangles become leaf children, the others new
//     BVH8 nodes; child boxes
//     are quantized to 8 bits per axis against a power-of-two grid, rounded
//     outward in f64 so every dequantized box contains its child.
#include <math.h>

#include <vector>

#include "kernels.hpp"

namespace wsrt {

namespace {

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ void mul_point(const float* m, float x, float y, float z, float* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float s = m[0 + i] * x;
        s = s + m[4 + i] * y;
        s = s + m[8 + i] * z;
        s = s + m[12 + i] * 1.0f;
        r[i] = s;
    }
}

__global__ __launch_bounds__(256) void k_world_tris(const Instance* __restrict__ inst,
                                                    uint32_t n_inst,
                                                    const wsrt_vertex* __restrict__ verts,
                                                    const uint32_t* __restrict__ indices,
                                                    const wsrt_mesh_info* __restrict__ meshes,
                                                    uint32_t n_tris, float* __restrict__ wv,
                                                    float* __restrict__ box, float* __restrict__ cent,
                                                    uint32_t* __restrict__ tri_inst,
                                                    uint32_t* __restrict__ cbounds) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float c[3] = {0, 0, 0};
    bool ok = t < n_tris;
    if (ok) {
        uint32_t lo = 0, hi = n_inst - 1;
        while (lo < hi) {  // last instance with first_tri <= t
            uint32_t mid = (lo + hi + 1) >> 1;
            if (inst[mid].first_tri <= t) lo = mid;
            else hi = mid - 1;
        }
        const Instance& in = inst[lo];
        uint32_t prim = t - in.first_tri;
        const wsrt_mesh_info mi = meshes[in.mesh];
        float w[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float* p = verts[mi.vertex_offset + indices[mi.index_offset + 3 * prim + k]].pos;
            mul_point(in.m, p[0], p[1], p[2], w[k]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) wv[9ull * t + 3 * k + a] = w[k][a];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float l = fminf(w[0][a], fminf(w[1][a], w[2][a]));
            float h = fmaxf(w[0][a], fmaxf(w[1][a], w[2][a]));
            box[6ull * t + a] = l;
            box[6ull * t + 3 + a] = h;
            c[a] = (l + h) * 0.5f;
            cent[3ull * t + a] = c[a];
        }
        tri_inst[t] = lo;
    }
    // centroid bounds: wave min/max then one atomic per wave
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t mn = ok ? f2ord(c[a]) : 0xFFFFFFFFu;
        uint32_t mx = ok ? f2ord(c[a]) : 0u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, (uint32_t)__shfl_xor(mn, off, 64));
            mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
        }
        if (__lane_id() == 0) {
            atomicMin(&cbounds[a], mn);
            atomicMax(&cbounds[3 + a], mx);
        }
    }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of each triangle's box centroid over the centroid
// bounds (codes), and its sort key: the code with bit 30 set for a large
// triangle, one whose box spans more than 1/8 of the centroid bounds on some
// axis. The radix tree's first split then puts the large triangles (walls,
// floors, ground planes) in a subtree of their own, so their boxes do not
// inflate every ancestor of the small triangles near them.
__global__ void k_morton(const float* __restrict__ cent, const float* __restrict__ box, uint32_t n,
                         const uint32_t* __restrict__ cbounds, uint32_t* __restrict__ keys,
                         uint32_t* __restrict__ codes, uint32_t* __restrict__ vals) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t q[3];
    bool large = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float mn = ord2f(cbounds[a]), mx = ord2f(cbounds[3 + a]);
        float ext = mx - mn;
        float x = ext > 0.0f ? (cent[3ull * t + a] - mn) / ext : 0.0f;
        float y = x * 1024.0f;
        uint32_t qi = (uint32_t)y;
        q[a] = qi > 1023u ? 1023u : qi;
        const float bext = box[6ull * t + 3 + a] - box[6ull * t + a];
        large = large || (ext > 0.0f && bext > 0.125f * ext);
    }
    const uint32_t code = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    codes[t] = code;
    keys[t] = code | (large ? 1u << 30 : 0u);
    vals[t] = t;
}

// Karras 2012 binary radix tree over n >= 2 sorted codes. Children are
// encoded with bit 31 set for leaves (sorted positions).
__device__ __forceinline__ int delta(const uint32_t* codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

__global__ void k_radix_tree(const uint32_t* __restrict__ codes, int n, uint32_t* __restrict__ left,
                             uint32_t* __restrict__ right, uint32_t* __restrict__ parent_int,
                             uint32_t* __restrict__ parent_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(codes, n, i, j);
    int s = 0;
    int len = l;
    for (int t = (len + 1) >> 1;; t = (t + 1) >> 1) {
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t lc = (lo == gamma) ? (0x80000000u | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t rc = (hi == gamma + 1) ? (0x80000000u | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    if (lc & 0x80000000u) parent_leaf[lc & 0x7FFFFFFFu] = i;
    else parent_int[lc] = i;
    if (rc & 0x80000000u) parent_leaf[rc & 0x7FFFFFFFu] = i;
    else parent_int[rc] = i;
}

__device__ __forceinline__ void child_box(uint32_t c, const float* ibox, const float* tbox,
                                          const uint32_t* order, float* b, const uint32_t* icnt,
                                          uint32_t* cnt) {
    if (c & 0x80000000u) {
        uint32_t t = order[c & 0x7FFFFFFFu];
#pragma unroll
        for (int k = 0; k < 6; ++k) b[k] = tbox[6ull * t + k];
        *cnt = 1;
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) b[k] = ibox[6ull * c + k];
        *cnt = icnt[c];
    }
}

__global__ void k_bottom_up(int n, const uint32_t* __restrict__ left,
                            const uint32_t* __restrict__ right,
                            const uint32_t* __restrict__ parent_int,
                            const uint32_t* __restrict__ parent_leaf,
                            const uint32_t* __restrict__ order, const float* __restrict__ tbox,
                            float* ibox, uint32_t* icnt, uint32_t* flags) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = parent_leaf[j];
    while (true) {
        // acq_rel at agent scope: publishes this thread's box writes and, for
        // the second arrival, invalidates stale L1 lines before the reads.
        if (__hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            return;  // first arrival: the sibling subtree is not finished yet
        float a[6], b[6];
        uint32_t ca, cb;
        child_box(left[p], ibox, tbox, order, a, icnt, &ca);
        child_box(right[p], ibox, tbox, order, b, icnt, &cb);
        float* o = ibox + 6ull * p;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[k] = fminf(a[k], b[k]);
            o[3 + k] = fmaxf(a[3 + k], b[3 + k]);
        }
        icnt[p] = ca + cb;
        if (p == 0) return;
        p = parent_int[p];
    }
}

struct CollapseArgs {
    const uint32_t* left;
    const uint32_t* right;
    const uint32_t* order;
    const float* ibox;
    const uint32_t* icnt;
    const float* tbox;
    const float* wv;
    const uint32_t* tri_inst;
    const Instance* inst;
    wsrt_bvh8_node* nodes;
    float4* tris;
    uint32_t* counters;  // [0] nodes, [1] tris, [2] max depth, [3] next queue size
    int n_leaves;
};

__device__ __forceinline__ float half_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ void write_tri(const CollapseArgs& A, uint32_t slot, uint32_t t) {
    const float* w = A.wv + 9ull * t;
    uint32_t in = A.tri_inst[t];
    float4* r = A.tris + 3ull * slot;
    r[0] = make_float4(w[0], w[1], w[2], __uint_as_float(t));
    r[1] = make_float4(w[3] - w[0], w[4] - w[1], w[5] - w[2], __uint_as_float(in));
    r[2] = make_float4(w[6] - w[0], w[7] - w[1], w[8] - w[2], __uint_as_float(A.inst[in].material));
}

// One thread per queued (binary node, BVH8 node, depth).
__global__ void k_collapse(CollapseArgs A, const uint4* __restrict__ qin, uint32_t n_in,
                           uint4* __restrict__ qout) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_in) return;
    uint4 item = qin[w];
    uint32_t b = item.x, node = item.y, depth = item.z;
    atomicMax(&A.counters[2], depth);
    uint32_t ent[8];
    float eb[8][6];
    uint32_t ec[8];
    int m = 0;
    float nb[6];
    if (b == kInvalid) {  // single-triangle scene: one leaf child
        ent[0] = 0x80000000u;
        child_box(ent[0], A.ibox, A.tbox, A.order, eb[0], A.icnt, &ec[0]);
        m = 1;
#pragma unroll
        for (int k = 0; k < 6; ++k) nb[k] = eb[0][k];
    } else {
        ent[0] = A.left[b];
        ent[1] = A.right[b];
        child_box(ent[0], A.ibox, A.tbox, A.order, eb[0], A.icnt, &ec[0]);
        child_box(ent[1], A.ibox, A.tbox, A.order, eb[1], A.icnt, &ec[1]);
        m = 2;
        for (int k = 0; k < 6; ++k) nb[k] = A.ibox[6ull * b + k];
        while (m < 8) {
            int best = -1;
            float best_s = -1.0f;
            for (int k = 0; k < m; ++k) {
                if (ent[k] & 0x80000000u) continue;
                float s = half_area(eb[k]);
                if (s > best_s) { best_s = s; best = k; }
            }
            if (best < 0) break;
            uint32_t e = ent[best];
            ent[best] = A.left[e];
            ent[m] = A.right[e];
            child_box(ent[best], A.ibox, A.tbox, A.order, eb[best], A.icnt, &ec[best]);
            child_box(ent[m], A.ibox, A.tbox, A.order, eb[m], A.icnt, &ec[m]);
            ++m;
        }
    }
    uint32_t n_int = 0, n_leaf_tris = 0;
    for (int k = 0; k < m; ++k) {
        if (ec[k] > 3u) n_int++;
        else n_leaf_tris += ec[k];
    }
    uint32_t child_base = n_int ? atomicAdd(&A.counters[0], n_int) : 0u;
    uint32_t tri_base = n_leaf_tris ? atomicAdd(&A.counters[1], n_leaf_tris) : 0u;
    uint32_t qbase = n_int ? atomicAdd(&A.counters[3], n_int) : 0u;

    wsrt_bvh8_node out;
    int ex[3];
    double scale[3];
    for (int a = 0; a < 3; ++a) {
        double ext = (double)nb[3 + a] - (double)nb[a];
        int e = -100;
        if (ext > 0.0) {
            e = (int)ceil(log2(ext / 255.0));
            if (ldexp(255.0, e) < ext) e++;
            if (e < -100) e = -100;
            if (e > 127) e = 127;
        }
        ex[a] = e;
        scale[a] = ldexp(1.0, e);
        out.origin[a] = nb[a];
        out.exp[a] = (int8_t)e;
    }
    out.child_base = child_base;
    out.tri_base = tri_base;
    // Slot of each child = the octant of its box centre around the node's
    // centre (bit a set: upper half along axis a), resolved greedily, largest
    // child first, to the free slot differing in the fewest axes. Traversal
    // then visits slots in the order (k ^ ray octant), near to far, without
    // sorting distances.
    int slot_of[8];
    {
        uint32_t used = 0;
        bool done[8] = {false, false, false, false, false, false, false, false};
        for (int it = 0; it < m; ++it) {
            int pick = -1;
            float best_a = -1.0f;
            for (int k = 0; k < m; ++k)
                if (!done[k] && half_area(eb[k]) > best_a) { best_a = half_area(eb[k]); pick = k; }
            done[pick] = true;
            uint32_t want = 0;
            for (int a = 0; a < 3; ++a)
                if ((eb[pick][a] + eb[pick][3 + a]) > (nb[a] + nb[3 + a])) want |= 1u << a;
            int sbest = -1, cbest = 99;
            for (int sl = 0; sl < 8; ++sl) {
                if (used & (1u << sl)) continue;
                int cost = __popc((uint32_t)sl ^ want);
                if (cost < cbest) { cbest = cost; sbest = sl; }
            }
            slot_of[pick] = sbest;
            used |= 1u << sbest;
        }
        for (int sl = 0; sl < 8; ++sl) {
            if (used & (1u << sl)) continue;
            // empty: an inverted box (lo above hi), which the slab test rejects
            out.meta[sl] = 0;
            for (int a = 0; a < 3; ++a) { out.qlo[a][sl] = 255; out.qhi[a][sl] = 0; }
        }
    }
    uint8_t imask = 0;
    uint32_t ri = 0, off = 0;
    for (int k = 0; k < m; ++k) {
        const int sl = slot_of[k];
        for (int a = 0; a < 3; ++a) {
            double lo = floor(((double)eb[k][a] - (double)nb[a]) / scale[a]);
            double hi = ceil(((double)eb[k][3 + a] - (double)nb[a]) / scale[a]);
            lo = lo < 0.0 ? 0.0 : (lo > 255.0 ? 255.0 : lo);
            hi = hi < 0.0 ? 0.0 : (hi > 255.0 ? 255.0 : hi);
            out.qlo[a][sl] = (uint8_t)lo;
            out.qhi[a][sl] = (uint8_t)hi;
        }
        if (ec[k] > 3u) {
            out.meta[sl] = (uint8_t)(0x80u | ri);
            imask |= (uint8_t)(1u << sl);
            qout[qbase + ri] = make_uint4(ent[k], child_base + ri, depth + 1, 0);
            ri++;
        } else {
            out.meta[sl] = (uint8_t)((ec[k] << 5) | off);
            // gather the <= 3 triangles of this subtree in left-to-right order
            uint32_t stk[4];
            int sp = 0;
            stk[sp++] = ent[k];
            uint32_t placed = 0;
            while (sp) {
                uint32_t e = stk[--sp];
                if (e & 0x80000000u) {
                    write_tri(A, tri_base + off + placed, A.order[e & 0x7FFFFFFFu]);
                    placed++;
                } else {
                    stk[sp++] = A.right[e];
                    stk[sp++] = A.left[e];
                }
            }
            off += ec[k];
        }
    }
    out.imask = imask;
    A.nodes[node] = out;
}

}  // namespace

int bvh_build(const BvhInputs& in, BvhOutputs* out, hipStream_t s) {
    const uint32_t n = in.n_tris;
    out->n_nodes = 0;
    out->max_depth = 0;
    out->placed_tris = 0;
    out->nodes = nullptr;
    out->tris = nullptr;
    out->wv = nullptr;
    out->tri_inst = nullptr;
    out->codes = nullptr;
    out->sorted = nullptr;
    if (n == 0) return 0;
    // temporaries and results
    auto alloc = [](void** p, size_t bytes) { return hipMalloc(p, bytes < 16 ? 16 : bytes); };
    float *wv, *box, *cent, *ibox;
    uint32_t *tri_inst, *cbounds, *codes, *vals, *left, *right, *pint, *pleaf, *icnt, *flags,
        *sort_tmp, *counters;
    uint4 *qa, *qb;
    wsrt_bvh8_node* nodes;
    float4* tris;
    size_t sort_words = radix_temp_words(n);
    hipError_t e = hipSuccess;
#define A_(p, bytes) \
    if (e == hipSuccess) e = alloc((void**)&p, bytes)
    A_(wv, sizeof(float) * 9 * n);
    A_(box, sizeof(float) * 6 * n);
    A_(cent, sizeof(float) * 3 * n);
    A_(tri_inst, sizeof(uint32_t) * n);
    A_(cbounds, sizeof(uint32_t) * 8);
    A_(codes, sizeof(uint32_t) * n);
    A_(vals, sizeof(uint32_t) * n);
    A_(left, sizeof(uint32_t) * n);
    A_(right, sizeof(uint32_t) * n);
    A_(pint, sizeof(uint32_t) * n);
    A_(pleaf, sizeof(uint32_t) * n);
    A_(ibox, sizeof(float) * 6 * n);
    A_(icnt, sizeof(uint32_t) * n);
    A_(flags, sizeof(uint32_t) * n);
    A_(sort_tmp, sizeof(uint32_t) * sort_words);
    A_(counters, sizeof(uint32_t) * 8);
    A_(qa, sizeof(uint4) * n);
    A_(qb, sizeof(uint4) * n);
    A_(nodes, sizeof(wsrt_bvh8_node) * n);
    A_(tris, sizeof(float4) * 3 * n);
#undef A_
    if (e != hipSuccess) return -2;
    const uint32_t init_b[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(cbounds, init_b, sizeof init_b, hipMemcpyHostToDevice, s)) return -3;
    uint32_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_world_tris, dim3(blocks), dim3(256), 0, s, in.instances, in.n_instances,
                       in.verts, in.indices, in.meshes, n, wv, box, cent, tri_inst, cbounds);
    // the unsorted codes are kept for inspection; the sort keys go to codes
    uint32_t* codes_copy;
    if (hipMalloc((void**)&codes_copy, sizeof(uint32_t) * n) != hipSuccess) return -2;
    hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, s, cent, box, n, cbounds, codes, codes_copy, vals);
    bool in_tmp = false;
    if (radix_sort_pairs(codes, vals, n, 31, sort_tmp, s, &in_tmp)) return -1;
    uint32_t* scodes = in_tmp ? sort_tmp : codes;
    uint32_t* order = in_tmp ? sort_tmp + n : vals;
    fill_u32(flags, n, 0u, s);
    fill_u32(counters, 8, 0u, s);
    if (n >= 2) {
        hipLaunchKernelGGL(k_radix_tree, dim3((n - 1 + 255) / 256), dim3(256), 0, s, scodes, (int)n,
                           left, right, pint, pleaf);
        hipLaunchKernelGGL(k_bottom_up, dim3(blocks), dim3(256), 0, s, (int)n, left, right, pint,
                           pleaf, order, box, ibox, icnt, flags);
    }
    CollapseArgs ca;
    ca.left = left;
    ca.right = right;
    ca.order = order;
    ca.ibox = ibox;
    ca.icnt = icnt;
    ca.tbox = box;
    ca.wv = wv;
    ca.tri_inst = tri_inst;
    ca.inst = in.instances;
    ca.nodes = nodes;
    ca.tris = tris;
    ca.counters = counters;
    ca.n_leaves = (int)n;
    uint4 root = make_uint4(n >= 2 ? 0u : kInvalid, 0u, 1u, 0u);
    const uint32_t c_init[4] = {1u, 0u, 0u, 0u};
    if (hipMemcpyAsync(qa, &root, sizeof root, hipMemcpyHostToDevice, s)) return -3;
    if (hipMemcpyAsync(counters, c_init, sizeof c_init, hipMemcpyHostToDevice, s)) return -3;
    uint32_t n_in = 1;
    while (n_in) {
        fill_u32(counters + 3, 1, 0u, s);
        hipLaunchKernelGGL(k_collapse, dim3((n_in + 63) / 64), dim3(64), 0, s, ca, qa, n_in, qb);
        uint32_t next = 0;
        if (hipMemcpyAsync(&next, counters + 3, sizeof next, hipMemcpyDeviceToHost, s)) return -3;
        if (hipStreamSynchronize(s) != hipSuccess) return -3;
        uint4* t = qa;
        qa = qb;
        qb = t;
        n_in = next;
    }
    uint32_t cnt[4];
    if (hipMemcpyAsync(cnt, counters, sizeof cnt, hipMemcpyDeviceToHost, s)) return -3;
    if (hipStreamSynchronize(s) != hipSuccess) return -3;
    out->n_nodes = cnt[0];
    out->max_depth = cnt[2];
    out->placed_tris = cnt[1];
    out->nodes = nodes;
    out->tris = tris;
    out->wv = wv;
    out->tri_inst = tri_inst;
    out->codes = codes_copy;
    // sorted order (triangle ids by code)
    uint32_t* sorted;
    if (hipMalloc((void**)&sorted, sizeof(uint32_t) * n) != hipSuccess) return -2;
    if (hipMemcpyAsync(sorted, order, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s)) return -3;
    out->sorted = sorted;
    if (hipStreamSynchronize(s)) return -3;
    (void)hipFree(box);
    (void)hipFree(cent);
    (void)hipFree(cbounds);
    (void)hipFree(codes);
    (void)hipFree(vals);
    (void)hipFree(left);
    (void)hipFree(right);
    (void)hipFree(pint);
    (void)hipFree(pleaf);
    (void)hipFree(ibox);
    (void)hipFree(icnt);
    (void)hipFree(flags);
    (void)hipFree(sort_tmp);
    (void)hipFree(counters);
    (void)hipFree(qa);
    (void)hipFree(qb);
    return 0;
}

}  // namespace wsrt
