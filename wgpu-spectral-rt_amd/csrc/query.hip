// query.hip - batched ray queries through the C ABI (closest hit, any hit
// and traversal counts), the same traversal the render kernels run.
#include "kernels.hpp"
#include "traverse.hpp"

namespace wsrt {

namespace {

__global__ __launch_bounds__(kTraceThreads) void k_query(SceneView sv, const float* __restrict__ rays,
                                                         uint32_t n, int mode,
                                                         uint32_t* __restrict__ out0,
                                                         uint32_t* __restrict__ out1,
                                                         uint2* __restrict__ gstack) {
    __shared__ uint2 lds_stack[kShortStack * kTraceThreads];
    __shared__ TriShare share_all[kTraceThreads / 64];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st;
    st.lds = lds_slot(lds_stack + threadIdx.x);
    st.lds_stride = kTraceThreads;
    st.glob = gstack + (uint64_t)gtid * (sv.stack_size > kShortStack ? sv.stack_size - kShortStack : 0);
    const uint32_t stride = gridDim.x * blockDim.x;
    if (mode == 0) {
        // closest hits as the render kernels trace them: the wave traverses
        // its rays together, triangle tests shared (trace_wave)
        for (uint32_t base = gtid & ~63u; base < n; base += stride) {
            const uint32_t i = base + __lane_id();
            const bool have = i < n;
            const uint32_t k = have ? i : base;
            const float4 a = reinterpret_cast<const float4*>(rays)[2ull * k];
            const float4 b = reinterpret_cast<const float4*>(rays)[2ull * k + 1];
            uint32_t nn = 0, nt = 0;
            const Hit h = trace_wave<false>(sv, xyz(a), xyz(b), a.w, b.w, have, st, share_all[threadIdx.x >> 6], nn, nt);
            if (have) {
                const bool hit = h.tri != kInvalid;
                reinterpret_cast<uint4*>(out0)[i] =
                    make_uint4(__float_as_uint(hit ? h.t : INFINITY), h.tri, __float_as_uint(hit ? h.u : 0.0f),
                               __float_as_uint(hit ? h.v : 0.0f));
                out1[i] = hit ? (1u | (h.front ? 2u : 0u)) : 0u;
            }
        }
        return;
    }
    for (uint32_t i = gtid; i < n; i += stride) {
        const float4 a = reinterpret_cast<const float4*>(rays)[2ull * i];
        const float4 b = reinterpret_cast<const float4*>(rays)[2ull * i + 1];
        if (mode == 1) {
            Hit h = traverse<true, false>(sv, xyz(a), xyz(b), a.w, b.w, st);
            out0[i] = h.tri != kInvalid ? 1u : 0u;
        } else {
            uint32_t nv = 0, nt = 0;
            (void)traverse<false, true>(sv, xyz(a), xyz(b), a.w, b.w, st, &nv, &nt);
            out0[i] = nv;
            out1[i] = nt;
        }
    }
}

}  // namespace

void launch_query(const SceneView& sv, const float* rays, uint32_t n, int mode, uint32_t* out0,
                  uint32_t* out1, uint2* gstack, hipStream_t s) {
    hipLaunchKernelGGL(k_query, dim3(kTraceBlocks), dim3(kTraceThreads), 0, s, sv, rays, n, mode, out0,
                       out1, gstack);
}

}  // namespace wsrt
