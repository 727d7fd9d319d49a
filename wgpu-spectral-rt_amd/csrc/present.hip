// present.hip - the presented frame: the blit pass (blit_pass.rs:114-180,
// blit.wgsl) and, on request, the photon overlay (debug_photons_pass.rs,
// debug_photons.wgsl), drawn into a window-sized RGBA8 image.
//
// The blit draws the storage image as a quad scaled to keep its aspect
// (scale_x or scale_y below 1, no offset) over a blue clear, sampled with
// wgpu's default sampler (nearest, clamp to edge), into the surface's
// default sRGB format. Texture row 0, the image bottom, lands at the bottom
// of the window. The overlay draws each stored photon as a one-pixel point
// of colour (1, 0.6, 0.1) and alpha 0.3, blended additively (src * alpha +
// dst, alpha replaced) in linear space, one point after another.
#include "kernels.hpp"

namespace wsrt {

namespace {

// Pixel (x, y) of the window (y = 0 at the top) that each stored photon's
// point covers; counts per pixel. Photon slot i holds a photon of this
// frame when bit first + i of stored is set.
__global__ __launch_bounds__(256) void k_present_points(const wsrt_photon* __restrict__ photons,
                                                        const uint32_t* __restrict__ stored, uint32_t first,
                                                        uint32_t n, PresentArgs a, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = first + i;
    if (!((stored[b >> 5] >> (b & 31u)) & 1u)) return;
    const float4 p0 = reinterpret_cast<const float4*>(photons + i)[0];
    const float e = reinterpret_cast<const float4*>(photons + i)[1].w;
    if (e == 0.0f) return;  // not written this frame: moved off screen
    float v[4], c[4];
    const float p[4] = {p0.x, p0.y, p0.z, 1.0f};
    for (int r = 0; r < 4; ++r) {
        float s = a.view[r] * p[0];
        s = s + a.view[4 + r] * p[1];
        s = s + a.view[8 + r] * p[2];
        s = s + a.view[12 + r] * p[3];
        v[r] = s;
    }
    for (int r = 0; r < 4; ++r) {
        float s = a.proj[r] * v[0];
        s = s + a.proj[4 + r] * v[1];
        s = s + a.proj[8 + r] * v[2];
        s = s + a.proj[12 + r] * v[3];
        c[r] = s;
    }
    // view-volume clip of a point: -w <= x, y <= w and 0 <= z <= w
    if (!(c[3] > 0.0f) || c[0] < -c[3] || c[0] > c[3] || c[1] < -c[3] || c[1] > c[3] || c[2] < 0.0f ||
        c[2] > c[3])
        return;
    const float nx = c[0] / c[3], ny = c[1] / c[3];
    const float fx = (nx + 1.0f) * 0.5f * (float)a.win_w;
    const float fy = (1.0f - ny) * 0.5f * (float)a.win_h;
    const int px = (int)floorf(fx), py = (int)floorf(fy);
    if (px < 0 || py < 0 || px >= (int)a.win_w || py >= (int)a.win_h) return;
    atomicAdd(&count[(uint32_t)py * a.win_w + (uint32_t)px], 1u);
}

__global__ __launch_bounds__(256) void k_present(const uchar4* __restrict__ image, PresentArgs a,
                                                 const uint32_t* __restrict__ count, uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.win_w * a.win_h) return;
    const uint32_t x = i % a.win_w, y = i / a.win_w;
    // fragment centre in normalized device coordinates (y up)
    const float ndx = ((float)x + 0.5f) / (float)a.win_w * 2.0f - 1.0f;
    const float ndy = 1.0f - ((float)y + 0.5f) / (float)a.win_h * 2.0f;
    uchar4 px = make_uchar4(0, 0, 255, 255);  // the blue clear, sRGB-encoded
    // the quad spans [-sx, sx] x [-sy, sy]; left and top edges included
    if (ndx >= -a.sx && ndx < a.sx && ndy > -a.sy && ndy <= a.sy) {
        const float u = (ndx / a.sx + 1.0f) * 0.5f, v = (ndy / a.sy + 1.0f) * 0.5f;
        // clamp to edge (u or v can round a hair below 0 at the quad's edge)
        const uint32_t tx = min((uint32_t)fmaxf(floorf(u * (float)a.img_w), 0.0f), a.img_w - 1u);
        const uint32_t ty = min((uint32_t)fmaxf(floorf(v * (float)a.img_h), 0.0f), a.img_h - 1u);
        const uchar4 t = image[ty * a.img_w + tx];
        px = make_uchar4(a.lut[t.x], a.lut[t.y], a.lut[t.z], t.w);
    }
    if (count) {
        const uint32_t k = min(count[i], 256u);  // every channel saturates within 256 points
        if (k) {
            uint32_t r = px.x, g = px.y, b = px.z;
            for (uint32_t j = 0; j < k; ++j) {
                r = a.lut[256 + r];
                g = a.lut[512 + g];
                b = a.lut[768 + b];
            }
            px = make_uchar4((uint8_t)r, (uint8_t)g, (uint8_t)b, a.point_alpha);
        }
    }
    out[i] = px;
}

}  // namespace

void launch_present(const uint8_t* image, const wsrt_photon* photons, const uint32_t* stored, uint32_t first,
                    uint32_t n_photons, const PresentArgs& a, uint32_t* count, uint8_t* out, hipStream_t s) {
    if (count && n_photons)
        hipLaunchKernelGGL(k_present_points, dim3((n_photons + 255) / 256), dim3(256), 0, s, photons, stored, first,
                           n_photons, a, count);
    const uint32_t npx = a.win_w * a.win_h;
    hipLaunchKernelGGL(k_present, dim3((npx + 255) / 256), dim3(256), 0, s, reinterpret_cast<const uchar4*>(image),
                       a, count, reinterpret_cast<uchar4*>(out));
}

}  // namespace wsrt
