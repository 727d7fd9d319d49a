// host_mirror.cpp - drives libwsrt through the C++ host mirror
// (wgpu-spectral-rt_amd/host/wsrt.hpp), for tests/test_host_cpp.py.
//
//   host_mirror cpu                 meshes, a camera uniform, the default
//                                   config and an error path, one line each
//   host_mirror gpu W H FRAMES OUT  renders the Cornell box W x H for FRAMES
//                                   frames and writes the accumulation
//                                   (float4 per pixel) to OUT
#include <cstdio>
#include <cstdlib>
#include <string>

#include "wsrt.hpp"

namespace {

void print_mesh(const char* name, const wsrt::scene::geometry::Mesh& m) {
    double vsum = 0.0;
    for (const auto& v : m.first)
        for (int k = 0; k < 3; ++k) vsum += (double)v.pos[k] + 2.0 * (double)v.normal[k];
    unsigned long long isum = 0;
    for (uint32_t i : m.second) isum += i;
    std::printf("%s %zu %zu %.9e %llu\n", name, m.first.size(), m.second.size(), vsum, isum);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") {
            namespace geo = wsrt::scene::geometry;
            print_mesh("plane", geo::create_plane());
            print_mesh("cube", geo::create_cube());
            print_mesh("sphere4", geo::create_sphere(4));
            const wsrt::CameraUniform u = wsrt::scene::CameraController{}.build_uniform(1.5f, 7);
            std::printf("camera");
            for (float f : u.view_inverse) std::printf(" %.9e", f);
            std::printf(" %u\n", u.frame_count);
            const wsrt_config c = wsrt::renderer::default_config();
            std::printf("config %u %u %u %u\n", c.max_depth, c.photon_max_depth, c.photons_per_pass, c.hash_size);
            try {
                wsrt::scene::SceneBuilder b;
                b.add_instance(3, 0, wsrt::mat4::identity());
                std::printf("error none\n");
            } catch (const wsrt::Error& e) {
                std::printf("error %d\n", e.code());
            }
            return 0;
        }
        if (mode == "gpu" && argc == 6) {
            const uint32_t W = (uint32_t)std::atoi(argv[2]), H = (uint32_t)std::atoi(argv[3]);
            const uint32_t F = (uint32_t)std::atoi(argv[4]);
            wsrt::Device dev(0);
            auto scene = wsrt::scene::create_cornell_box(dev);
            wsrt_config cfg = wsrt::renderer::default_config();
            cfg.photons_per_pass = 1u << 16;
            wsrt::renderer::Renderer r(dev, scene, W, H, cfg);
            r.set_camera(wsrt::scene::CameraController{}.build_uniform((float)W / (float)H, 0));
            r.render_frames(F);
            const std::vector<float> acc = r.read_accum();
            std::printf("frames %u\n", r.frame_count());
            FILE* f = std::fopen(argv[5], "wb");
            if (!f || std::fwrite(acc.data(), sizeof(float), acc.size(), f) != acc.size()) return 3;
            std::fclose(f);
            return 0;
        }
        std::fprintf(stderr, "usage: host_mirror cpu | gpu W H FRAMES OUT\n");
        return 1;
    } catch (const wsrt::Error& e) {
        std::fprintf(stderr, "wsrt error %d: %s\n", e.code(), e.what());
        return 2;
    }
}
