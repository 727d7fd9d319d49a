// wsrt.hpp - C++ host API over the C ABI of libwsrt, mirroring the
// reference's Rust host interface for the render loop:
//
//   scene::geometry::{create_plane, create_cube, create_sphere}   geometry.rs:12-232
//   scene::SceneBuilder::{add_material, add_mesh, add_instance,
//                         add_light_instance, build}              builder.rs:57-266
//   scene::create_cornell_box                                      cornell_box.rs:6-110
//   scene::CameraController::build_uniform                         camera.rs:178-194
//   renderer::Renderer::{new, resize, render}, frame_count         renderer/mod.rs:21-207
//
// Where the Rust code panics or unwraps, these wrappers throw wsrt::Error
// carrying the library's status code and message. Header-only; link with
// -lwsrt.
#pragma once

#include <stdint.h>

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wsrt.h"

namespace wsrt {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline void check(int rc) {
    if (rc != WSRT_OK) throw Error(rc, std::string("wsrt: ") + wsrt_last_error());
}

using Vertex = wsrt_vertex;
using MaterialUniform = wsrt_material;
using LightInfo = wsrt_light_info;
using CameraUniform = wsrt_camera_uniform;
using Photon = wsrt_photon;
using Mat4 = std::array<float, 16>;  // column-major, glam's to_cols_array order

// ---- glam-style Mat4 helpers (f32, glam's summation order) ------------------
namespace mat4 {
inline Mat4 identity() { return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }
inline Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r{};
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i) {
            float s = a[0 + i] * b[4 * c + 0];
            s = s + a[4 + i] * b[4 * c + 1];
            s = s + a[8 + i] * b[4 * c + 2];
            s = s + a[12 + i] * b[4 * c + 3];
            r[4 * c + i] = s;
        }
    return r;
}
inline Mat4 from_translation(float x, float y, float z) {
    Mat4 m = identity();
    m[12] = x; m[13] = y; m[14] = z;
    return m;
}
inline Mat4 from_scale(float x, float y, float z) {
    Mat4 m = identity();
    m[0] = x; m[5] = y; m[10] = z;
    return m;
}
inline Mat4 from_rotation_x(float a) {
    Mat4 m = identity();
    float s = std::sin(a), c = std::cos(a);
    m[5] = c; m[6] = s; m[9] = -s; m[10] = c;
    return m;
}
inline Mat4 from_rotation_y(float a) {
    Mat4 m = identity();
    float s = std::sin(a), c = std::cos(a);
    m[0] = c; m[2] = -s; m[8] = s; m[10] = c;
    return m;
}
inline Mat4 from_rotation_z(float a) {
    Mat4 m = identity();
    float s = std::sin(a), c = std::cos(a);
    m[0] = c; m[1] = s; m[4] = -s; m[5] = c;
    return m;
}
}  // namespace mat4

// ---- device (WgpuContext) ----------------------------------------------------
class Device {
   public:
    explicit Device(int ordinal = 0) { check(wsrt_device_create(ordinal, &h_)); }
    ~Device() { wsrt_device_destroy(h_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    wsrt_device* handle() const { return h_; }

   private:
    wsrt_device* h_ = nullptr;
};

namespace scene {

namespace geometry {
using Mesh = std::pair<std::vector<Vertex>, std::vector<uint32_t>>;
inline Mesh fetch(int (*fn)(wsrt_vertex*, uint32_t*, uint32_t*, uint32_t*)) {
    uint32_t nv = 0, ni = 0;
    check(fn(nullptr, nullptr, &nv, &ni));
    Mesh m{std::vector<Vertex>(nv), std::vector<uint32_t>(ni)};
    check(fn(m.first.data(), m.second.data(), &nv, &ni));
    return m;
}
inline Mesh create_plane() { return fetch(wsrt_geometry_plane); }
inline Mesh create_cube() { return fetch(wsrt_geometry_cube); }
inline Mesh create_sphere(uint32_t subdivisions) {
    uint32_t nv = 0, ni = 0;
    check(wsrt_geometry_sphere(subdivisions, nullptr, nullptr, &nv, &ni));
    Mesh m{std::vector<Vertex>(nv), std::vector<uint32_t>(ni)};
    check(wsrt_geometry_sphere(subdivisions, m.first.data(), m.second.data(), &nv, &ni));
    return m;
}
}  // namespace geometry

// SceneResources: device buffers + BVH8, immutable after build.
class SceneResources {
   public:
    explicit SceneResources(wsrt_scene* h) : h_(h) {}
    ~SceneResources() { wsrt_scene_destroy(h_); }
    SceneResources(SceneResources&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    SceneResources(const SceneResources&) = delete;
    SceneResources& operator=(const SceneResources&) = delete;
    const wsrt_scene* handle() const { return h_; }
    wsrt_scene_info info() const {
        wsrt_scene_info i;
        check(wsrt_scene_get_info(h_, &i));
        return i;
    }

   private:
    wsrt_scene* h_ = nullptr;
};

class SceneBuilder {
   public:
    SceneBuilder() { check(wsrt_scene_builder_create(&h_)); }
    explicit SceneBuilder(wsrt_scene_builder* h) : h_(h) {}
    ~SceneBuilder() { wsrt_scene_builder_destroy(h_); }
    SceneBuilder(const SceneBuilder&) = delete;
    SceneBuilder& operator=(const SceneBuilder&) = delete;

    uint32_t add_material(const MaterialUniform& m) {
        uint32_t id = 0;
        check(wsrt_scene_add_material(h_, &m, &id));
        return id;
    }
    uint32_t add_mesh(const std::vector<Vertex>& v, const std::vector<uint32_t>& i) {
        uint32_t id = 0;
        check(wsrt_scene_add_mesh(h_, v.data(), v.size(), i.data(), i.size(), &id));
        return id;
    }
    void add_instance(uint32_t mesh, uint32_t mat, const Mat4& transform) {
        check(wsrt_scene_add_instance(h_, mesh, mat, transform.data()));
    }
    void add_light_instance(uint32_t mesh, uint32_t mat, const Mat4& transform,
                            const std::array<float, 4>& params) {
        check(wsrt_scene_add_light_instance(h_, mesh, mat, transform.data(), params.data()));
    }
    void add_point_light(const std::array<float, 3>& pos, const std::array<float, 4>& params) {
        check(wsrt_scene_add_point_light(h_, pos.data(), params.data()));
    }
    SceneResources build(const Device& dev) {
        wsrt_scene* s = nullptr;
        check(wsrt_scene_build(h_, dev.handle(), &s));
        return SceneResources(s);
    }

   private:
    wsrt_scene_builder* h_ = nullptr;
};

inline SceneResources create_cornell_box(const Device& dev) {
    wsrt_scene_builder* b = nullptr;
    check(wsrt_scene_builder_cornell_box(&b));
    SceneBuilder builder(b);
    return builder.build(dev);
}

// camera.rs:125-194 without the keyboard handling.
struct CameraController {
    std::array<float, 3> position{0.0f, 0.0f, 3.0f};
    float yaw = -90.0f * (3.14159265359f / 180.0f);
    float pitch = 0.0f;
    CameraUniform build_uniform(float aspect, uint32_t frame_count) const {
        CameraUniform u;
        check(wsrt_camera_build_uniform(position.data(), yaw, pitch, aspect, frame_count, &u));
        return u;
    }
};

}  // namespace scene

namespace renderer {

constexpr uint32_t MAX_PHOTONS = 1024 * 1024;  // renderer/mod.rs:18
constexpr uint32_t HASH_SIZE = 4 * 1024 * 1024;  // renderer/mod.rs:19

inline wsrt_config default_config() {
    wsrt_config c;
    check(wsrt_config_default(&c));
    return c;
}

class Renderer {
   public:
    // Renderer::new (mod.rs:45-147). The camera is set per frame with
    // set_camera, which stands for the Engine's camera buffer write.
    Renderer(const Device& dev, const scene::SceneResources& scene, uint32_t target_width,
             uint32_t target_height, const wsrt_config& cfg = default_config())
        : target_width(target_width), target_height(target_height) {
        check(wsrt_renderer_create(dev.handle(), scene.handle(), target_width, target_height, &cfg, &h_));
    }
    ~Renderer() { wsrt_renderer_destroy(h_); }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void set_camera(const CameraUniform& cam) { check(wsrt_renderer_set_camera(h_, &cam)); }
    // mod.rs:168-207: one frame, then frame_count += 1
    void render() { check(wsrt_renderer_render(h_)); }
    void render_frames(uint32_t n) { check(wsrt_renderer_render_frames(h_, n)); }
    void resize(uint32_t w, uint32_t h) { check(wsrt_renderer_resize(h_, w, h)); }
    void sync() { check(wsrt_renderer_sync(h_)); }
    uint32_t frame_count() const {
        uint32_t f = 0;
        check(wsrt_renderer_get_frame_count(h_, &f));
        return f;
    }
    void set_frame_count(uint32_t f) { check(wsrt_renderer_set_frame_count(h_, f)); }
    std::vector<float> read_accum() {
        std::vector<float> a(4ull * target_width * target_height);
        check(wsrt_renderer_read_accum(h_, a.data()));
        return a;
    }
    // storage_texture contents (RGBA8, row 0 = image bottom)
    std::vector<uint8_t> read_rgba8() {
        std::vector<uint8_t> a(4ull * target_width * target_height);
        check(wsrt_renderer_read_rgba8(h_, a.data()));
        return a;
    }
    wsrt_renderer* handle() const { return h_; }

    const uint32_t target_width;
    const uint32_t target_height;

   private:
    wsrt_renderer* h_ = nullptr;
};

}  // namespace renderer

}  // namespace wsrt
