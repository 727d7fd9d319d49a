// api.cpp - the C ABI of libwsrt (include/wsrt.h): devices, the scene
// builder (src/scene/builder.rs), the renderer (src/renderer/mod.rs) and the
// debug taps. Host code; kernels are reached through kernels.hpp.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "shading.hpp"

using namespace wsrt;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_OR_FAIL(expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            if (e_ == hipErrorOutOfMemory)                                                  \
                return fail(WSRT_E_OOM, "%s: %s", #expr, hipGetErrorString(e_));            \
            return fail(WSRT_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));                \
        }                                                                                   \
    } while (0)

#define CHECK_ARG(cond)                                                                     \
    do {                                                                                    \
        if (!(cond)) return fail(WSRT_E_INVALID, "invalid argument: %s", #cond);            \
    } while (0)

extern "C" const char* wsrt_last_error(void) { return g_last_error.c_str(); }
extern "C" int wsrt_abi_version(void) { return WSRT_ABI_VERSION; }

// ---------------------------------------------------------------------------
// device memory helper
// ---------------------------------------------------------------------------
namespace {

// Device memory the library's objects hold, per device ordinal, and the cap
// wsrt_device_set_memory_limit puts on it (0: none). An allocation past the
// cap fails as hipMalloc would when the device is full.
constexpr int kMaxOrdinals = 64;
std::atomic<uint64_t> g_mem_used[kMaxOrdinals];
std::atomic<uint64_t> g_mem_limit[kMaxOrdinals];

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t held = 0;  // bytes counted against the device (the allocation's size)
    int ordinal = -1;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), held(o.held), ordinal(o.ordinal) {
        o.p = nullptr;
        o.bytes = o.held = 0;
        o.ordinal = -1;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            held = o.held;
            ordinal = o.ordinal;
            o.p = nullptr;
            o.bytes = o.held = 0;
            o.ordinal = -1;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            if (ordinal >= 0) g_mem_used[ordinal] -= held;
        }
        p = nullptr;
        bytes = held = 0;
        ordinal = -1;
    }
    // On the current device.
    hipError_t alloc(size_t n) {
        release();
        int d = -1;
        if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxOrdinals) d = -1;
        const size_t size = n < 256 ? 256 : n;
        if (d >= 0) {
            const uint64_t used = g_mem_used[d].fetch_add(size) + size, cap = g_mem_limit[d].load();
            if (cap && used > cap) {
                g_mem_used[d] -= size;
                return hipErrorOutOfMemory;
            }
        }
        const hipError_t e = hipMalloc(&p, size);
        if (e != hipSuccess) {
            p = nullptr;
            if (d >= 0) g_mem_used[d] -= size;
            return e;
        }
        bytes = n;
        held = size;
        ordinal = d;
        return hipSuccess;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

void mat4_mul_vec4(const float* a, const float* v, float* out) {
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float s = a[0 + i] * v[0];
        s = s + a[4 + i] * v[1];
        s = s + a[8 + i] * v[2];
        s = s + a[12 + i] * v[3];
        r[i] = s;
    }
    memcpy(out, r, sizeof r);
}
void mat4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int c = 0; c < 4; ++c) mat4_mul_vec4(a, b + 4 * c, r + 4 * c);
    memcpy(out, r, sizeof r);
}

}  // namespace

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------
// Work memory of wsrt_photon_grid_build_noalloc: one copy per GPU, held
// until the process ends. Host code uses only the device address that
// hipGetSymbolAddress returns for it (wsrt_device::scratch), never the name.
constexpr uint32_t kNoAllocMaxPhotons = 1u << 20;  // MAX_PHOTONS, renderer/mod.rs:18
constexpr uint32_t kNoAllocMaxCells = 1u << 22;    // HASH_SIZE, renderer/mod.rs:19
constexpr size_t kScratchBytes = 32ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct wsrt_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;  // builds and debug queries
    char name[256] = {0};
    std::atomic<int> live{0};      // scenes and renderers created on this device
    uint8_t* scratch = nullptr;    // this GPU's copy of the file-scope scratch array
};

extern "C" int wsrt_device_set_memory_limit(wsrt_device* dev, uint64_t bytes) {
    CHECK_ARG(dev && dev->ordinal >= 0 && dev->ordinal < kMaxOrdinals);
    g_mem_limit[dev->ordinal] = bytes;
    return WSRT_OK;
}

extern "C" int wsrt_device_get_memory_used(const wsrt_device* dev, uint64_t* out) {
    CHECK_ARG(dev && out && dev->ordinal >= 0 && dev->ordinal < kMaxOrdinals);
    *out = g_mem_used[dev->ordinal].load();
    return WSRT_OK;
}

extern "C" int wsrt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int wsrt_device_create(int ordinal, wsrt_device** out) {
    CHECK_ARG(out);
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(WSRT_E_UNSUPPORTED, "no HIP device available");
    if (ordinal < 0 || ordinal >= n) return fail(WSRT_E_INVALID, "device ordinal %d of %d", ordinal, n);
    HIP_OR_FAIL(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    HIP_OR_FAIL(hipGetDeviceProperties(&prop, ordinal));
    auto d = std::make_unique<wsrt_device>();
    d->ordinal = ordinal;
    snprintf(d->name, sizeof d->name, "%s", prop.gcnArchName);
    HIP_OR_FAIL(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    // this device's address of the scratch array (the device is current)
    void* sp = nullptr;
    HIP_OR_FAIL(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    d->scratch = static_cast<uint8_t*>(sp);
    *out = d.release();
    return WSRT_OK;
}

extern "C" int wsrt_device_destroy(wsrt_device* dev) {
    if (!dev) return WSRT_OK;
    if (dev->live.load() > 0)
        return fail(WSRT_E_STATE, "device still has %d live scenes/renderers; destroy them first",
                    dev->live.load());
    (void)hipSetDevice(dev->ordinal);
    if (dev->stream) (void)hipStreamDestroy(dev->stream);
    delete dev;
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// geometry generators (src/scene/geometry.rs)
// ---------------------------------------------------------------------------
static wsrt_vertex vtx(float px, float py, float pz, float nx, float ny, float nz) {
    wsrt_vertex v;
    v.pos[0] = px; v.pos[1] = py; v.pos[2] = pz; v.pos[3] = 1.0f;
    v.normal[0] = nx; v.normal[1] = ny; v.normal[2] = nz; v.normal[3] = 0.0f;
    return v;
}

static void gen_plane(std::vector<wsrt_vertex>& v, std::vector<uint32_t>& i) {
    // 1 x 1 quad on the XZ plane facing +Y (geometry.rs:12-35)
    v = {vtx(-0.5f, 0, 0.5f, 0, 1, 0), vtx(0.5f, 0, 0.5f, 0, 1, 0), vtx(-0.5f, 0, -0.5f, 0, 1, 0),
         vtx(0.5f, 0, -0.5f, 0, 1, 0)};
    i = {0, 1, 2, 2, 1, 3};
}

static void gen_cube(std::vector<wsrt_vertex>& v, std::vector<uint32_t>& idx) {
    // six quads, outward normals, two triangles each (geometry.rs:38-118)
    struct Face {
        float n[3];
        float c[4][3];
    };
    static const Face faces[6] = {
        {{0, 0, 1}, {{-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}}},
        {{0, 0, -1}, {{0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}}},
        {{0, 1, 0}, {{-0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, 0.5f}, {0.5f, 0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}}},
        {{0, -1, 0}, {{-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, 0.5f}, {-0.5f, -0.5f, 0.5f}}},
        {{1, 0, 0}, {{0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, 0.5f}}},
        {{-1, 0, 0}, {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f}}},
    };
    v.clear();
    idx.clear();
    for (const Face& f : faces) {
        uint32_t b = (uint32_t)v.size();
        for (const auto& c : f.c) v.push_back(vtx(c[0], c[1], c[2], f.n[0], f.n[1], f.n[2]));
        const uint32_t q[6] = {b, b + 1, b + 2, b, b + 2, b + 3};
        idx.insert(idx.end(), q, q + 6);
    }
}

static void gen_sphere(uint32_t subdiv, std::vector<wsrt_vertex>& v, std::vector<uint32_t>& idx) {
    // icosphere of radius 0.5 (geometry.rs:121-232): 12 icosahedron vertices,
    // each level splits every face into four at cached edge midpoints that are
    // pushed back onto the sphere
    v.clear();
    auto push = [&v](float x, float y, float z) {
        float len = sqrtf(x * x + y * y + z * z);
        float nx = x / len, ny = y / len, nz = z / len;
        v.push_back(vtx(nx * 0.5f, ny * 0.5f, nz * 0.5f, nx, ny, nz));
        return (uint32_t)v.size() - 1;
    };
    const float t = (1.0f + sqrtf(5.0f)) / 2.0f;
    const float ico[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (const auto& p : ico) push(p[0], p[1], p[2]);
    std::vector<std::array<uint32_t, 3>> faces = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
        auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        const float* pa = v[a].pos;
        const float* pb = v[b].pos;
        uint32_t id = push((pa[0] + pb[0]) * 0.5f, (pa[1] + pb[1]) * 0.5f, (pa[2] + pb[2]) * 0.5f);
        mid.emplace(key, id);
        return id;
    };
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::vector<std::array<uint32_t, 3>> next;
        next.reserve(faces.size() * 4);
        for (const auto& f : faces) {
            uint32_t a = midpoint(f[0], f[1]);
            uint32_t b = midpoint(f[1], f[2]);
            uint32_t c = midpoint(f[2], f[0]);
            next.push_back({f[0], a, c});
            next.push_back({f[1], b, a});
            next.push_back({f[2], c, b});
            next.push_back({a, b, c});
        }
        faces.swap(next);
    }
    idx.clear();
    for (const auto& f : faces) idx.insert(idx.end(), f.begin(), f.end());
}

static int copy_geometry(const std::vector<wsrt_vertex>& v, const std::vector<uint32_t>& i,
                         wsrt_vertex* verts, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    CHECK_ARG(nv && ni);
    if (verts && idx) {
        if (*nv < v.size() || *ni < i.size())
            return fail(WSRT_E_INVALID, "geometry buffers too small (%u/%zu vertices, %u/%zu indices)",
                        *nv, v.size(), *ni, i.size());
        memcpy(verts, v.data(), v.size() * sizeof(wsrt_vertex));
        memcpy(idx, i.data(), i.size() * sizeof(uint32_t));
    }
    *nv = (uint32_t)v.size();
    *ni = (uint32_t)i.size();
    return WSRT_OK;
}

extern "C" int wsrt_geometry_plane(wsrt_vertex* verts, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    std::vector<wsrt_vertex> v;
    std::vector<uint32_t> i;
    gen_plane(v, i);
    return copy_geometry(v, i, verts, idx, nv, ni);
}
extern "C" int wsrt_geometry_cube(wsrt_vertex* verts, uint32_t* idx, uint32_t* nv, uint32_t* ni) {
    std::vector<wsrt_vertex> v;
    std::vector<uint32_t> i;
    gen_cube(v, i);
    return copy_geometry(v, i, verts, idx, nv, ni);
}
extern "C" int wsrt_geometry_sphere(uint32_t subdiv, wsrt_vertex* verts, uint32_t* idx, uint32_t* nv,
                                    uint32_t* ni) {
    if (subdiv > 8) return fail(WSRT_E_INVALID, "sphere subdivisions %u > 8", subdiv);
    std::vector<wsrt_vertex> v;
    std::vector<uint32_t> i;
    gen_sphere(subdiv, v, i);
    return copy_geometry(v, i, verts, idx, nv, ni);
}

// ---------------------------------------------------------------------------
// scene builder (src/scene/builder.rs:47-266)
// ---------------------------------------------------------------------------
struct MeshRec {
    uint32_t vertex_offset, index_offset, n_verts, n_idx;
};
struct InstRec {
    uint32_t mesh, mat;
    float m[16];
};

struct wsrt_scene_builder {
    std::vector<wsrt_vertex> verts;
    std::vector<uint32_t> idx;
    std::vector<MeshRec> meshes;
    std::vector<wsrt_material> materials;
    std::vector<InstRec> instances;
    std::vector<wsrt_light_info> lights;
    std::vector<float> env;
    uint32_t env_w = 0, env_h = 0;
};

extern "C" int wsrt_scene_builder_create(wsrt_scene_builder** out) {
    CHECK_ARG(out);
    *out = new (std::nothrow) wsrt_scene_builder();
    return *out ? WSRT_OK : fail(WSRT_E_OOM, "scene builder allocation failed");
}
extern "C" int wsrt_scene_builder_destroy(wsrt_scene_builder* b) {
    delete b;
    return WSRT_OK;
}
extern "C" int wsrt_scene_add_material(wsrt_scene_builder* b, const wsrt_material* m, uint32_t* id) {
    CHECK_ARG(b && m);
    uint32_t t = (uint32_t)m->extra[0];
    if (!(m->extra[0] >= 0.0f) || t > 5u)
        return fail(WSRT_E_INVALID, "material type %g not in 0..5", (double)m->extra[0]);
    if (id) *id = (uint32_t)b->materials.size();
    b->materials.push_back(*m);
    return WSRT_OK;
}
extern "C" int wsrt_scene_add_mesh(wsrt_scene_builder* b, const wsrt_vertex* verts, size_t n_verts,
                                   const uint32_t* idx, size_t n_idx, uint32_t* id) {
    CHECK_ARG(b && (verts || n_verts == 0) && (idx || n_idx == 0));
    if (n_idx % 3) return fail(WSRT_E_INVALID, "index count %zu is not a multiple of 3", n_idx);
    for (size_t k = 0; k < n_idx; ++k)
        if (idx[k] >= n_verts)
            return fail(WSRT_E_INVALID, "index %u at %zu out of range (%zu vertices)", idx[k], k, n_verts);
    if (b->verts.size() + n_verts > 0xFFFFFFFFull || b->idx.size() + n_idx > 0xFFFFFFFFull)
        return fail(WSRT_E_INVALID, "scene exceeds 2^32 vertices or indices");
    MeshRec r{(uint32_t)b->verts.size(), (uint32_t)b->idx.size(), (uint32_t)n_verts, (uint32_t)n_idx};
    if (id) *id = (uint32_t)b->meshes.size();
    b->meshes.push_back(r);
    b->verts.insert(b->verts.end(), verts, verts + n_verts);
    b->idx.insert(b->idx.end(), idx, idx + n_idx);
    return WSRT_OK;
}
extern "C" int wsrt_scene_add_instance(wsrt_scene_builder* b, uint32_t mesh, uint32_t mat,
                                       const float transform[16]) {
    CHECK_ARG(b && transform);
    if (mesh >= b->meshes.size()) return fail(WSRT_E_INVALID, "mesh %u of %zu", mesh, b->meshes.size());
    if (mat >= b->materials.size())
        return fail(WSRT_E_INVALID, "material %u of %zu", mat, b->materials.size());
    InstRec r;
    r.mesh = mesh;
    r.mat = mat;
    memcpy(r.m, transform, sizeof r.m);
    b->instances.push_back(r);
    return WSRT_OK;
}
extern "C" int wsrt_scene_add_light_instance(wsrt_scene_builder* b, uint32_t mesh, uint32_t mat,
                                             const float transform[16], const float params[4]) {
    CHECK_ARG(params);
    int rc = wsrt_scene_add_instance(b, mesh, mat, transform);
    if (rc) return rc;
    // one world-space light per triangle (builder.rs:110-128)
    const MeshRec& mr = b->meshes[mesh];
    for (uint32_t i = 0; i + 2 < mr.n_idx; i += 3) {
        wsrt_light_info L;
        float* dst[3] = {L.v0, L.v1, L.v2};
        for (int k = 0; k < 3; ++k) {
            const float* p = b->verts[mr.vertex_offset + b->idx[mr.index_offset + i + k]].pos;
            float v[4] = {p[0], p[1], p[2], 1.0f}, r[4];
            mat4_mul_vec4(transform, v, r);
            dst[k][0] = r[0]; dst[k][1] = r[1]; dst[k][2] = r[2]; dst[k][3] = 1.0f;
        }
        memcpy(L.params, params, sizeof L.params);
        b->lights.push_back(L);
    }
    return WSRT_OK;
}
extern "C" int wsrt_scene_add_point_light(wsrt_scene_builder* b, const float position[3],
                                          const float params[4]) {
    CHECK_ARG(b && position && params);
    wsrt_light_info L;
    memset(&L, 0, sizeof L);
    L.v0[0] = position[0]; L.v0[1] = position[1]; L.v0[2] = position[2]; L.v0[3] = 1.0f;
    memcpy(L.params, params, sizeof L.params);
    L.params[0] = 1.0f;
    b->lights.push_back(L);
    return WSRT_OK;
}
extern "C" int wsrt_scene_set_environment(wsrt_scene_builder* b, const float* rgb, uint32_t w,
                                          uint32_t h) {
    CHECK_ARG(b && rgb && w > 0 && h > 0 && (uint64_t)w * h <= (1ull << 26));
    b->env.assign(rgb, rgb + (size_t)w * h * 3);
    b->env_w = w;
    b->env_h = h;
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
struct wsrt_scene {
    wsrt_device* dev = nullptr;
    std::atomic<int> renderers{0};
    DevBuf verts, idx, meshes, materials, lights, instances, env;
    DevBuf nodes, tris, wv, tri_inst, codes, sorted;
    DevBuf spec_instances, spec_nodes, spec_tris;
    SceneView view{};
    wsrt_scene_info info{};
    uint32_t n_point_lights = 0;  // lights of kind 1 (the guided light proposal needs none)
};

static void linear_inverse(const float* m, float* w) {
    // inverse of the 3x3 linear part in f64, rounded to f32 (column-major)
    double a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double D = -(b * i - c * h), E = a * i - c * g, F = -(a * h - b * g);
    double G = b * f - c * e, H = -(a * f - c * d), I = a * e - b * d;
    double det = a * A + b * B + c * C;
    double r[3][3] = {{A / det, D / det, G / det}, {B / det, E / det, H / det}, {C / det, F / det, I / det}};
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) w[col * 3 + row] = (float)r[row][col];
}

template <class T>
static hipError_t upload(DevBuf& buf, const std::vector<T>& v, hipStream_t s) {
    hipError_t e = buf.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

extern "C" int wsrt_scene_build(wsrt_scene_builder* b, wsrt_device* dev, wsrt_scene** out) {
    CHECK_ARG(b && dev && out);
    *out = nullptr;
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    auto sc = std::make_unique<wsrt_scene>();
    sc->dev = dev;
    hipStream_t s = dev->stream;
    // instance table with world-triangle ranges and world-to-object matrices
    std::vector<Instance> inst(b->instances.size());
    uint64_t n_tris = 0;
    for (size_t k = 0; k < b->instances.size(); ++k) {
        const InstRec& r = b->instances[k];
        Instance& I = inst[k];
        memset(&I, 0, sizeof I);
        I.mesh = r.mesh;
        I.material = r.mat;
        I.first_tri = (uint32_t)n_tris;
        I.n_tris = b->meshes[r.mesh].n_idx / 3;
        memcpy(I.m, r.m, sizeof I.m);
        linear_inverse(r.m, I.w2o);
        n_tris += I.n_tris;
    }
    if (n_tris >= 0x7FFFFFFFull) return fail(WSRT_E_INVALID, "scene has %llu triangles (max 2^31-1)",
                                             (unsigned long long)n_tris);
    std::vector<wsrt_mesh_info> mi(b->meshes.size());
    for (size_t k = 0; k < b->meshes.size(); ++k) {
        mi[k].vertex_offset = b->meshes[k].vertex_offset;
        mi[k].index_offset = b->meshes[k].index_offset;
        mi[k].pad[0] = mi[k].pad[1] = 0;
    }
    HIP_OR_FAIL(upload(sc->verts, b->verts, s));
    HIP_OR_FAIL(upload(sc->idx, b->idx, s));
    HIP_OR_FAIL(upload(sc->meshes, mi, s));
    HIP_OR_FAIL(upload(sc->materials, b->materials, s));
    HIP_OR_FAIL(upload(sc->lights, b->lights, s));
    for (const wsrt_light_info& L : b->lights) sc->n_point_lights += L.params[0] == 1.0f ? 1u : 0u;
    HIP_OR_FAIL(upload(sc->instances, inst, s));
    HIP_OR_FAIL(upload(sc->env, b->env, s));
    hipEvent_t e0, e1;
    HIP_OR_FAIL(hipEventCreate(&e0));
    HIP_OR_FAIL(hipEventCreate(&e1));
    HIP_OR_FAIL(hipEventRecord(e0, s));
    BvhInputs bi;
    bi.instances = sc->instances.as<Instance>();
    bi.n_instances = (uint32_t)inst.size();
    bi.verts = sc->verts.as<wsrt_vertex>();
    bi.indices = sc->idx.as<uint32_t>();
    bi.meshes = sc->meshes.as<wsrt_mesh_info>();
    bi.n_tris = (uint32_t)n_tris;
    BvhOutputs bo;
    int rc = bvh_build(bi, &bo, s);
    HIP_OR_FAIL(hipEventRecord(e1, s));
    HIP_OR_FAIL(hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == -2) return fail(WSRT_E_OOM, "BVH build: device allocation failed");
    if (rc != 0) return fail(WSRT_E_HIP, "BVH build failed (%d): %s", rc, hipGetErrorString(hipGetLastError()));
    if (bo.placed_tris != n_tris)
        return fail(WSRT_E_HIP, "BVH build placed %u of %llu triangles", bo.placed_tris,
                    (unsigned long long)n_tris);
    sc->nodes.p = bo.nodes;
    sc->nodes.bytes = (size_t)bo.n_nodes * sizeof(wsrt_bvh8_node);
    sc->tris.p = bo.tris;
    sc->tris.bytes = (size_t)n_tris * sizeof(wsrt_bvh_tri);
    sc->wv.p = bo.wv;
    sc->wv.bytes = (size_t)n_tris * 9 * sizeof(float);
    sc->tri_inst.p = bo.tri_inst;
    sc->tri_inst.bytes = (size_t)n_tris * sizeof(uint32_t);
    sc->codes.p = bo.codes;
    sc->codes.bytes = (size_t)n_tris * sizeof(uint32_t);
    sc->sorted.p = bo.sorted;
    sc->sorted.bytes = (size_t)n_tris * sizeof(uint32_t);
    SceneView& v = sc->view;
    v.nodes = sc->nodes.as<wsrt_bvh8_node>();
    v.tris = sc->tris.as<float4>();
    v.verts = sc->verts.as<wsrt_vertex>();
    v.indices = sc->idx.as<uint32_t>();
    v.meshes = sc->meshes.as<wsrt_mesh_info>();
    v.materials = sc->materials.as<wsrt_material>();
    v.lights = sc->lights.as<wsrt_light_info>();
    v.instances = sc->instances.as<Instance>();
    v.tri_instance = sc->tri_inst.as<uint32_t>();
    v.env = b->env.empty() ? nullptr : sc->env.as<float>();
    v.n_nodes = bo.n_nodes;
    v.n_tris = (uint32_t)n_tris;
    v.n_lights = (uint32_t)b->lights.size();
    v.n_instances = (uint32_t)inst.size();
    v.env_w = b->env_w;
    v.env_h = b->env_h;
    v.stack_size = 1u + 7u * bo.max_depth;
    // the specular BVH: instances of metal (1), glass (2) and thin-film (4)
    // materials, numbered from 0 (the photon pass's first-bounce filter)
    {
        std::vector<Instance> spec;
        uint32_t n_spec = 0;
        for (const Instance& in : inst) {
            const uint32_t type = (uint32_t)b->materials[in.material].extra[0];
            if (type != 1u && type != 2u && type != 4u) continue;
            Instance c = in;
            c.first_tri = n_spec;
            n_spec += c.n_tris;
            spec.push_back(c);
        }
        BvhOutputs so;
        so.n_nodes = 0;
        so.max_depth = 0;
        so.nodes = nullptr;
        so.tris = nullptr;
        if (n_spec) {
            HIP_OR_FAIL(upload(sc->spec_instances, spec, s));
            BvhInputs sb = bi;
            sb.instances = sc->spec_instances.as<Instance>();
            sb.n_instances = (uint32_t)spec.size();
            sb.n_tris = n_spec;
            rc = bvh_build(sb, &so, s);
            HIP_OR_FAIL(hipStreamSynchronize(s));
            // only the nodes and leaf triangles are kept
            for (void* q : {(void*)so.wv, (void*)so.tri_inst, (void*)so.codes, (void*)so.sorted})
                if (q) (void)hipFree(q);
            if (rc == -2) return fail(WSRT_E_OOM, "specular BVH build: device allocation failed");
            if (rc != 0 || so.placed_tris != n_spec) return fail(WSRT_E_HIP, "specular BVH build failed (%d)", rc);
            sc->spec_nodes.p = so.nodes;
            sc->spec_nodes.bytes = (size_t)so.n_nodes * sizeof(wsrt_bvh8_node);
            sc->spec_tris.p = so.tris;
            sc->spec_tris.bytes = (size_t)n_spec * sizeof(wsrt_bvh_tri);
        }
        v.spec_n_nodes = so.n_nodes;
        v.spec_nodes = sc->spec_nodes.as<wsrt_bvh8_node>();
        v.spec_tris = sc->spec_tris.as<float4>();
        v.stack_size = std::max(v.stack_size, 1u + 7u * so.max_depth);
    }
    wsrt_scene_info& I = sc->info;
    I.n_triangles = (uint32_t)n_tris;
    I.n_instances = (uint32_t)inst.size();
    I.n_lights = (uint32_t)b->lights.size();
    I.n_materials = (uint32_t)b->materials.size();
    I.n_nodes = bo.n_nodes;
    I.max_depth = bo.max_depth;
    I.stack_size = 1u + 7u * bo.max_depth;
    I.node_bytes = sizeof(wsrt_bvh8_node);
    I.tri_bytes = sizeof(wsrt_bvh_tri);
    I.build_ms = ms;
    if (bo.n_nodes) {
        wsrt_bvh8_node root;
        HIP_OR_FAIL(hipMemcpy(&root, bo.nodes, sizeof root, hipMemcpyDeviceToHost));
        for (int a = 0; a < 3; ++a) {
            I.bounds[a] = root.origin[a];
            float s = ldexpf(1.0f, root.exp[a]);
            float hi = root.origin[a];
            for (int c = 0; c < 8; ++c)
                if (root.meta[c]) hi = fmaxf(hi, root.origin[a] + (float)root.qhi[a][c] * s);
            I.bounds[3 + a] = hi;
        }
    }
    dev->live++;
    *out = sc.release();
    return WSRT_OK;
}

extern "C" int wsrt_scene_destroy(wsrt_scene* s) {
    if (!s) return WSRT_OK;
    if (s->renderers.load() > 0)
        return fail(WSRT_E_STATE, "scene still has %d live renderers; destroy them first",
                    s->renderers.load());
    s->dev->live--;
    (void)hipSetDevice(s->dev->ordinal);
    (void)hipStreamSynchronize(s->dev->stream);
    delete s;
    return WSRT_OK;
}

extern "C" int wsrt_scene_get_info(const wsrt_scene* s, wsrt_scene_info* out) {
    CHECK_ARG(s && out);
    *out = s->info;
    return WSRT_OK;
}

// cornell_box.rs:6-110
static void trs(float tx, float ty, float tz, int axis, float ang, float sx, float sy, float sz,
                float* out) {
    float T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, tx, ty, tz, 1};
    float R[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float S[16] = {sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, sz, 0, 0, 0, 0, 1};
    float sn = sinf(ang), cs = cosf(ang);
    if (axis == 0) { R[5] = cs; R[6] = sn; R[9] = -sn; R[10] = cs; }
    if (axis == 1) { R[0] = cs; R[2] = -sn; R[8] = sn; R[10] = cs; }
    if (axis == 2) { R[0] = cs; R[1] = sn; R[4] = -sn; R[5] = cs; }
    float TR[16];
    if (axis < 0) {
        mat4_mul(T, S, out);
        return;
    }
    mat4_mul(T, R, TR);
    mat4_mul(TR, S, out);
}

extern "C" int wsrt_scene_builder_cornell_box(wsrt_scene_builder** out) {
    CHECK_ARG(out);
    int rc = wsrt_scene_builder_create(out);
    if (rc) return rc;
    wsrt_scene_builder* b = *out;
    const float PI_F = 3.14159265358979323846f, HALF_PI = 1.57079632679489661923f;
    const wsrt_material mats[6] = {
        {{6500.0f, 1.0f, 0.0f, 1.0f}, {3, 0, 0, 0}},           // light
        {{0.65f, 0.05f, 0.05f, 1}, {0, 0, 0, 0}},              // red
        {{0.12f, 0.45f, 0.15f, 1}, {0, 0, 0, 0}},              // green
        {{0.73f, 0.73f, 0.73f, 1}, {0, 0, 0, 0}},              // white
        {{1, 1, 1, 1}, {2, 0, 1.5f, 0.01f}},                   // glass
        {{1, 1, 1, 1}, {4, 500.0f, 1.33f, 1.0f}},              // bubble
    };
    uint32_t id[6];
    for (int k = 0; k < 6; ++k) wsrt_scene_add_material(b, &mats[k], &id[k]);
    std::vector<wsrt_vertex> v;
    std::vector<uint32_t> i;
    uint32_t plane, cube, sphere;
    gen_plane(v, i);
    wsrt_scene_add_mesh(b, v.data(), v.size(), i.data(), i.size(), &plane);
    gen_cube(v, i);
    wsrt_scene_add_mesh(b, v.data(), v.size(), i.data(), i.size(), &cube);
    gen_sphere(4, v, i);
    wsrt_scene_add_mesh(b, v.data(), v.size(), i.data(), i.size(), &sphere);
    float M[16];
    trs(0, -1, 0, -1, 0, 2, 2, 2, M);
    wsrt_scene_add_instance(b, plane, id[3], M);
    trs(0, 1, 0, 0, PI_F, 2, 2, 2, M);
    wsrt_scene_add_instance(b, plane, id[3], M);
    trs(0, 0, -1, 0, HALF_PI, 2, 2, 2, M);
    wsrt_scene_add_instance(b, plane, id[3], M);
    trs(-1, 0, 0, 2, -HALF_PI, 2, 2, 2, M);
    wsrt_scene_add_instance(b, plane, id[1], M);
    trs(1, 0, 0, 2, HALF_PI, 2, 2, 2, M);
    wsrt_scene_add_instance(b, plane, id[2], M);
    trs(0, 0.99f, 0, 0, PI_F, 0.5f, 0.5f, 0.5f, M);
    const float lp[4] = {0.0f, 6500.0f, 1.0f, 0.0f};
    wsrt_scene_add_light_instance(b, plane, id[0], M, lp);
    trs(-0.35f, -0.4f + 0.002f, -0.3f, 1, 0.4f, 0.6f, 1.2f, 0.6f, M);
    wsrt_scene_add_instance(b, cube, id[4], M);
    trs(0.4f, -0.64f, 0.3f, -1, 0, 0.75f, 0.75f, 0.75f, M);
    wsrt_scene_add_instance(b, sphere, id[5], M);
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// camera (src/scene/camera.rs:178-194)
// ---------------------------------------------------------------------------
static void inverse4(const float* mf, float* out) {
    // cofactor expansion in f64, rounded to f32
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = mf[i];
    auto cof = [&](int r, int c) {
        double a[9];
        int k = 0;
        for (int cc = 0; cc < 4; ++cc) {
            if (cc == c) continue;
            for (int rr = 0; rr < 4; ++rr) {
                if (rr == r) continue;
                a[k++] = m[cc * 4 + rr];
            }
        }
        double det3 = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                      a[6] * (a[1] * a[5] - a[4] * a[2]);
        return ((r + c) & 1) ? -det3 : det3;
    };
    double det = 0.0;
    for (int c = 0; c < 4; ++c) det += m[c * 4 + 0] * cof(0, c);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[r * 4 + c] = cof(r, c) / det;  // transpose of cofactors
    for (int i = 0; i < 16; ++i) out[i] = (float)inv[i];
}

extern "C" int wsrt_camera_build_uniform(const float position[3], float yaw, float pitch,
                                         float aspect, uint32_t frame_count,
                                         wsrt_camera_uniform* out) {
    CHECK_ARG(position && out && aspect > 0.0f);
    float sy = sinf(yaw), cy = cosf(yaw), sp = sinf(pitch), cp = cosf(pitch);
    float fw[3] = {cp * cy, sp, cp * sy};
    float il = 1.0f / sqrtf(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
    for (float& x : fw) x = x * il;
    const float* eye = position;
    float f[3] = {(eye[0] + fw[0]) - eye[0], (eye[1] + fw[1]) - eye[1], (eye[2] + fw[2]) - eye[2]};
    il = 1.0f / sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& x : f) x = x * il;
    // s = f x up, u = s x f (up = +Y)
    float s[3] = {f[1] * 0.0f - f[2] * 1.0f, f[2] * 0.0f - f[0] * 0.0f, f[0] * 1.0f - f[1] * 0.0f};
    il = 1.0f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (float& x : s) x = x * il;
    float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    auto dot3 = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    float view[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                      -dot3(eye, s), -dot3(eye, u), dot3(eye, f), 1};
    float fov = 45.0f * (3.14159265359f / 180.0f);
    float sf = sinf(0.5f * fov), cf = cosf(0.5f * fov);
    float h = cf / sf, w = h / aspect, r = 100.0f / (0.1f - 100.0f);
    float proj[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1, 0, 0, r * 0.1f, 0};
    memcpy(out->view, view, sizeof view);
    memcpy(out->proj, proj, sizeof proj);
    inverse4(view, out->view_inverse);
    inverse4(proj, out->proj_inverse);
    out->frame_count = frame_count;
    out->pad[0] = out->pad[1] = out->pad[2] = 0;
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// config
// ---------------------------------------------------------------------------
extern "C" int wsrt_config_default(wsrt_config* c) {
    CHECK_ARG(c);
    memset(c, 0, sizeof *c);
    c->struct_size = sizeof(wsrt_config);
    c->max_depth = 12;
    c->photon_max_depth = 5;
    c->photons_per_pass = 1u << 20;
    c->hash_size = 1u << 22;
    c->cell_size = 0.02f;
    c->gather_radius = 0.02f;
    c->gather_radius_exponent = -0.15f;
    c->spectral_radius = 20.0f;
    c->gather_cap = 64;
    c->rr_after = 3;
    c->enable_photons = 1;
    c->photon_norm = 1u << 20;
    c->guiding = 0;
    c->guide_mix = 0.5f;
    c->guide_lr = 1e-3f;
    c->guide_train_samples = 1u << 18;
    c->env_map = 0;
    c->frame_graphs = 1;
    c->frames_in_flight = 8;
    c->tail_paths = 1u << 18;
    return WSRT_OK;
}

static int validate_config(const wsrt_config* c) {
    if (c->struct_size != sizeof(wsrt_config))
        return fail(WSRT_E_INVALID, "wsrt_config.struct_size %u != %zu", c->struct_size, sizeof(wsrt_config));
    if (c->max_depth == 0 || c->max_depth > 64) return fail(WSRT_E_INVALID, "max_depth %u", c->max_depth);
    if (c->photon_max_depth > 64) return fail(WSRT_E_INVALID, "photon_max_depth %u", c->photon_max_depth);
    if (c->photons_per_pass > (1u << 24))
        return fail(WSRT_E_INVALID, "photons_per_pass %u > 2^24", c->photons_per_pass);
    if (c->hash_size == 0 || (c->hash_size & (c->hash_size - 1)) || c->hash_size > (1u << 26))
        return fail(WSRT_E_INVALID, "hash_size %u must be a power of two <= 2^26", c->hash_size);
    if (!(c->cell_size > 0.0f) || !(c->gather_radius > 0.0f) || !(c->spectral_radius > 0.0f))
        return fail(WSRT_E_INVALID, "cell_size, gather_radius and spectral_radius must be > 0");
    if (c->photon_norm == 0) return fail(WSRT_E_INVALID, "photon_norm must be > 0");
    if (c->frames_in_flight == 0 || c->frames_in_flight > 256)
        return fail(WSRT_E_INVALID, "frames_in_flight %u not in 1..256", c->frames_in_flight);
    if ((uint64_t)c->photons_per_pass * c->frames_in_flight > (1ull << 28) ||
        (uint64_t)c->hash_size * c->frames_in_flight > (1ull << 30))
        return fail(WSRT_E_INVALID, "photons or cells times frames_in_flight too large");
    if (c->guiding > 1) return fail(WSRT_E_INVALID, "guiding %u not 0 or 1", c->guiding);
    if (c->trace_rounds > 4) return fail(WSRT_E_INVALID, "trace_rounds %u not in 0..4", c->trace_rounds);
    if (c->extend_refill > 64) return fail(WSRT_E_INVALID, "extend_refill %u not in 0..64", c->extend_refill);
    if (c->ordered_queues > 1) return fail(WSRT_E_INVALID, "ordered_queues %u not 0 or 1", c->ordered_queues);
    const uint32_t known = WSRT_FIX_EMISSION_MIS | WSRT_FIX_RR_CLAMP | WSRT_FIX_ONB_SIGN;
    if (c->quirks & ~known) return fail(WSRT_E_INVALID, "quirks 0x%x has unknown bits", c->quirks);
    if (c->guiding) {
        if (!(c->guide_mix >= 0.0f && c->guide_mix <= 1.0f))
            return fail(WSRT_E_INVALID, "guide_mix %g not in [0, 1]", (double)c->guide_mix);
        if (!(c->guide_lr > 0.0f)) return fail(WSRT_E_INVALID, "guide_lr must be > 0");
        if (c->guide_train_samples < 64 || c->guide_train_samples > (1u << 22))
            return fail(WSRT_E_INVALID, "guide_train_samples %u not in 64..2^22", c->guide_train_samples);
    }
    return WSRT_OK;
}

// Host-side constants of one frame (hero-wavelength CDF, gather radius and
// density, grid jitter, camera unprojection matrix).
static FrameConst make_frame_const(const wsrt_config& c, const wsrt_camera_uniform& cam, uint32_t W,
                                   uint32_t H) {
    FrameConst fc;
    memset(&fc, 0, sizeof fc);
    const float pi = 3.14159265359f;
    float mn = (1.0f / pi) * atanf((380.0f - 555.0f) / 40.0f) + 0.5f;
    float mx = (1.0f / pi) * atanf((780.0f - 555.0f) / 40.0f) + 0.5f;
    fc.cdf_min = mn;
    fc.cdf_range = mx - mn;
    const uint32_t frame = cam.frame_count;
    fc.radius = c.gather_radius * powf((float)(frame + 1u), c.gather_radius_exponent);
    fc.radius_sq = fc.radius * fc.radius;
    fc.density = (3.0f / (pi * fc.radius_sq)) * (1.0f / c.spectral_radius) * (1.0f / (float)c.photon_norm);
    uint32_t state = (frame * 114514u) * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    uint32_t seed = (word >> 22u) ^ word;
    float rx = (float)(seed & 0xFFu) / 255.0f, ry = (float)((seed >> 8u) & 0xFFu) / 255.0f,
          rz = (float)((seed >> 16u) & 0xFFu) / 255.0f;
    fc.jitter[0] = (rx - 0.5f) * c.cell_size;
    fc.jitter[1] = (ry - 0.5f) * c.cell_size;
    fc.jitter[2] = (rz - 0.5f) * c.cell_size;
    fc.spectral_radius = c.spectral_radius;
    fc.cell_size = c.cell_size;
    fc.hash_size = c.hash_size;
    fc.gather_cap = c.gather_cap;
    fc.frame = frame;
    fc.width = W;
    fc.height = H;
    fc.max_depth = c.max_depth;
    fc.photon_max_depth = c.photon_max_depth;
    fc.rr_after = c.rr_after;
    fc.enable_photons = c.enable_photons;
    fc.n_photons = c.enable_photons ? c.photons_per_pass : 0u;
    fc.quirks = c.quirks;
    fc.guide_light_mix = 1.0f;
    mat4_mul(cam.view_inverse, cam.proj_inverse, fc.M);
    fc.origin[0] = cam.view_inverse[12];
    fc.origin[1] = cam.view_inverse[13];
    fc.origin[2] = cam.view_inverse[14];
    return fc;
}

// ---------------------------------------------------------------------------
// renderer
// ---------------------------------------------------------------------------
namespace {
enum KernelId {
    K_FRAME_PREP, K_PHOTON_EMIT, K_PHOTON_TRACE, K_GRID_BUILD, K_GENERATE, K_EXTEND, K_SHADE, K_SHADOW,
    K_ACCUMULATE, K_GUIDE_SCATTER, K_GUIDE_TRAIN, K_TAIL, K_ORDER, K_COUNT
};
const char* kKernelNames[K_COUNT] = {"frame_prep", "photon_emit", "photon_trace", "grid_build", "generate",
                                     "extend", "shade", "shadow", "accumulate", "guide_scatter",
                                     "guide_train", "tail", "order"};
constexpr uint32_t kUploadHalf = kFrameTable / 2;  // frames per upload (staging half)
struct Pending {
    int kid;
    hipEvent_t a, b;
};
}  // namespace

// ---------------------------------------------------------------------------
// path-guiding network (guide.hip): parameters, Adam state, packed operands
// and the buffers of one training step
// ---------------------------------------------------------------------------
struct wsrt_guide {
    wsrt_device* dev = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool owned_by_renderer = false;
    DevBuf wf, wt, params, m1, m2, step, act, partial, tin0, tin1, tdir, out;
    DevBuf params_avg, wf_avg;  // running average of the parameters and its forward operands (inference)
    uint32_t ns_max = 0;  // training samples per step the buffers hold
    GuideView gv{};
    GuideTrain gt{};
};

namespace {

constexpr uint32_t kGuideChunk = 128;    // samples per partial gradient sum (768 gradient waves per 16384-sample step)
constexpr uint32_t kGuideBatch = 16384;  // samples per Adam step while rendering

// Training-step size for n samples: a multiple of 64, and of the partial-sum
// chunk (which divides it).
void guide_step_shape(uint32_t n, uint32_t* ns, uint32_t* chunk) {
    if (n <= kGuideChunk) {
        *ns = (n + 63u) & ~63u;
        *chunk = *ns;
    } else {
        *ns = (n + kGuideChunk - 1) / kGuideChunk * kGuideChunk;
        *chunk = kGuideChunk;
    }
}

// Initial parameters: uniform Xavier weights from a PCG stream; output
// biases spread the 8 lobes over the cube diagonals with kappa = 2 and equal
// weights.
std::vector<float> guide_initial_params(uint32_t seed) {
    using namespace guide;
    std::vector<float> w(kParams, 0.0f);
    uint32_t st = seed * 747796405u + 2891336453u;
    auto next = [&st]() {
        st = st * 747796405u + 2891336453u;
        uint32_t x = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
        x = (x >> 22u) ^ x;
        return (float)x * 2.3283064365386963e-10f;
    };
    auto fill = [&](int off, int rows, int cols, int used_rows, int used_cols, float scale) {
        const float a = scale * sqrtf(6.0f / (float)(used_rows + used_cols));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c)
                w[off + r * cols + c] = (r < used_rows && c < used_cols) ? a * (2.0f * next() - 1.0f) : 0.0f;
    };
    fill(kW1, kHid, kIn, kHid, kFeat, 1.0f);
    fill(kW2, kHid, kHid, kHid, kHid, 1.0f);
    fill(kW3, kHid, kHid, kHid, kHid, 1.0f);
    fill(kW4, kHid, kHid, kHid, kHid, 1.0f);
    fill(kW5, kOutPad, kHid, kOut, kHid, 0.1f);
    const float d = 0.57735027f;
    for (int k = 0; k < kLobes; ++k) {
        w[kB5 + 5 * k + 0] = (k & 1) ? -d : d;
        w[kB5 + 5 * k + 1] = (k & 2) ? -d : d;
        w[kB5 + 5 * k + 2] = (k & 4) ? -d : d;
        w[kB5 + 5 * k + 3] = 0.6931472f / kKappaScale;  // kappa = 2
        w[kB5 + 5 * k + 4] = 0.0f;
    }
    return w;
}

// The running average restarts at the parameters (and its operands at
// theirs), in stream order after the parameters are packed.
hipError_t guide_reset_average(wsrt_guide* g) {
    hipError_t e = hipMemcpyAsync(g->params_avg.p, g->params.p, 4ull * guide::kParams, hipMemcpyDeviceToDevice,
                                  g->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->wf_avg.p, g->wf.p, 2ull * guide::kFwdPack, hipMemcpyDeviceToDevice, g->stream);
    return e;
}

// The view of the network the tracer samples with: the running average's
// operands and biases in place of the Adam iterate's.
GuideView guide_inference_view(const wsrt_guide* g) {
    GuideView v = g->gv;
    v.wf = g->wf_avg.as<uint16_t>();
    v.params = g->params_avg.as<float>();
    return v;
}

int guide_setup(wsrt_guide* g, uint32_t max_samples, uint32_t seed, float lr, float mix) {
    using namespace guide;
    uint32_t ns, chunk;
    guide_step_shape(max_samples, &ns, &chunk);
    g->ns_max = ns;
    hipStream_t s = g->stream;
    HIP_OR_FAIL(g->wf.alloc(2ull * kFwdPack));
    HIP_OR_FAIL(g->wt.alloc(2ull * kBwdPack));
    HIP_OR_FAIL(g->params.alloc(4ull * kParams));
    HIP_OR_FAIL(g->params_avg.alloc(4ull * kParams));
    HIP_OR_FAIL(g->wf_avg.alloc(2ull * kFwdPack));
    HIP_OR_FAIL(g->m1.alloc(4ull * kParams));
    HIP_OR_FAIL(g->m2.alloc(4ull * kParams));
    HIP_OR_FAIL(g->step.alloc(4));
    HIP_OR_FAIL(g->act.alloc(4ull * guide_act_words(ns)));
    HIP_OR_FAIL(g->partial.alloc(4ull * (ns / chunk) * kParams));
    HIP_OR_FAIL(g->tin0.alloc(16ull * ns));
    HIP_OR_FAIL(g->tin1.alloc(16ull * ns));
    HIP_OR_FAIL(g->tdir.alloc(16ull * ns));
    HIP_OR_FAIL(hipMemsetAsync(g->wf.p, 0, g->wf.bytes, s));
    HIP_OR_FAIL(hipMemsetAsync(g->wt.p, 0, g->wt.bytes, s));
    HIP_OR_FAIL(hipMemsetAsync(g->m1.p, 0, g->m1.bytes, s));
    HIP_OR_FAIL(hipMemsetAsync(g->m2.p, 0, g->m2.bytes, s));
    HIP_OR_FAIL(hipMemsetAsync(g->step.p, 0, 4, s));
    const std::vector<float> w = guide_initial_params(seed);
    HIP_OR_FAIL(hipMemcpyAsync(g->params.p, w.data(), 4ull * kParams, hipMemcpyHostToDevice, s));
    g->gv.wf = g->wf.as<uint16_t>();
    g->gv.wt = g->wt.as<uint16_t>();
    g->gv.params = g->params.as<float>();
    g->gv.mix = mix;
    for (int a = 0; a < 3; ++a) {
        g->gv.bmin[a] = 0.0f;
        g->gv.binv[a] = 1.0f;
    }
    g->gv.rec = guide::Records{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    g->gt.act = g->act.as<uint32_t>();
    g->gt.partial = g->partial.as<float>();
    g->gt.params = g->params.as<float>();
    g->gt.m1 = g->m1.as<float>();
    g->gt.m2 = g->m2.as<float>();
    g->gt.step = g->step.as<uint32_t>();
    g->gt.avg = g->params_avg.as<float>();
    g->gt.wf_avg = g->wf_avg.as<uint16_t>();
    g->gt.chunk = chunk;
    g->gt.lr = lr;
    launch_guide_pack(g->gv, g->params.as<float>(), s);
    HIP_OR_FAIL(guide_reset_average(g));
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipStreamSynchronize(s));
    return WSRT_OK;
}

}  // namespace

struct wsrt_renderer {
    wsrt_device* dev = nullptr;
    const wsrt_scene* scene = nullptr;
    wsrt_config cfg{};
    uint32_t W = 0, H = 0;
    hipStream_t stream = nullptr;
    wsrt_camera_uniform cam{};
    bool have_cam = false;
    uint32_t frame_count = 0;
    // buffers
    DevBuf photons, grid_count, grid_offset, grid_cell, grid_slot, grid_sorted, grid_nvalid, grid_temp, grid_touched;
    DevBuf path_tag;  // ordered queues: per-path bounce tags (whole 256-path tiles)
    DevBuf path_o, path_d, path_tp, path_L, path_wl, path_rng, path_flags, path_hit, q_a, q_b, shade_q,
        shadow;
    DevBuf accum, rgba, counters, gstack, qcounts;
    DevBuf ftable, dframe;           // per-frame constants, device frame counter
    DevBuf ph_id[2], ph_pos[2], ph_dir[2], ph_rng[2], ph_flags[2], ph_stored;
    PhotonState phs{};
    FrameConst* staging = nullptr;   // pinned, 2 halves of kUploadHalf entries
    hipEvent_t upload_ev[2] = {nullptr, nullptr};
    bool upload_pending[2] = {false, false};
    uint32_t upload_half = 0;
    GridBuffers gb{};
    PathLaunch pl{};
    std::map<uint32_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs;  // by batch size
    bool graph_failed = false;
    uint32_t S = 1;                  // frames in flight
    uint32_t last_slot = 0;          // slot of the last frame of the last batch
    // profiling
    bool profiling = false;
    bool count_traversal = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<Pending> pending;
    double k_ms[K_COUNT] = {0};
    uint64_t k_launches[K_COUNT] = {0};
    uint64_t frames = 0, samples = 0, photon_paths = 0;
    // path guiding (cfg.guiding)
    std::unique_ptr<wsrt_guide> guide;
    DevBuf g_in0, g_in1, g_dir, g_tp, g_Lb, g_nee, g_n, g_alb, g_q;
    uint32_t guide_ns = 0, guide_stride = 1, guide_batch = 0;
    uint64_t frames_trained = 0;  // frames whose records trained the network
    // guided next-event estimation: probability of the uniform light sample
    // (1: the plain tracer's light sampling only)
    float guide_light_mix = 1.0f;
};

namespace {

hipEvent_t take_event(wsrt_renderer* r) {
    if (!r->event_pool.empty()) {
        hipEvent_t e = r->event_pool.back();
        r->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct ProfScope {
    wsrt_renderer* r;
    int kid;
    hipEvent_t a = nullptr;
    ProfScope(wsrt_renderer* r_, int k) : r(r_), kid(k) {
        if (r->profiling) {
            a = take_event(r);
            (void)hipEventRecord(a, r->stream);
        }
    }
    ~ProfScope() {
        if (r->profiling && a) {
            hipEvent_t b = take_event(r);
            (void)hipEventRecord(b, r->stream);
            r->pending.push_back({kid, a, b});
        }
    }
};

void drain_profile(wsrt_renderer* r) {
    for (const Pending& p : r->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            r->k_ms[p.kid] += ms;
            r->k_launches[p.kid]++;
        }
        r->event_pool.push_back(p.a);
        r->event_pool.push_back(p.b);
    }
    r->pending.clear();
}

}  // namespace

extern "C" int wsrt_renderer_create(wsrt_device* dev, const wsrt_scene* scene, uint32_t width,
                                    uint32_t height, const wsrt_config* cfg, wsrt_renderer** out) {
    CHECK_ARG(dev && scene && out && width > 0 && height > 0);
    *out = nullptr;
    if ((uint64_t)width * height > (1ull << 26)) return fail(WSRT_E_INVALID, "resolution %ux%u too large", width, height);
    if (scene->dev != dev) return fail(WSRT_E_INVALID, "scene was built on another device");
    wsrt_config c;
    if (cfg) c = *cfg;
    else wsrt_config_default(&c);
    int rc = validate_config(&c);
    if (rc) return rc;
    if (c.env_map && !scene->view.env) return fail(WSRT_E_INVALID, "env_map = 1 but the scene has no environment map");
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    auto r = std::make_unique<wsrt_renderer>();
    r->dev = dev;
    r->scene = scene;
    r->cfg = c;
    r->W = width;
    r->H = height;
    HIP_OR_FAIL(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    r->S = c.frames_in_flight;
    const uint64_t npx = (uint64_t)width * height;  // pixels of one frame
    const uint64_t npaths = npx * r->S;             // path slots of a batch
    if (npaths > (1ull << 30)) return fail(WSRT_E_INVALID, "resolution x frames_in_flight too large");
    const uint32_t np = c.enable_photons ? c.photons_per_pass : 0u;
    const uint64_t nph = (uint64_t)(np ? np : 1u) * r->S;
    const uint64_t cells = (uint64_t)c.hash_size * r->S;
    HIP_OR_FAIL(r->photons.alloc(sizeof(wsrt_photon) * nph));
    HIP_OR_FAIL(r->grid_count.alloc(sizeof(uint32_t) * cells));
    HIP_OR_FAIL(r->grid_offset.alloc(sizeof(uint32_t) * cells));
    HIP_OR_FAIL(r->grid_cell.alloc(sizeof(uint32_t) * nph));
    HIP_OR_FAIL(r->grid_slot.alloc(sizeof(uint32_t) * nph));
    HIP_OR_FAIL(r->grid_sorted.alloc(sizeof(wsrt_photon) * nph));
    HIP_OR_FAIL(r->grid_nvalid.alloc(sizeof(uint32_t)));
    HIP_OR_FAIL(r->grid_touched.alloc(sizeof(uint32_t) * nph));
    HIP_OR_FAIL(r->grid_temp.alloc(sizeof(uint32_t) * grid_temp_words((uint32_t)nph, (uint32_t)cells)));
    HIP_OR_FAIL(r->ph_stored.alloc(4 * ((nph + 31) / 32)));
    // photon paths in queue order, one set per depth parity
    const uint64_t pq = (uint64_t)kSegs * seg_capacity(nph);
    for (int k = 0; k < 2; ++k) {
        HIP_OR_FAIL(r->ph_id[k].alloc(4 * pq));
        HIP_OR_FAIL(r->ph_pos[k].alloc(16 * pq));
        HIP_OR_FAIL(r->ph_dir[k].alloc(16 * pq));
        HIP_OR_FAIL(r->ph_rng[k].alloc(4 * pq));
        HIP_OR_FAIL(r->ph_flags[k].alloc(4 * pq));
        r->phs.q[k] = PhotonQueue{r->ph_id[k].as<uint32_t>(), r->ph_pos[k].as<float4>(), r->ph_dir[k].as<float4>(),
                                  r->ph_rng[k].as<uint32_t>(), r->ph_flags[k].as<uint32_t>()};
    }
    r->phs.stored = r->ph_stored.as<uint32_t>();
    HIP_OR_FAIL(r->path_o.alloc(16 * npaths));
    HIP_OR_FAIL(r->path_d.alloc(16 * npaths));
    HIP_OR_FAIL(r->path_tp.alloc(16 * npaths));
    HIP_OR_FAIL(r->path_L.alloc(16 * npaths));
    HIP_OR_FAIL(r->path_wl.alloc(16 * npaths));
    HIP_OR_FAIL(r->path_rng.alloc(4 * npaths));
    HIP_OR_FAIL(r->path_flags.alloc(4 * npaths));
    if (c.ordered_queues) {
        const uint64_t tag_bytes = (npaths + 255) / 256 * 256;
        HIP_OR_FAIL(r->path_tag.alloc(tag_bytes));
        HIP_OR_FAIL(hipMemsetAsync(r->path_tag.p, 0, tag_bytes, r->stream));
    }
    HIP_OR_FAIL(r->path_hit.alloc(16 * npaths));
    // with guiding, k_shade and k_guide_scatter both fill a bounce's active
    // queue, so each segment may receive two producers' worth of entries
    const uint32_t seg = seg_capacity(npaths) * (c.guiding ? 2u : 1u);
    HIP_OR_FAIL(r->q_a.alloc(4ull * kSegs * seg));
    HIP_OR_FAIL(r->q_b.alloc(4ull * kSegs * seg));
    HIP_OR_FAIL(r->shade_q.alloc(4ull * kSegs * seg * kMatBins));
    HIP_OR_FAIL(r->shadow.alloc(48ull * kSegs * seg));
    const uint32_t depths = std::max(c.max_depth, c.photon_max_depth) + 1;
    // the counters, then the tail word on a line of its own
    HIP_OR_FAIL(r->qcounts.alloc(4 * (queue_count_words(depths) + kSegStride)));
    HIP_OR_FAIL(hipMemsetAsync(r->qcounts.p, 0, 4 * (queue_count_words(depths) + kSegStride), r->stream));
    HIP_OR_FAIL(r->accum.alloc(16 * npx));
    HIP_OR_FAIL(r->rgba.alloc(4 * npx));
    HIP_OR_FAIL(r->counters.alloc(sizeof(Counters)));
    const uint32_t spill = scene->view.stack_size > (uint32_t)kShortStack ? scene->view.stack_size - kShortStack : 0u;
    HIP_OR_FAIL(r->gstack.alloc(sizeof(uint2) * kTraceLanes * (spill ? spill : 1u)));
    HIP_OR_FAIL(r->ftable.alloc(sizeof(FrameConst) * kFrameTable));
    HIP_OR_FAIL(r->dframe.alloc(sizeof(uint32_t)));
    HIP_OR_FAIL(hipHostMalloc((void**)&r->staging, sizeof(FrameConst) * kFrameTable, hipHostMallocDefault));
    HIP_OR_FAIL(hipEventCreateWithFlags(&r->upload_ev[0], hipEventDisableTiming));
    HIP_OR_FAIL(hipEventCreateWithFlags(&r->upload_ev[1], hipEventDisableTiming));
    HIP_OR_FAIL(hipMemsetAsync(r->dframe.p, 0, sizeof(uint32_t), r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->counters.p, 0, sizeof(Counters), r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->accum.p, 0, 16 * npx, r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->photons.p, 0, sizeof(wsrt_photon) * nph, r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->grid_count.p, 0, sizeof(uint32_t) * cells, r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->grid_offset.p, 0, sizeof(uint32_t) * cells, r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->grid_nvalid.p, 0, sizeof(uint32_t), r->stream));
    HIP_OR_FAIL(hipMemsetAsync(r->ph_stored.p, 0, 4 * ((nph + 31) / 32), r->stream));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    r->gb.count = r->grid_count.as<uint32_t>();
    r->gb.offset = r->grid_offset.as<uint32_t>();
    r->gb.cell_id = r->grid_cell.as<uint32_t>();
    r->gb.sorted_slot = r->grid_slot.as<uint32_t>();
    r->gb.sorted = r->grid_sorted.as<wsrt_photon>();
    r->gb.n_valid = r->grid_nvalid.as<uint32_t>();
    r->gb.temp = r->grid_temp.as<uint32_t>();
    r->gb.touched = r->grid_touched.as<uint32_t>();
    PathLaunch& pl = r->pl;
    pl.sv = scene->view;
    if (!c.env_map) pl.sv.env = nullptr;  // the environment lights misses only on request
    pl.ps.o = r->path_o.as<float4>();
    pl.ps.d = r->path_d.as<float4>();
    pl.ps.tp = r->path_tp.as<float4>();
    pl.ps.L = r->path_L.as<float4>();
    pl.ps.wl = r->path_wl.as<float4>();
    pl.ps.rng = r->path_rng.as<uint32_t>();
    pl.ps.flags = r->path_flags.as<uint32_t>();
    pl.ps.tag = c.ordered_queues ? r->path_tag.as<uint8_t>() : nullptr;
    pl.ps.hit = r->path_hit.as<uint4>();
    pl.ps.queue_a = r->q_a.as<uint32_t>();
    pl.ps.queue_b = r->q_b.as<uint32_t>();
    pl.ps.shade_q = r->shade_q.as<uint32_t>();
    pl.ps.shadow = r->shadow.as<float4>();
    pl.ps.capacity = (uint32_t)npaths;
    pl.grid.count = r->gb.count;
    pl.grid.offset = r->gb.offset;
    pl.grid.sorted = r->gb.sorted;
    pl.grid.hash_size = c.hash_size;
    pl.grid.cell_size = c.cell_size;
    pl.ctr = r->counters.as<Counters>();
    pl.gstack = r->gstack.as<uint2>();
    pl.accum = r->accum.as<float4>();
    pl.rgba = r->rgba.as<uint8_t>();
    pl.fa.table = r->ftable.as<FrameConst>();
    pl.fa.frame = r->dframe.as<uint32_t>();
    pl.fa.mask = kFrameTable - 1;
    pl.count_traversal = 0;
    pl.qs.counts = r->qcounts.as<uint32_t>();
    pl.qs.seg_cap = seg;
    // guided paths scatter in their own kernel between shade and shadow, so
    // they stay in the wavefront to the end
    pl.qs.tail = (c.guiding || c.tail_paths == 0) ? nullptr : pl.qs.counts + queue_count_words(depths);
    pl.qs.tail_paths = c.tail_paths;
    pl.qs.trace_rounds = c.trace_rounds ? c.trace_rounds : 1u;
    pl.guided = c.guiding;
    pl.extend_refill = c.extend_refill;
    pl.guide_mis = c.guiding ? 1u : 0u;  // the shadow kernel flags visible guided light samples
    if (c.guiding) {
        const uint32_t want = (uint32_t)std::min<uint64_t>(c.guide_train_samples, npx * guide::kRecBounces);
        // whole chunks of samples per step (at least one step of 64)
        const uint32_t ns = want >= kGuideBatch   ? want / kGuideBatch * kGuideBatch
                            : want >= kGuideChunk ? want / kGuideChunk * kGuideChunk
                                                  : (want + 63u) & ~63u;
        // the records a frame's training reads, per frame slot
        const uint64_t nrec = (uint64_t)ns * r->S;
        HIP_OR_FAIL(r->g_in0.alloc(16 * nrec));
        HIP_OR_FAIL(r->g_in1.alloc(16 * nrec));
        HIP_OR_FAIL(r->g_dir.alloc(16 * nrec));
        HIP_OR_FAIL(r->g_tp.alloc(16 * nrec));
        HIP_OR_FAIL(r->g_Lb.alloc(16 * nrec));
        HIP_OR_FAIL(r->g_nee.alloc(16 * nrec));
        // the light proposal draws from the guide when every light is an
        // area light and there are few (one triangle test per light)
        r->guide_light_mix = (scene->n_point_lights == 0 && scene->view.n_lights <= kGuideLightsMax) ? 0.5f : 1.0f;
        HIP_OR_FAIL(r->g_n.alloc(16 * npaths));
        HIP_OR_FAIL(r->g_alb.alloc(16 * npaths));
        HIP_OR_FAIL(r->g_q.alloc(4ull * kSegs * seg));
        r->guide = std::make_unique<wsrt_guide>();
        wsrt_guide* g = r->guide.get();
        g->dev = dev;
        g->stream = r->stream;
        g->owned_by_renderer = true;
        int grc = guide_setup(g, ns, 0x5eedu, c.guide_lr, c.guide_mix);
        if (grc) return grc;
        r->guide_ns = ns;
        r->guide_stride = (uint32_t)std::max<uint64_t>(1, npx * guide::kRecBounces / ns);
        r->guide_batch = std::min(ns, kGuideBatch);  // divides ns
        const wsrt_scene_info& si = scene->info;
        for (int a = 0; a < 3; ++a) {
            const float lo = si.bounds[a], ext = si.bounds[3 + a] - si.bounds[a];
            g->gv.bmin[a] = lo;
            g->gv.binv[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
        }
        g->gv.rec = guide::Records{r->g_in0.as<float4>(), r->g_in1.as<float4>(), r->g_dir.as<float4>(),
                                   r->g_tp.as<float4>(), r->g_Lb.as<float4>(), r->g_nee.as<float4>(),
                                   (uint32_t)npx, ns,
                                   r->guide_stride, guide::fast_div((uint32_t)npx),
                                   guide::fast_div(r->guide_stride)};
        pl.gv = guide_inference_view(g);
    }
    pl.ps.guide_q = r->g_q.as<uint32_t>();
    pl.ps.gn = r->g_n.as<float4>();
    pl.ps.galb = r->g_alb.as<float4>();
    pl.qs.pseg_cap = seg_capacity(nph);
    // default camera of CameraController::new (camera.rs:144-162)
    const float pos[3] = {0.0f, 0.0f, 3.0f};
    wsrt_camera_build_uniform(pos, -90.0f * (3.14159265359f / 180.0f), 0.0f, (float)width / (float)height,
                              0, &r->cam);
    dev->live++;
    const_cast<wsrt_scene*>(scene)->renderers++;
    *out = r.release();
    return WSRT_OK;
}

extern "C" int wsrt_renderer_destroy(wsrt_renderer* r) {
    if (!r) return WSRT_OK;
    r->dev->live--;
    const_cast<wsrt_scene*>(r->scene)->renderers--;
    (void)hipSetDevice(r->dev->ordinal);
    (void)hipStreamSynchronize(r->stream);
    for (const Pending& p : r->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : r->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : r->upload_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& kv : r->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    if (r->staging) (void)hipHostFree(r->staging);
    hipStream_t s = r->stream;
    delete r;
    (void)hipStreamDestroy(s);
    return WSRT_OK;
}

extern "C" int wsrt_renderer_set_camera(wsrt_renderer* r, const wsrt_camera_uniform* cam) {
    CHECK_ARG(r && cam);
    r->cam = *cam;
    r->have_cam = true;
    return WSRT_OK;
}

// Enqueues a batch of nf consecutive frames (mod.rs:168-207 each) traced in
// one wavefront; kernels read each frame's constants through the device
// frame counter, which the batch advances by nf at its end. Used directly
// when profiling or counting, and captured into one graph per batch size.
// Whether the next batch trains the guide network (guide_train_frames).
static bool batch_trains(const wsrt_renderer* r) {
    return r->pl.guided && (r->cfg.guide_train_frames == 0 || r->frames_trained < r->cfg.guide_train_frames);
}

static int enqueue_batch(wsrt_renderer* r, uint32_t nf, bool train) {
    hipStream_t s = r->stream;
    const PathLaunch& pl = r->pl;
    const uint32_t npx = r->W * r->H;
    const uint32_t np = r->cfg.enable_photons ? r->cfg.photons_per_pass : 0u;
    {
        ProfScope p(r, K_FRAME_PREP);
        launch_clear_queues(pl.qs, std::max(r->cfg.max_depth, r->cfg.photon_max_depth) + 1, s);
    }
    if (np) {
        {
            ProfScope p(r, K_PHOTON_EMIT);
            launch_photon_emit(pl.sv, pl.fa, np, nf, pl.qs, r->phs, pl.count_traversal, pl.ctr, s);
        }
        for (uint32_t d = 0; d < r->cfg.photon_max_depth; ++d) {
            ProfScope p(r, K_PHOTON_TRACE);
            launch_photon_bounce(pl.sv, d, pl.count_traversal, r->phs, pl.qs, r->photons.as<wsrt_photon>(), pl.ctr,
                                 pl.gstack, s);
        }
        ProfScope p(r, K_GRID_BUILD);
        if (grid_build_stored(r->photons.as<wsrt_photon>(), r->phs.stored, np, nf, r->cfg.hash_size, pl.fa, r->gb,
                              s))
            return fail(WSRT_E_HIP, "grid build launch failed");
    }
    {
        ProfScope p(r, K_GENERATE);
        launch_generate(pl, npx, nf, s);
    }
    for (uint32_t b = 0; b < r->cfg.max_depth; ++b) {
        uint32_t* q = (b & 1u) ? pl.ps.queue_b : pl.ps.queue_a;
        uint32_t* nq = (b & 1u) ? pl.ps.queue_a : pl.ps.queue_b;
        {
            ProfScope p(r, K_EXTEND);
            launch_extend(pl, b, npx * nf, q, s);
        }
        {
            ProfScope p(r, K_SHADE);
            launch_shade(pl, npx, b, nq, s);
        }
        if (pl.guided) {
            // scattering and next-event estimation of the Lambert vertices,
            // whose shadow rays the shadow kernel then traces
            ProfScope p(r, K_GUIDE_SCATTER);
            launch_guide_scatter(pl.sv, pl.ps, pl.qs, pl.gv, pl.fa, b, nq, pl.ctr, s);
        }
        {
            ProfScope p(r, K_SHADOW);
            launch_shadow(pl, b, s);
        }
        if (pl.ps.tag && b + 1 < r->cfg.max_depth) {
            ProfScope p(r, K_ORDER);
            launch_order_active(pl, b + 1, npx * nf, nq, s);
        }
    }
    if (pl.qs.tail && r->cfg.max_depth > kTailFirstBounce) {
        ProfScope p(r, K_TAIL);
        launch_tail(pl, npx, r->cfg.max_depth, s);
    }
    {
        ProfScope p(r, K_ACCUMULATE);
        launch_accumulate(pl, npx, nf, s);
    }
    if (pl.guided && train) {
        // each frame's records train the network in Adam steps over
        // minibatches of guide_batch samples
        ProfScope p(r, K_GUIDE_TRAIN);
        wsrt_guide* g = r->guide.get();
        const uint32_t mb = r->guide_batch;
        GuideTrain gt = g->gt;
        gt.chunk = std::min(mb, kGuideChunk);
        for (uint32_t slot = 0; slot < nf; ++slot) {
            launch_guide_prepare(pl.ps, pl.gv, pl.fa, npx, slot, r->guide_ns, r->guide_stride,
                                 g->tin0.as<float4>(), g->tin1.as<float4>(), g->tdir.as<float4>(), s);
            for (uint32_t off = 0; off < r->guide_ns; off += mb)
                launch_guide_train_step(g->gv, gt, g->tin0.as<float4>() + off, g->tin1.as<float4>() + off,
                                        g->tdir.as<float4>() + off, mb, 1.0f / (float)mb, true, s);
        }
    }
    launch_advance_frame(r->dframe.as<uint32_t>(), nf, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(WSRT_E_HIP, "frame launch: %s", hipGetErrorString(e));
    return WSRT_OK;
}

static hipGraphExec_t batch_graph(wsrt_renderer* r, uint32_t nf, bool train) {
    const uint32_t key = nf | (train ? 0x10000u : 0u);
    auto it = r->graphs.find(key);
    if (it != r->graphs.end()) return it->second.second;
    if (r->graph_failed) return nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    bool prof = r->profiling;
    r->profiling = false;
    int rc = ok ? enqueue_batch(r, nf, train) : WSRT_E_HIP;
    r->profiling = prof;
    if (ok) ok = hipStreamEndCapture(r->stream, &g) == hipSuccess && rc == WSRT_OK;
    if (ok) ok = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {
        if (g) (void)hipGraphDestroy(g);
        r->graph_failed = true;
        (void)hipGetLastError();
        return nullptr;
    }
    r->graphs[key] = {g, ge};
    return ge;
}

// Frames [first, first + n) with camera `cam`: the host computes each
// frame's constants, uploads them in stream order into the device table
// (so no frame in flight has its entry overwritten), sets the device frame
// counter and submits batches of up to frames_in_flight frames, each one
// graph (or eager launches when profiling or counting).
static int submit_frames(wsrt_renderer* r, uint32_t first, uint32_t n, const wsrt_camera_uniform& cam) {
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    hipStream_t s = r->stream;
    const bool eager = r->profiling || r->count_traversal || !r->cfg.frame_graphs;
    r->pl.count_traversal = r->count_traversal ? 1u : 0u;
    for (uint32_t done = 0; done < n;) {
        const uint32_t m = std::min(n - done, kUploadHalf);
        const uint32_t half = r->upload_half;
        r->upload_half ^= 1u;
        if (r->upload_pending[half]) HIP_OR_FAIL(hipEventSynchronize(r->upload_ev[half]));
        FrameConst* st = r->staging + (size_t)half * kUploadHalf;
        const uint32_t f0 = first + done;
        for (uint32_t i = 0; i < m; ++i) {
            wsrt_camera_uniform c = cam;
            c.frame_count = f0 + i;
            st[i] = make_frame_const(r->cfg, c, r->W, r->H);
            st[i].guide_light_mix = r->guide_light_mix;
        }
        // device slots f & mask: at most two contiguous ranges
        uint32_t slot = f0 & (kFrameTable - 1);
        uint32_t m1 = std::min(m, kFrameTable - slot);
        FrameConst* dt = r->ftable.as<FrameConst>();
        HIP_OR_FAIL(hipMemcpyAsync(dt + slot, st, sizeof(FrameConst) * m1, hipMemcpyHostToDevice, s));
        if (m1 < m)
            HIP_OR_FAIL(hipMemcpyAsync(dt, st + m1, sizeof(FrameConst) * (m - m1), hipMemcpyHostToDevice, s));
        HIP_OR_FAIL(hipEventRecord(r->upload_ev[half], s));
        r->upload_pending[half] = true;
        launch_set_frame(r->dframe.as<uint32_t>(), f0, s);
        for (uint32_t i = 0; i < m;) {
            const uint32_t nf = std::min(m - i, r->S);
            const bool train = batch_trains(r);
            hipGraphExec_t ge = eager ? nullptr : batch_graph(r, nf, train);
            if (ge) {
                HIP_OR_FAIL(hipGraphLaunch(ge, s));
            } else {
                int rc = enqueue_batch(r, nf, train);
                if (rc) return rc;
            }
            if (train) r->frames_trained += nf;
            r->last_slot = nf - 1;
            i += nf;
        }
        done += m;
    }
    const uint32_t npx = r->W * r->H;
    r->frames += n;
    r->samples += (uint64_t)npx * n;
    r->photon_paths += (uint64_t)(r->cfg.enable_photons ? r->cfg.photons_per_pass : 0u) * n;
    return WSRT_OK;
}

extern "C" int wsrt_renderer_render(wsrt_renderer* r) {
    CHECK_ARG(r);
    int rc = submit_frames(r, r->cam.frame_count, 1, r->cam);
    if (rc) return rc;
    r->frame_count++;  // mod.rs:204
    return WSRT_OK;
}

extern "C" int wsrt_renderer_render_frames(wsrt_renderer* r, uint32_t n) {
    CHECK_ARG(r);
    // frame k is shaded with frame_count = the renderer's counter (engine.rs:158-166)
    int rc = submit_frames(r, r->frame_count, n, r->cam);
    if (rc) return rc;
    r->frame_count += n;
    return WSRT_OK;
}

extern "C" int wsrt_renderer_get_frame_count(const wsrt_renderer* r, uint32_t* out) {
    CHECK_ARG(r && out);
    *out = r->frame_count;
    return WSRT_OK;
}
extern "C" int wsrt_renderer_set_frame_count(wsrt_renderer* r, uint32_t fc) {
    CHECK_ARG(r);
    r->frame_count = fc;
    return WSRT_OK;
}
extern "C" int wsrt_renderer_resize(wsrt_renderer* r, uint32_t ww, uint32_t wh) {
    CHECK_ARG(r && ww > 0 && wh > 0);
    return WSRT_OK;  // the internal resolution is fixed; frame_count is kept (mod.rs:149-166)
}
extern "C" int wsrt_renderer_sync(wsrt_renderer* r) {
    CHECK_ARG(r);
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    return WSRT_OK;
}

static int read_back(wsrt_renderer* r, void* dst, const void* src, size_t bytes) {
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, r->stream));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    return WSRT_OK;
}

extern "C" int wsrt_renderer_read_accum(wsrt_renderer* r, float* out) {
    CHECK_ARG(r && out);
    return read_back(r, out, r->accum.p, 16ull * r->W * r->H);
}
extern "C" int wsrt_renderer_read_rgba8(wsrt_renderer* r, uint8_t* out) {
    CHECK_ARG(r && out);
    return read_back(r, out, r->rgba.p, 4ull * r->W * r->H);
}
// screenshot.rs:47-73: rows reversed (row 0 = image top), each channel
// c / 255 through the sRGB OETF, rounded back to 8 bits, alpha 255.
extern "C" int wsrt_renderer_read_screenshot(wsrt_renderer* r, uint8_t* out) {
    CHECK_ARG(r && out);
    const size_t row = 4ull * r->W;
    std::vector<uint8_t> lin(row * r->H);
    int rc = read_back(r, lin.data(), r->rgba.p, lin.size());
    if (rc) return rc;
    uint8_t lut[256];
    for (int c = 0; c < 256; ++c) {
        const float cf = (float)c / 255.0f;
        const float v = cf <= 0.0031308f ? cf * 12.92f : 1.055f * powf(cf, 1.0f / 2.4f) - 0.055f;
        lut[c] = (uint8_t)roundf(std::min(std::max(v, 0.0f), 1.0f) * 255.0f);
    }
    for (uint32_t y = 0; y < r->H; ++y) {
        const uint8_t* src = lin.data() + (size_t)(r->H - 1 - y) * row;
        uint8_t* dst = out + (size_t)y * row;
        for (uint32_t x = 0; x < r->W; ++x) {
            dst[4 * x + 0] = lut[src[4 * x + 0]];
            dst[4 * x + 1] = lut[src[4 * x + 1]];
            dst[4 * x + 2] = lut[src[4 * x + 2]];
            dst[4 * x + 3] = 255;
        }
    }
    return WSRT_OK;
}

// The presented frame. The blit (blit_pass.rs:114-180) keeps the image's
// aspect: scale_x = target / window when the window is wider, else scale_y =
// window / target, both in f32. The surface is wgpu's default sRGB format,
// so every value written is encoded with the sRGB OETF; the encodings and
// the overlay's per-point blend step are 256-entry tables computed here in
// double precision.
extern "C" int wsrt_renderer_present(wsrt_renderer* r, uint32_t win_w, uint32_t win_h, int overlay_photons,
                                     uint8_t* out) {
    CHECK_ARG(r && out);
    if (win_w == 0 || win_h == 0 || win_w > 16384 || win_h > 16384)
        return fail(WSRT_E_INVALID, "window %ux%u outside 1..16384", win_w, win_h);
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    auto oetf8 = [](double v) {
        v = std::min(std::max(v, 0.0), 1.0);
        const double e = v <= 0.0031308 ? v * 12.92 : 1.055 * std::pow(v, 1.0 / 2.4) - 0.055;
        return (uint8_t)std::lround(std::min(std::max(e, 0.0), 1.0) * 255.0);
    };
    auto eotf = [](int c) {
        const double e = c / 255.0;
        return e <= 0.04045 ? e / 12.92 : std::pow((e + 0.055) / 1.055, 2.4);
    };
    uint8_t lut[1024];
    const double add[3] = {0.3 * 1.0, 0.3 * 0.6, 0.3 * 0.1};  // debug_photons.wgsl colour times alpha
    for (int c = 0; c < 256; ++c) {
        lut[c] = oetf8(c / 255.0);
        for (int k = 0; k < 3; ++k) lut[256 * (k + 1) + c] = oetf8(eotf(c) + add[k]);
    }
    PresentArgs a{};
    memcpy(a.view, r->cam.view, sizeof a.view);
    memcpy(a.proj, r->cam.proj, sizeof a.proj);
    const float target_aspect = (float)r->W / (float)r->H, window_aspect = (float)win_w / (float)win_h;
    a.sx = 1.0f;
    a.sy = 1.0f;
    if (window_aspect > target_aspect) a.sx = target_aspect / window_aspect;
    else a.sy = window_aspect / target_aspect;
    a.win_w = win_w;
    a.win_h = win_h;
    a.img_w = r->W;
    a.img_h = r->H;
    a.point_alpha = (uint8_t)std::lround((double)(0.3f * 255.0f));  // the shader's f32 alpha 0.3 as unorm8: 77
    const size_t npx = (size_t)win_w * win_h;
    const uint32_t np = r->cfg.enable_photons ? r->cfg.photons_per_pass : 0u;
    DevBuf dlut, dout, dcount;
    HIP_OR_FAIL(dlut.alloc(sizeof lut));
    HIP_OR_FAIL(dout.alloc(4 * npx));
    HIP_OR_FAIL(hipMemcpyAsync(dlut.p, lut, sizeof lut, hipMemcpyHostToDevice, r->stream));
    a.lut = dlut.as<uint8_t>();
    uint32_t* count = nullptr;
    if (overlay_photons) {
        HIP_OR_FAIL(dcount.alloc(4 * npx));
        HIP_OR_FAIL(hipMemsetAsync(dcount.p, 0, 4 * npx, r->stream));
        count = dcount.as<uint32_t>();
    }
    const size_t first = (size_t)r->last_slot * np;
    launch_present(r->rgba.as<uint8_t>(), r->photons.as<wsrt_photon>() + first, r->ph_stored.as<uint32_t>(),
                   (uint32_t)first, np, a, count, dout.as<uint8_t>(), r->stream);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipMemcpyAsync(out, dout.p, 4 * npx, hipMemcpyDeviceToHost, r->stream));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    return WSRT_OK;
}

extern "C" int wsrt_renderer_write_accum(wsrt_renderer* r, const float* in) {
    CHECK_ARG(r && in);
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(r->accum.p, in, 16ull * r->W * r->H, hipMemcpyHostToDevice, r->stream));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    return WSRT_OK;
}

extern "C" int wsrt_renderer_get_stats(wsrt_renderer* r, wsrt_frame_stats* out) {
    CHECK_ARG(r && out);
    std::vector<Counters> cv(1);
    Counters& c = cv[0];
    int rc = read_back(r, &c, r->counters.p, sizeof c);
    if (rc) return rc;
    out->frames = r->frames;
    out->camera_paths = r->samples;
    out->extend_rays = c.extend_rays;
    out->shadow_rays = c.shadow_rays;
    out->photon_paths = r->photon_paths;
    out->photon_rays = 0;
    out->photons_stored = 0;
    for (int b = 0; b < kStatBlocks; ++b) {
        out->photon_rays += c.photon_rays_part[b];
        out->photons_stored += c.photons_stored_part[b];
    }
    out->guide_samples = c.guide_samples;
    // by ray type: photon (bounce + emit root tests), camera extend (wavefront
    // + tail), shadow (wavefront + tail)
    const int by_type[3][2] = {{kRkPhotonBounce, kRkPhotonEmit}, {kRkExtend, kRkTailExtend},
                               {kRkShadow, kRkTailShadow}};
    for (int k = 0; k < 3; ++k) {
        out->trav_nodes[k] = c.trav_nodes[by_type[k][0]] + c.trav_nodes[by_type[k][1]];
        out->trav_tris[k] = c.trav_tris[by_type[k][0]] + c.trav_tris[by_type[k][1]];
        out->trav_rays[k] = c.trav_rays[by_type[k][0]] + c.trav_rays[by_type[k][1]];
    }
    for (int k = 0; k < 6; ++k) {
        out->kernel_rays[k] = c.kernel_rays[k];
        out->kernel_nodes[k] = c.trav_nodes[k];
        out->kernel_tris[k] = c.trav_tris[k];
        out->kernel_counted_rays[k] = c.trav_rays[k];
    }
    out->extend_wave_nodes = c.extend_wave_nodes;
    for (int k = 0; k < 6; ++k) out->ray_max_nodes[k] = c.ray_max_nodes[k];
    return WSRT_OK;
}

extern "C" int wsrt_renderer_set_traversal_counting(wsrt_renderer* r, int enable) {
    CHECK_ARG(r);
    r->count_traversal = enable != 0;
    return WSRT_OK;
}
extern "C" int wsrt_renderer_reset_stats(wsrt_renderer* r) {
    CHECK_ARG(r);
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    HIP_OR_FAIL(hipMemsetAsync(r->counters.p, 0, sizeof(Counters), r->stream));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    r->frames = r->samples = r->photon_paths = 0;
    return WSRT_OK;
}

extern "C" int wsrt_renderer_set_profiling(wsrt_renderer* r, int enable) {
    CHECK_ARG(r);
    r->profiling = enable != 0;
    return WSRT_OK;
}
extern "C" int wsrt_renderer_get_profile(wsrt_renderer* r, wsrt_kernel_stat* out, uint32_t cap,
                                         uint32_t* n) {
    CHECK_ARG(r && n);
    HIP_OR_FAIL(hipSetDevice(r->dev->ordinal));
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    drain_profile(r);
    *n = K_COUNT;
    if (!out) return WSRT_OK;
    for (uint32_t k = 0; k < K_COUNT && k < cap; ++k) {
        memset(&out[k], 0, sizeof out[k]);
        snprintf(out[k].name, sizeof out[k].name, "%s", kKernelNames[k]);
        out[k].launches = r->k_launches[k];
        out[k].total_ms = r->k_ms[k];
    }
    return WSRT_OK;
}
extern "C" int wsrt_renderer_reset_profile(wsrt_renderer* r) {
    CHECK_ARG(r);
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    drain_profile(r);
    for (int k = 0; k < K_COUNT; ++k) {
        r->k_ms[k] = 0;
        r->k_launches[k] = 0;
    }
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// debug / parity taps
// ---------------------------------------------------------------------------
// The debug taps read the last frame of the last batch (frame slot
// last_slot of the batch's buffers).
// The stored bits of photon slots first .. first+np-1, one flag per slot.
static int read_stored_flags(wsrt_renderer* r, size_t first, uint32_t np, std::vector<uint8_t>& flag) {
    flag.assign(np, 0);
    if (np == 0) return WSRT_OK;
    const size_t w0 = first / 32, w1 = (first + np + 31) / 32;
    std::vector<uint32_t> words(w1 - w0);
    int rc = read_back(r, words.data(), r->ph_stored.as<uint32_t>() + w0, 4 * (w1 - w0));
    if (rc) return rc;
    for (uint32_t i = 0; i < np; ++i) {
        const size_t b = first + i;
        flag[i] = (words[b / 32 - w0] >> (b % 32)) & 1u;
    }
    return WSRT_OK;
}

extern "C" int wsrt_debug_read_photons(wsrt_renderer* r, wsrt_photon* out, uint32_t cap,
                                       uint32_t* n_stored) {
    CHECK_ARG(r && out);
    uint32_t np = r->cfg.enable_photons ? r->cfg.photons_per_pass : 0u;
    if (cap < np) return fail(WSRT_E_INVALID, "photon buffer holds %u < %u", cap, np);
    const size_t first = (size_t)r->last_slot * np;
    int rc = read_back(r, out, r->photons.as<wsrt_photon>() + first, sizeof(wsrt_photon) * np);
    std::vector<uint8_t> flag;
    if (!rc) rc = read_stored_flags(r, first, np, flag);
    if (rc) return rc;
    // a slot holds a photon only when its stored bit is set; the others read
    // as the reference's skipped photons (all zero)
    uint32_t k = 0;
    for (uint32_t i = 0; i < np; ++i) {
        if (flag[i]) k++;
        else memset(&out[i], 0, sizeof(wsrt_photon));
    }
    if (n_stored) *n_stored = k;
    return WSRT_OK;
}

// The grid of the frame in the dense form of wsrt_photon_grid_build. The
// renderer writes cell ids for stored photons only and offsets for
// non-empty cells only (nothing else is read); here a slot without a photon
// reads kInvalid, and an empty cell's offset is the number of photons in the
// cells before it (what a dense exclusive scan gives). Non-empty cells report
// the device's offsets.
extern "C" int wsrt_debug_read_grid(wsrt_renderer* r, uint32_t* cell_id, uint32_t* count,
                                    uint32_t* offset, uint32_t* sorted_slot) {
    CHECK_ARG(r);
    const uint32_t np = r->cfg.enable_photons ? r->cfg.photons_per_pass : 0u;
    const uint32_t hs = r->cfg.hash_size, ls = r->last_slot;
    int rc = WSRT_OK;
    if (cell_id && np) {
        std::vector<uint8_t> flag;
        rc = read_back(r, cell_id, r->grid_cell.as<uint32_t>() + (size_t)ls * np, 4ull * np);
        if (!rc) rc = read_stored_flags(r, (size_t)ls * np, np, flag);
        if (rc) return rc;
        for (uint32_t i = 0; i < np; ++i)
            if (!flag[i]) cell_id[i] = kInvalid;
    }
    std::vector<uint32_t> cnt(hs), off(hs);
    if (!rc) rc = read_back(r, cnt.data(), r->grid_count.as<uint32_t>() + (size_t)ls * hs, 4ull * hs);
    if (!rc) rc = read_back(r, off.data(), r->grid_offset.as<uint32_t>() + (size_t)ls * hs, 4ull * hs);
    if (rc) return rc;
    // this frame's photons start at the offset of its first non-empty cell
    uint32_t base = 0;
    for (uint32_t c = 0; c < hs; ++c)
        if (cnt[c]) {
            base = off[c];
            break;
        }
    uint64_t valid = 0;
    for (uint32_t c = 0; c < hs; ++c) {
        off[c] = cnt[c] ? off[c] - base : (uint32_t)valid;
        valid += cnt[c];
    }
    if (count) memcpy(count, cnt.data(), 4ull * hs);
    if (offset) memcpy(offset, off.data(), 4ull * hs);
    if (sorted_slot && np) {
        std::vector<uint32_t> sl(valid);
        if (valid) rc = read_back(r, sl.data(), r->grid_slot.as<uint32_t>() + base, 4ull * valid);
        if (rc) return rc;
        for (uint64_t k = 0; k < np; ++k)
            sorted_slot[k] = k < valid ? sl[k] - ls * np : kInvalid;  // photon slot within the frame
    }
    return WSRT_OK;
}

extern "C" int wsrt_debug_read_paths(wsrt_renderer* r, float* wl, float* L) {
    CHECK_ARG(r);
    const size_t npx = (size_t)r->W * r->H, off = npx * r->last_slot;
    int rc = WSRT_OK;
    if (wl) rc = read_back(r, wl, r->path_wl.as<float4>() + off, 16 * npx);
    if (L && !rc) rc = read_back(r, L, r->path_L.as<float4>() + off, 16 * npx);
    return rc;
}

extern "C" int wsrt_debug_trace_paths(wsrt_renderer* r, const uint32_t* pixel_ids, uint32_t n, float* wl,
                                      float* L) {
    CHECK_ARG(r && (pixel_ids || n == 0) && ((wl && L) || n == 0));
    const size_t npx = (size_t)r->W * r->H;
    for (uint32_t k = 0; k < n; ++k)
        if (pixel_ids[k] >= npx) return fail(WSRT_E_INVALID, "pixel id %u >= %zu pixels", pixel_ids[k], npx);
    if (n == 0) return WSRT_OK;
    std::vector<float> all_wl(4 * npx), all_L(4 * npx);
    int rc = wsrt_debug_read_paths(r, all_wl.data(), all_L.data());
    if (rc) return rc;
    for (uint32_t k = 0; k < n; ++k) {
        memcpy(wl + 4ull * k, all_wl.data() + 4ull * pixel_ids[k], 16);
        memcpy(L + 4ull * k, all_L.data() + 4ull * pixel_ids[k], 16);
    }
    return WSRT_OK;
}

extern "C" int wsrt_photon_grid_build(wsrt_device* dev, const wsrt_photon* photons, uint32_t n,
                                      uint32_t frame_count, float cell_size, uint32_t hash_size,
                                      uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                                      uint32_t* sorted_slot) {
    CHECK_ARG(dev && (photons || n == 0) && count && offset);
    CHECK_ARG((cell_id && sorted_slot) || n == 0);
    if (hash_size == 0 || (hash_size & (hash_size - 1)) || hash_size > (1u << 26))
        return fail(WSRT_E_INVALID, "hash_size %u must be a power of two <= 2^26", hash_size);
    if (!(cell_size > 0.0f)) return fail(WSRT_E_INVALID, "cell_size must be > 0");
    if (n > (1u << 26)) return fail(WSRT_E_INVALID, "%u photons > 2^26", n);
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    hipStream_t s = dev->stream;
    wsrt_config c;
    wsrt_config_default(&c);
    c.cell_size = cell_size;
    c.hash_size = hash_size;
    wsrt_camera_uniform cam;
    memset(&cam, 0, sizeof cam);
    cam.frame_count = frame_count;
    FrameConst fc = make_frame_const(c, cam, 1, 1);
    const uint32_t n1 = n ? n : 1u;
    DevBuf ph, cnt, off, cell, slot, sorted, nvalid, temp, tab, fr;
    HIP_OR_FAIL(tab.alloc(sizeof fc));
    HIP_OR_FAIL(fr.alloc(sizeof(uint32_t)));
    HIP_OR_FAIL(hipMemcpyAsync(tab.p, &fc, sizeof fc, hipMemcpyHostToDevice, s));
    HIP_OR_FAIL(hipMemsetAsync(fr.p, 0, sizeof(uint32_t), s));
    FrameArgs fa{tab.as<FrameConst>(), fr.as<uint32_t>(), 0u};
    HIP_OR_FAIL(ph.alloc(sizeof(wsrt_photon) * n1));
    HIP_OR_FAIL(cnt.alloc(4ull * hash_size));
    HIP_OR_FAIL(off.alloc(4ull * hash_size));
    HIP_OR_FAIL(cell.alloc(4ull * n1));
    HIP_OR_FAIL(slot.alloc(4ull * n1));
    HIP_OR_FAIL(sorted.alloc(sizeof(wsrt_photon) * n1));
    HIP_OR_FAIL(nvalid.alloc(4));
    HIP_OR_FAIL(temp.alloc(4 * grid_temp_words(n1, hash_size)));
    if (n) HIP_OR_FAIL(hipMemcpyAsync(ph.p, photons, sizeof(wsrt_photon) * n, hipMemcpyHostToDevice, s));
    GridBuffers g;
    g.count = cnt.as<uint32_t>();
    g.offset = off.as<uint32_t>();
    g.cell_id = cell.as<uint32_t>();
    g.sorted_slot = slot.as<uint32_t>();
    g.sorted = sorted.as<wsrt_photon>();
    g.n_valid = nvalid.as<uint32_t>();
    g.temp = temp.as<uint32_t>();
    HIP_OR_FAIL(hipMemsetAsync(g.n_valid, 0, 4, s));
    if (grid_build(ph.as<wsrt_photon>(), n, 1, hash_size, fa, g, s)) return fail(WSRT_E_HIP, "grid build failed");
    uint32_t valid = 0;
    HIP_OR_FAIL(hipMemcpyAsync(count, g.count, 4ull * hash_size, hipMemcpyDeviceToHost, s));
    HIP_OR_FAIL(hipMemcpyAsync(offset, g.offset, 4ull * hash_size, hipMemcpyDeviceToHost, s));
    if (n) {
        HIP_OR_FAIL(hipMemcpyAsync(cell_id, g.cell_id, 4ull * n, hipMemcpyDeviceToHost, s));
        HIP_OR_FAIL(hipMemcpyAsync(sorted_slot, g.sorted_slot, 4ull * n, hipMemcpyDeviceToHost, s));
        HIP_OR_FAIL(hipMemcpyAsync(&valid, g.n_valid, 4, hipMemcpyDeviceToHost, s));
    }
    HIP_OR_FAIL(hipStreamSynchronize(s));
    // slots past the valid ones are not part of the grid
    for (uint32_t k = valid; k < n; ++k) sorted_slot[k] = kInvalid;
    return WSRT_OK;
}

// Scratch layout: the frame's constants, the device frame counter, the
// valid-photon count, then the grid's work words.
extern "C" int wsrt_photon_grid_build_noalloc(wsrt_device* dev, const wsrt_photon* photons, uint32_t n,
                                              uint32_t frame_count, float cell_size, uint32_t hash_size,
                                              uint32_t* cell_id, uint32_t* count, uint32_t* offset,
                                              uint32_t* sorted_slot, void* stream) {
    CHECK_ARG(dev && dev->scratch && (photons || n == 0) && count && offset);
    CHECK_ARG((cell_id && sorted_slot) || n == 0);
    if (hash_size == 0 || (hash_size & (hash_size - 1)) || hash_size > kNoAllocMaxCells)
        return fail(WSRT_E_INVALID, "hash_size %u must be a power of two <= 2^22 (scratch limit)", hash_size);
    if (n > kNoAllocMaxPhotons) return fail(WSRT_E_INVALID, "%u photons > 2^20 (scratch limit)", n);
    if (!(cell_size > 0.0f)) return fail(WSRT_E_INVALID, "cell_size must be > 0");
    const size_t words = grid_temp_words(n ? n : 1u, hash_size);
    if (768 + 4 * words > kScratchBytes) return fail(WSRT_E_INVALID, "grid work exceeds the scratch array");
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : dev->stream;
    wsrt_config c;
    wsrt_config_default(&c);
    c.cell_size = cell_size;
    c.hash_size = hash_size;
    wsrt_camera_uniform cam;
    memset(&cam, 0, sizeof cam);
    cam.frame_count = frame_count;
    const FrameConst fc = make_frame_const(c, cam, 1, 1);
    static_assert(sizeof(FrameConst) <= 256, "frame constants fit their scratch slot");
    FrameConst* tab = reinterpret_cast<FrameConst*>(dev->scratch);
    uint32_t* frame = reinterpret_cast<uint32_t*>(dev->scratch + 256);
    GridBuffers g;
    g.count = count;
    g.offset = offset;
    g.cell_id = cell_id;
    g.sorted_slot = sorted_slot;
    g.sorted = nullptr;
    g.n_valid = reinterpret_cast<uint32_t*>(dev->scratch + 512);
    g.temp = reinterpret_cast<uint32_t*>(dev->scratch + 768);
    launch_store_frame(fc, tab, frame, s);
    FrameArgs fa{tab, frame, 0u};
    if (grid_build(photons, n, 1, hash_size, fa, g, s)) return fail(WSRT_E_HIP, "grid build failed");
    // slots past the valid photons are not part of the grid
    if (n) launch_fill_after(sorted_slot, g.n_valid, n, kInvalid, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(WSRT_E_HIP, "grid build launch: %s", hipGetErrorString(e));
    return WSRT_OK;
}

static int scene_query(const wsrt_scene* s, const float* rays, uint32_t n, int mode, uint32_t* out0,
                       size_t out0_words, uint32_t* out1) {
    CHECK_ARG(s && (rays || n == 0) && (out0 || n == 0));
    if (n == 0) return WSRT_OK;
    if (n > (1u << 26)) return fail(WSRT_E_INVALID, "%u rays > 2^26", n);
    HIP_OR_FAIL(hipSetDevice(s->dev->ordinal));
    hipStream_t st = s->dev->stream;
    DevBuf dr, d0, d1, gs;
    HIP_OR_FAIL(dr.alloc(32ull * n));
    HIP_OR_FAIL(d0.alloc(4ull * out0_words * n));
    HIP_OR_FAIL(d1.alloc(4ull * n));
    const uint32_t spill = s->view.stack_size > (uint32_t)kShortStack ? s->view.stack_size - kShortStack : 0u;
    HIP_OR_FAIL(gs.alloc(sizeof(uint2) * kTraceLanes * (spill ? spill : 1u)));
    HIP_OR_FAIL(hipMemcpyAsync(dr.p, rays, 32ull * n, hipMemcpyHostToDevice, st));
    launch_query(s->view, dr.as<float>(), n, mode, d0.as<uint32_t>(), d1.as<uint32_t>(), gs.as<uint2>(), st);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipMemcpyAsync(out0, d0.p, 4ull * out0_words * n, hipMemcpyDeviceToHost, st));
    if (out1) HIP_OR_FAIL(hipMemcpyAsync(out1, d1.p, 4ull * n, hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    return WSRT_OK;
}

extern "C" int wsrt_debug_closest_hit(const wsrt_scene* s, const float* rays, uint32_t n,
                                      uint32_t* hits, uint32_t* flags) {
    CHECK_ARG(flags || n == 0);
    return scene_query(s, rays, n, 0, hits, 4, flags);
}
extern "C" int wsrt_debug_any_hit(const wsrt_scene* s, const float* rays, uint32_t n, uint32_t* occ) {
    return scene_query(s, rays, n, 1, occ, 1, nullptr);
}
extern "C" int wsrt_debug_traversal_counts(const wsrt_scene* s, const float* rays, uint32_t n,
                                           uint32_t* nodes, uint32_t* tris) {
    CHECK_ARG(tris || n == 0);
    return scene_query(s, rays, n, 2, nodes, 1, tris);
}

extern "C" int wsrt_debug_read_morton(const wsrt_scene* s, uint32_t* codes, uint32_t* sorted_tris) {
    CHECK_ARG(s);
    HIP_OR_FAIL(hipSetDevice(s->dev->ordinal));
    size_t n = s->info.n_triangles;
    if (codes && n) HIP_OR_FAIL(hipMemcpy(codes, s->codes.p, 4 * n, hipMemcpyDeviceToHost));
    if (sorted_tris && n) HIP_OR_FAIL(hipMemcpy(sorted_tris, s->sorted.p, 4 * n, hipMemcpyDeviceToHost));
    return WSRT_OK;
}
extern "C" int wsrt_debug_read_bvh(const wsrt_scene* s, void* nodes, void* tris) {
    CHECK_ARG(s);
    HIP_OR_FAIL(hipSetDevice(s->dev->ordinal));
    if (nodes && s->nodes.bytes) HIP_OR_FAIL(hipMemcpy(nodes, s->nodes.p, s->nodes.bytes, hipMemcpyDeviceToHost));
    if (tris && s->tris.bytes) HIP_OR_FAIL(hipMemcpy(tris, s->tris.p, s->tris.bytes, hipMemcpyDeviceToHost));
    return WSRT_OK;
}
extern "C" int wsrt_debug_read_world(const wsrt_scene* s, float* tris, uint32_t* ids) {
    CHECK_ARG(s);
    HIP_OR_FAIL(hipSetDevice(s->dev->ordinal));
    size_t n = s->info.n_triangles;
    if (tris && n) HIP_OR_FAIL(hipMemcpy(tris, s->wv.p, 36 * n, hipMemcpyDeviceToHost));
    if (ids && n) {
        std::vector<uint32_t> ti(n);
        std::vector<Instance> inst(s->info.n_instances);
        HIP_OR_FAIL(hipMemcpy(ti.data(), s->tri_inst.p, 4 * n, hipMemcpyDeviceToHost));
        HIP_OR_FAIL(hipMemcpy(inst.data(), s->instances.p, sizeof(Instance) * inst.size(), hipMemcpyDeviceToHost));
        for (size_t t = 0; t < n; ++t) {
            ids[3 * t] = ti[t];
            ids[3 * t + 1] = inst[ti[t]].material;
            ids[3 * t + 2] = (uint32_t)t - inst[ti[t]].first_tri;
        }
    }
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// path-guiding network
// ---------------------------------------------------------------------------
extern "C" int wsrt_guide_create(wsrt_device* dev, uint32_t max_samples, uint32_t seed, wsrt_guide** out) {
    CHECK_ARG(dev && out && max_samples >= 1 && max_samples <= (1u << 22));
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    auto g = std::make_unique<wsrt_guide>();
    g->dev = dev;
    HIP_OR_FAIL(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    g->own_stream = true;
    int rc = guide_setup(g.get(), max_samples, seed, 1e-3f, 0.5f);
    if (rc) {
        (void)hipStreamDestroy(g->stream);
        return rc;
    }
    dev->live++;
    *out = g.release();
    return WSRT_OK;
}

extern "C" int wsrt_guide_destroy(wsrt_guide* g) {
    if (!g) return WSRT_OK;
    if (g->owned_by_renderer) return fail(WSRT_E_STATE, "the guide belongs to a renderer");
    (void)hipSetDevice(g->dev->ordinal);
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
    g->dev->live--;
    delete g;
    return WSRT_OK;
}

extern "C" int wsrt_guide_get_params(wsrt_guide* g, float* out) {
    CHECK_ARG(g && out);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(out, g->params.p, 4ull * guide::kParams, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_guide_get_state(wsrt_guide* g, float* params, float* m1, float* m2, uint32_t* step) {
    CHECK_ARG(g && params && m1 && m2 && step);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(params, g->params.p, 4ull * guide::kParams, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(m1, g->m1.p, 4ull * guide::kParams, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(m2, g->m2.p, 4ull * guide::kParams, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(step, g->step.p, 4, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_guide_set_state(wsrt_guide* g, const float* params, const float* m1, const float* m2,
                                    uint32_t step) {
    CHECK_ARG(g && params && m1 && m2);
    int rc = wsrt_guide_set_params(g, params);  // packs the operands, zeroes the padding
    if (rc) return rc;
    HIP_OR_FAIL(hipMemcpyAsync(g->m1.p, m1, 4ull * guide::kParams, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(g->m2.p, m2, 4ull * guide::kParams, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(g->step.p, &step, 4, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_guide_get_average(wsrt_guide* g, float* out) {
    CHECK_ARG(g && out);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(out, g->params_avg.p, 4ull * guide::kParams, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_guide_set_average(wsrt_guide* g, const float* in) {
    CHECK_ARG(g && in);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    std::vector<float> w(in, in + guide::kParams);
    for (int r = 0; r < guide::kHid; ++r)
        for (int c = guide::kFeat; c < guide::kIn; ++c) w[guide::kW1 + r * guide::kIn + c] = 0.0f;
    for (int r = guide::kOut; r < guide::kOutPad; ++r) {
        for (int c = 0; c < guide::kHid; ++c) w[guide::kW5 + r * guide::kHid + c] = 0.0f;
        w[guide::kB5 + r] = 0.0f;
    }
    HIP_OR_FAIL(hipMemcpyAsync(g->params_avg.p, w.data(), 4ull * guide::kParams, hipMemcpyHostToDevice, g->stream));
    GuideView v = g->gv;  // pack the forward operands of the average only
    v.wf = g->wf_avg.as<uint16_t>();
    v.wt = nullptr;
    launch_guide_pack(v, g->params_avg.as<float>(), g->stream);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_renderer_get_guide_frames_trained(const wsrt_renderer* r, uint64_t* out) {
    CHECK_ARG(r && out);
    *out = r->frames_trained;
    return WSRT_OK;
}

extern "C" int wsrt_renderer_set_guide_frames_trained(wsrt_renderer* r, uint64_t frames) {
    CHECK_ARG(r);
    if (!r->guide) return fail(WSRT_E_STATE, "the renderer was created without guiding");
    r->frames_trained = frames;
    return WSRT_OK;
}

extern "C" int wsrt_guide_set_params(wsrt_guide* g, const float* in) {
    CHECK_ARG(g && in);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    std::vector<float> w(in, in + guide::kParams);
    // padding (W1 columns >= 168, output rows >= 40) stays zero
    for (int r = 0; r < guide::kHid; ++r)
        for (int c = guide::kFeat; c < guide::kIn; ++c) w[guide::kW1 + r * guide::kIn + c] = 0.0f;
    for (int r = guide::kOut; r < guide::kOutPad; ++r) {
        for (int c = 0; c < guide::kHid; ++c) w[guide::kW5 + r * guide::kHid + c] = 0.0f;
        w[guide::kB5 + r] = 0.0f;
    }
    HIP_OR_FAIL(hipMemcpyAsync(g->params.p, w.data(), 4ull * guide::kParams, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipMemsetAsync(g->m1.p, 0, g->m1.bytes, g->stream));
    HIP_OR_FAIL(hipMemsetAsync(g->m2.p, 0, g->m2.bytes, g->stream));
    HIP_OR_FAIL(hipMemsetAsync(g->step.p, 0, 4, g->stream));
    launch_guide_pack(g->gv, g->params.as<float>(), g->stream);
    HIP_OR_FAIL(guide_reset_average(g));
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

namespace {
// Uploads n samples (scalars n x 7, optional dirs n x 3 and weights n),
// zero-padded to ns, into the guide's training inputs.
int guide_upload(wsrt_guide* g, const float* scalars, const float* dirs, const float* weights, uint32_t n,
                 uint32_t ns) {
    std::vector<float4> a(ns), b(ns), d(ns);
    for (uint32_t i = 0; i < ns; ++i) {
        if (i < n) {
            const float* x = scalars + 7ull * i;
            a[i] = make_float4(x[0], x[1], x[2], x[3]);
            b[i] = make_float4(x[4], x[5], x[6], weights ? weights[i] : 0.0f);
            d[i] = dirs ? make_float4(dirs[3ull * i], dirs[3ull * i + 1], dirs[3ull * i + 2], 0.0f)
                        : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        } else {
            a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            b[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            d[i] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        }
    }
    HIP_OR_FAIL(hipMemcpyAsync(g->tin0.p, a.data(), 16ull * ns, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(g->tin1.p, b.data(), 16ull * ns, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipMemcpyAsync(g->tdir.p, d.data(), 16ull * ns, hipMemcpyHostToDevice, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}
}  // namespace

extern "C" int wsrt_guide_forward(wsrt_guide* g, const float* scalars, uint32_t n, float* out) {
    CHECK_ARG(g && (scalars || n == 0) && (out || n == 0));
    if (n == 0) return WSRT_OK;
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    if (n > g->ns_max) return fail(WSRT_E_INVALID, "n %u > max_samples %u", n, g->ns_max);
    int rc = guide_upload(g, scalars, nullptr, nullptr, n, n);
    if (rc) return rc;
    HIP_OR_FAIL(g->out.alloc(4ull * guide::kOut * n));
    launch_guide_infer(g->gv, g->tin0.as<float4>(), g->tin1.as<float4>(), n, g->out.as<float>(), g->stream);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipMemcpyAsync(out, g->out.p, 4ull * guide::kOut * n, hipMemcpyDeviceToHost, g->stream));
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

static int guide_train_call(wsrt_guide* g, const float* scalars, const float* dirs, const float* weights,
                            uint32_t n, bool update, float lr, float* grads) {
    CHECK_ARG(g && scalars && dirs && weights && n > 0);
    HIP_OR_FAIL(hipSetDevice(g->dev->ordinal));
    uint32_t ns, chunk;
    guide_step_shape(n, &ns, &chunk);
    if (ns > g->ns_max) return fail(WSRT_E_INVALID, "n %u > max_samples %u", n, g->ns_max);
    int rc = guide_upload(g, scalars, dirs, weights, n, ns);
    if (rc) return rc;
    GuideTrain gt = g->gt;
    gt.chunk = chunk;
    if (update) gt.lr = lr;
    launch_guide_train_step(g->gv, gt, g->tin0.as<float4>(), g->tin1.as<float4>(), g->tdir.as<float4>(), ns,
                            1.0f / (float)n, update, g->stream);
    HIP_OR_FAIL(hipGetLastError());
    if (grads) {
        const uint32_t nc = ns / chunk;
        std::vector<float> part((size_t)nc * guide::kParams);
        HIP_OR_FAIL(hipMemcpyAsync(part.data(), g->partial.p, 4ull * part.size(), hipMemcpyDeviceToHost,
                                   g->stream));
        HIP_OR_FAIL(hipStreamSynchronize(g->stream));
        for (int i = 0; i < guide::kParams; ++i) {
            double acc = 0.0;
            for (uint32_t c = 0; c < nc; ++c) acc += part[(size_t)c * guide::kParams + i];
            grads[i] = (float)acc;
        }
    }
    HIP_OR_FAIL(hipStreamSynchronize(g->stream));
    return WSRT_OK;
}

extern "C" int wsrt_guide_gradients(wsrt_guide* g, const float* scalars, const float* dirs,
                                    const float* weights, uint32_t n, float* grads) {
    CHECK_ARG(grads);
    return guide_train_call(g, scalars, dirs, weights, n, false, 0.0f, grads);
}

extern "C" int wsrt_guide_train_step(wsrt_guide* g, const float* scalars, const float* dirs,
                                     const float* weights, uint32_t n, float lr) {
    CHECK_ARG(lr > 0.0f);
    return guide_train_call(g, scalars, dirs, weights, n, true, lr, nullptr);
}

extern "C" int wsrt_renderer_get_guide(wsrt_renderer* r, wsrt_guide** out) {
    CHECK_ARG(r && out);
    if (!r->guide) return fail(WSRT_E_STATE, "the renderer was created without guiding");
    HIP_OR_FAIL(hipStreamSynchronize(r->stream));
    *out = r->guide.get();
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// autotuner: frames in flight
// ---------------------------------------------------------------------------
// Milliseconds per frame of a renderer with config c (first batch warms and
// captures, then 2 batches timed with events), and its device memory; 0 ms
// when it does not fit the budget or fails.
static float autotune_time(wsrt_device* dev, const wsrt_scene* scene, uint32_t width, uint32_t height,
                           const wsrt_camera_uniform* cam, const wsrt_config& c, uint64_t mem_budget,
                           uint64_t* bytes) {
    size_t fa = 0, fb = 0, total = 0;
    (void)hipMemGetInfo(&fa, &total);
    wsrt_renderer* r = nullptr;
    *bytes = 0;
    if (wsrt_renderer_create(dev, scene, width, height, &c, &r) != WSRT_OK) return 0.0f;
    (void)hipMemGetInfo(&fb, &total);
    *bytes = fa > fb ? (uint64_t)(fa - fb) : 0;
    const uint32_t S = c.frames_in_flight;
    bool ok = *bytes <= mem_budget;
    if (ok && cam) ok = wsrt_renderer_set_camera(r, cam) == WSRT_OK;
    float ms = 0.0f;
    if (ok) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        ok = ok && wsrt_renderer_render_frames(r, S) == WSRT_OK;  // first batch captures the graph
        ok = ok && hipEventRecord(e0, r->stream) == hipSuccess;
        ok = ok && wsrt_renderer_render_frames(r, 2 * S) == WSRT_OK;
        ok = ok && hipEventRecord(e1, r->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) ms /= (float)(2 * S);
        else ms = 0.0f;
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    wsrt_renderer_destroy(r);
    (void)hipGetLastError();
    return ok ? ms : 0.0f;
}

extern "C" int wsrt_autotune(wsrt_device* dev, const wsrt_scene* scene, uint32_t width, uint32_t height,
                             const wsrt_camera_uniform* cam, wsrt_config* cfg, uint64_t mem_budget,
                             wsrt_autotune_result* out) {
    CHECK_ARG(dev && scene && cfg && out && width > 0 && height > 0);
    int rc = validate_config(cfg);
    if (rc) return rc;
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    size_t free0 = 0, total = 0;
    HIP_OR_FAIL(hipMemGetInfo(&free0, &total));
    if (mem_budget == 0) mem_budget = total / 4;
    memset(out, 0, sizeof *out);
    const uint32_t cands[9] = {1, 2, 4, 8, 16, 32, 64, 128, 256};
    float best_ms = 0.0f;
    uint32_t best = cfg->frames_in_flight, slower = 0;
    for (uint32_t k = 0; k < 9; ++k) {
        const uint32_t S = cands[k];
        out->frames_in_flight[k] = S;
        out->n_candidates = k + 1;
        wsrt_config c = *cfg;
        c.frames_in_flight = S;
        if (validate_config(&c) != WSRT_OK) break;
        const float ms = autotune_time(dev, scene, width, height, cam, c, mem_budget, &out->device_bytes[k]);
        if (ms == 0.0f) break;
        out->ms_per_frame[k] = ms;
        if (best_ms == 0.0f || ms < best_ms * 0.98f) {  // a larger batch must win by 2 %
            best_ms = ms;
            best = S;
            slower = 0;
        } else if (++slower == 2) {
            break;
        }
    }
    if (best_ms == 0.0f) return fail(WSRT_E_OOM, "no frames_in_flight candidate fits the memory budget");
    out->best = best;
    cfg->frames_in_flight = best;
    // second stage at the chosen frames in flight: the tail launch's queue
    // threshold and the traversal grids; the config as given (a measured
    // first-stage candidate) is kept unless another is 2 % faster
    const uint32_t tails[4] = {0u, 1u << 16, 1u << 18, 1u << 20};
    float best2 = best_ms;
    uint32_t bt = cfg->tail_paths, br = cfg->trace_rounds ? cfg->trace_rounds : 1u;
    for (uint32_t rounds = 1; rounds <= 2; ++rounds) {
        for (uint32_t t = 0; t < 4; ++t) {
            if (rounds == br && tails[t] == bt) continue;  // measured in the first stage
            wsrt_config c = *cfg;
            c.tail_paths = tails[t];
            c.trace_rounds = rounds;
            const uint32_t i = out->n_stage2++;
            out->tail_paths[i] = tails[t];
            out->trace_rounds[i] = rounds;
            uint64_t bytes = 0;
            const float ms = autotune_time(dev, scene, width, height, cam, c, mem_budget, &bytes);
            out->stage2_ms_per_frame[i] = ms;
            if (ms > 0.0f && ms < best2 * 0.98f) {
                best2 = ms;
                bt = tails[t];
                br = rounds;
            }
        }
    }
    cfg->tail_paths = bt;
    cfg->trace_rounds = br;
    out->best_tail_paths = bt;
    out->best_trace_rounds = br;
    // third stage: lane refill in the closest-hit kernel, which pays where
    // ray costs within a wave differ widely (occluded interiors)
    // and path-ordered queues, which shorten shading where paths scatter
    // widely
    const uint32_t refills[4] = {0u, 16u, 24u, 32u};
    float best3 = best2;
    uint32_t bx = cfg->extend_refill, bo = cfg->ordered_queues;
    for (uint32_t o = 0; o < 2; ++o) {
        for (uint32_t k = 0; k < 4; ++k) {
            if (refills[k] == cfg->extend_refill && o == cfg->ordered_queues) continue;  // measured above
            wsrt_config c = *cfg;
            c.extend_refill = refills[k];
            c.ordered_queues = o;
            const uint32_t i = out->n_stage3++;
            out->extend_refill[i] = refills[k];
            out->ordered_queues[i] = o;
            uint64_t bytes = 0;
            const float ms = autotune_time(dev, scene, width, height, cam, c, mem_budget, &bytes);
            out->stage3_ms_per_frame[i] = ms;
            if (ms > 0.0f && ms < best3 * 0.98f) {
                best3 = ms;
                bx = refills[k];
                bo = o;
            }
        }
    }
    cfg->extend_refill = bx;
    cfg->ordered_queues = bo;
    out->best_extend_refill = bx;
    out->best_ordered_queues = bo;
    return WSRT_OK;
}

// ---------------------------------------------------------------------------
// persistent multi-scene batch
// ---------------------------------------------------------------------------
struct wsrt_batch {
    wsrt_device* dev = nullptr;
    hipStream_t stream = nullptr;
    wsrt_config cfg{};
    std::vector<wsrt_batch_job> jobs;
    std::vector<BatchJob> hjobs;  // host copies; first_item follows the tile mode
    std::vector<DevBuf> accum, rgba;
    DevBuf djobs, djobs_tune, table, head, gstack;
    BatchArgs args{};
    wsrt_batch_tuning tuning{1u, 2u};
    int cus = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // photon pass (cfg.enable_photons): each job's grid of the frame being
    // traced, and the trace buffers the jobs' passes share (one job at a time)
    bool photons = false;
    uint32_t max_spp = 0;
    uint32_t photon_frames = 1;  // frames per window (batch_alloc_photons)
    std::vector<DevBuf> gcount, goffset, gsorted, gtouched, gnvalid;
    DevBuf ph, ph_stored, g_cell, g_slot, g_temp, qcounts, ctr, pgstack, dframe;
    DevBuf ph_id[2], ph_pos[2], ph_dir[2], ph_rng[2], ph_flags[2];
    PhotonState phs{};
    Queues qs{};
    std::vector<std::pair<hipGraph_t, hipGraphExec_t>> frame_graph;  // by frame, photon mode
    // stopping a render: the pinned stop flag, the launch's abort counter and
    // start time, and the device's wall-clock rate (kHz)
    uint32_t* stop_host = nullptr;
    DevBuf abort_buf;
    int wall_khz = 100000;
};

// Most frames of every job one photon pass and one persistent launch cover
// in photon mode (the window). Each job's pass and grid build is about 15
// small launches, so a window of W frames takes 1/W of the per-frame
// launches: on the 64-scene batch at 2^20 photons per frame, 339 Msamples/s
// frame by frame, 634 with windows of 4, 725 of 8, 791 of 16. The window
// holds one grid per frame and job (68 MB per job-frame at 2^20 photons and
// 2^22 cells), so batch_alloc_photons picks the largest window up to this
// whose grids fit in a quarter of the free device memory.
constexpr uint32_t kBatchPhotonFramesMax = 16;

static void batch_drop_graph(wsrt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (auto& g : b->frame_graph) {
        if (g.second) (void)hipGraphExecDestroy(g.second);
        if (g.first) (void)hipGraphDestroy(g.first);
    }
    b->frame_graph.clear();
}

// Frames [f0, f0 + photon_frames) of every job that renders them: each
// job's photon pass and grid build for those frames at once (photon_emit.wgsl,
// build_grid.wgsl: the renderer's kernels with that many frames in flight,
// one grid per frame), then one persistent launch tracing the window's frames
// of all jobs' pixels, each pixel's samples in frame order.
static void batch_enqueue_frames(wsrt_batch* b, uint32_t f0, const BatchArgs& base, uint32_t blocks) {
    hipStream_t s = b->stream;
    const uint32_t np = b->cfg.photons_per_pass;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const BatchJob& B = b->hjobs[j];
        if (f0 >= B.frames) continue;
        const uint32_t nf = std::min(b->photon_frames, B.frames - f0);
        const FrameArgs fa{b->table.as<FrameConst>() + B.table_base, b->dframe.as<uint32_t>(), 0xFFFFFFFFu};
        launch_set_frame(b->dframe.as<uint32_t>(), f0, s);
        launch_clear_queues(b->qs, b->cfg.photon_max_depth + 1, s);
        launch_photon_emit(B.sv, fa, np, nf, b->qs, b->phs, 0u, b->ctr.as<Counters>(), s);
        for (uint32_t d = 0; d < b->cfg.photon_max_depth; ++d)
            launch_photon_bounce(B.sv, d, 0u, b->phs, b->qs, b->ph.as<wsrt_photon>(), b->ctr.as<Counters>(),
                                 b->pgstack.as<uint2>(), s);
        GridBuffers g;
        g.count = b->gcount[j].as<uint32_t>();
        g.offset = b->goffset[j].as<uint32_t>();
        g.cell_id = b->g_cell.as<uint32_t>();
        g.sorted_slot = b->g_slot.as<uint32_t>();
        g.sorted = b->gsorted[j].as<wsrt_photon>();
        g.n_valid = b->gnvalid[j].as<uint32_t>();
        g.temp = b->g_temp.as<uint32_t>();
        g.touched = b->gtouched[j].as<uint32_t>();
        grid_build_stored(b->ph.as<wsrt_photon>(), b->phs.stored, np, nf, b->cfg.hash_size, fa, g, s);
    }
    BatchArgs a = base;
    a.one_frame = 1u;
    a.frame = f0;
    a.frame_span = b->photon_frames;
    launch_batch(a, blocks, s, f0 == 0);
}

// The photon pass's buffers: per job, the grids of a window of
// photon_frames frames (counts, offsets, photons in cell order, the cells the
// last build touched); shared, the photon trace state of one pass over a
// window and the grid build's work memory.
static int batch_alloc_photons(wsrt_batch* b, uint32_t max_stack) {
    const wsrt_config& c = b->cfg;
    const uint32_t np1 = c.photons_per_pass ? c.photons_per_pass : 1u;
    const size_t nj = b->jobs.size();
    {
        // the window: the most frames (up to kBatchPhotonFramesMax, within
        // 2^30 photons and cells) whose grids take at most a quarter of the
        // free device memory (or of what a memory cap leaves)
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        const uint64_t cap = g_mem_limit[b->dev->ordinal].load(), used = g_mem_used[b->dev->ordinal].load();
        if (cap) free_b = std::min<uint64_t>(free_b, cap > used ? cap - used : 0u);
        const uint64_t per_frame =
            (uint64_t)nj * (8ull * c.hash_size + (sizeof(wsrt_photon) + 4ull) * np1) + 160ull * np1;
        uint32_t w = kBatchPhotonFramesMax;
        while (w > 1 && ((uint64_t)w * per_frame > free_b / 4 || (uint64_t)np1 * w > (1ull << 30) ||
                         (uint64_t)c.hash_size * w > (1ull << 30)))
            w >>= 1;
        b->photon_frames = w;
    }
    const uint32_t np = np1 * b->photon_frames;          // photon paths of one window
    const uint32_t cells = c.hash_size * b->photon_frames;  // grid cells of one window
    b->gcount.resize(nj);
    b->goffset.resize(nj);
    b->gsorted.resize(nj);
    b->gtouched.resize(nj);
    b->gnvalid.resize(nj);
    for (size_t j = 0; j < nj; ++j) {
        HIP_OR_FAIL(b->gcount[j].alloc(4ull * cells));
        HIP_OR_FAIL(b->goffset[j].alloc(4ull * cells));
        HIP_OR_FAIL(b->gsorted[j].alloc(sizeof(wsrt_photon) * np));
        HIP_OR_FAIL(b->gtouched[j].alloc(4ull * np));
        HIP_OR_FAIL(b->gnvalid[j].alloc(4));
        HIP_OR_FAIL(hipMemsetAsync(b->gcount[j].p, 0, 4ull * cells, b->stream));
        HIP_OR_FAIL(hipMemsetAsync(b->goffset[j].p, 0, 4ull * cells, b->stream));
        HIP_OR_FAIL(hipMemsetAsync(b->gnvalid[j].p, 0, 4, b->stream));
        BatchJob& B = b->hjobs[j];
        B.grid = GridView{b->gcount[j].as<uint32_t>(), b->goffset[j].as<uint32_t>(),
                          b->gsorted[j].as<wsrt_photon>(), c.hash_size, c.cell_size};
    }
    HIP_OR_FAIL(b->ph.alloc(sizeof(wsrt_photon) * np));
    HIP_OR_FAIL(b->ph_stored.alloc(4ull * ((np + 31) / 32)));
    HIP_OR_FAIL(b->g_cell.alloc(4ull * np));
    HIP_OR_FAIL(b->g_slot.alloc(4ull * np));
    HIP_OR_FAIL(b->g_temp.alloc(4 * grid_temp_words(np, cells)));
    const uint64_t pq = (uint64_t)kSegs * seg_capacity(np);
    for (int k = 0; k < 2; ++k) {
        HIP_OR_FAIL(b->ph_id[k].alloc(4 * pq));
        HIP_OR_FAIL(b->ph_pos[k].alloc(16 * pq));
        HIP_OR_FAIL(b->ph_dir[k].alloc(16 * pq));
        HIP_OR_FAIL(b->ph_rng[k].alloc(4 * pq));
        HIP_OR_FAIL(b->ph_flags[k].alloc(4 * pq));
        b->phs.q[k] = PhotonQueue{b->ph_id[k].as<uint32_t>(), b->ph_pos[k].as<float4>(), b->ph_dir[k].as<float4>(),
                                  b->ph_rng[k].as<uint32_t>(), b->ph_flags[k].as<uint32_t>()};
    }
    b->phs.stored = b->ph_stored.as<uint32_t>();
    const uint32_t depths = c.photon_max_depth + 1;
    HIP_OR_FAIL(b->qcounts.alloc(4 * (queue_count_words(depths) + kSegStride)));
    HIP_OR_FAIL(hipMemsetAsync(b->qcounts.p, 0, b->qcounts.bytes, b->stream));
    b->qs.counts = b->qcounts.as<uint32_t>();
    b->qs.seg_cap = 0;
    b->qs.pseg_cap = seg_capacity(np);
    b->qs.tail = nullptr;
    b->qs.trace_rounds = c.trace_rounds ? c.trace_rounds : 1u;
    HIP_OR_FAIL(b->ctr.alloc(sizeof(Counters)));
    HIP_OR_FAIL(hipMemsetAsync(b->ctr.p, 0, sizeof(Counters), b->stream));
    const uint32_t spill = max_stack > (uint32_t)kShortStack ? max_stack - kShortStack : 1u;
    HIP_OR_FAIL(b->pgstack.alloc(sizeof(uint2) * kTraceLanes * spill));
    HIP_OR_FAIL(b->dframe.alloc(4));
    return WSRT_OK;
}

static void batch_release(wsrt_batch* b) {
    (void)hipSetDevice(b->dev->ordinal);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (const wsrt_batch_job& J : b->jobs) const_cast<wsrt_scene*>(J.scene)->renderers--;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->stop_host) (void)hipHostFree(b->stop_host);
    b->dev->live--;
    delete b;
}

extern "C" int wsrt_batch_set_tuning(wsrt_batch* b, const wsrt_batch_tuning* t) {
    CHECK_ARG(b && t && t->tile_mode <= 1u && t->blocks_per_cu >= 1u && t->blocks_per_cu <= 8u);
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    uint32_t items = 0;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        BatchJob& B = b->hjobs[j];
        B.first_item = items;
        const uint64_t n = t->tile_mode == 0u ? ((uint64_t)B.width * B.height + 63u) / 64u
                                              : (uint64_t)B.tiles_x * ((B.height + 7u) / 8u);
        items += (uint32_t)n;
    }
    std::vector<BatchJob> tune = b->hjobs;
    for (BatchJob& B : tune) B.frames = 1;
    HIP_OR_FAIL(hipMemcpyAsync(b->djobs.p, b->hjobs.data(), sizeof(BatchJob) * b->hjobs.size(),
                               hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipMemcpyAsync(b->djobs_tune.p, tune.data(), sizeof(BatchJob) * tune.size(), hipMemcpyHostToDevice,
                               b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    b->args.n_items = items;
    b->args.tile_mode = t->tile_mode;
    b->tuning = *t;
    batch_drop_graph(b);
    return WSRT_OK;
}

extern "C" int wsrt_batch_create(wsrt_device* dev, const wsrt_batch_job* jobs, uint32_t n_jobs,
                                 const wsrt_config* cfg, wsrt_batch** out) {
    CHECK_ARG(dev && jobs && out && n_jobs > 0 && n_jobs <= kMaxBatchJobs);
    wsrt_config c;
    if (cfg) c = *cfg;
    else wsrt_config_default(&c);
    int rc = validate_config(&c);
    if (rc) return rc;
    if (c.guiding) return fail(WSRT_E_UNSUPPORTED, "batches trace without guiding");
    uint64_t total_frames = 0;
    uint32_t max_stack = 1;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const wsrt_batch_job& J = jobs[j];
        CHECK_ARG(J.scene && J.scene->dev == dev && J.width > 0 && J.height > 0 && J.spp > 0);
        if ((uint64_t)J.width * J.height > (1ull << 26)) return fail(WSRT_E_INVALID, "job %u too large", j);
        if (c.env_map && !J.scene->view.env) return fail(WSRT_E_INVALID, "job %u: scene without environment", j);
        total_frames += J.spp;
        max_stack = std::max(max_stack, J.scene->view.stack_size);
    }
    if (total_frames > (1u << 20)) return fail(WSRT_E_INVALID, "too many samples per pixel over the jobs");
    HIP_OR_FAIL(hipSetDevice(dev->ordinal));
    wsrt_batch* b = new wsrt_batch();
    b->dev = dev;
    b->cfg = c;
    b->jobs.assign(jobs, jobs + n_jobs);
    for (const wsrt_batch_job& J : b->jobs) const_cast<wsrt_scene*>(J.scene)->renderers++;
    dev->live++;
    auto bail = [b](int code) {
        batch_release(b);
        return code;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(WSRT_E_HIP, "stream creation failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev->ordinal) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    std::vector<FrameConst> table;
    table.reserve(total_frames);
    b->accum.resize(n_jobs);
    b->rgba.resize(n_jobs);
    b->hjobs.resize(n_jobs);
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const wsrt_batch_job& J = jobs[j];
        const size_t npx = (size_t)J.width * J.height;
        if (b->accum[j].alloc(16 * npx) != hipSuccess || b->rgba[j].alloc(4 * npx) != hipSuccess ||
            hipMemsetAsync(b->accum[j].p, 0, 16 * npx, b->stream) != hipSuccess)
            return bail(fail(WSRT_E_OOM, "batch images"));
        BatchJob& B = b->hjobs[j];
        B.sv = J.scene->view;
        if (!c.env_map) B.sv.env = nullptr;
        B.accum = b->accum[j].as<float4>();
        B.rgba = b->rgba[j].as<uint8_t>();
        B.width = J.width;
        B.height = J.height;
        B.frames = J.spp;
        B.table_base = (uint32_t)table.size();
        B.tiles_x = (J.width + 7u) / 8u;
        B.first_item = 0;
        B.grid = GridView{nullptr, nullptr, nullptr, c.hash_size, c.cell_size};
        b->max_spp = std::max(b->max_spp, J.spp);
        for (uint32_t f = 0; f < J.spp; ++f) {
            wsrt_camera_uniform cam = J.camera;
            cam.frame_count = f;
            table.push_back(make_frame_const(c, cam, J.width, J.height));
        }
    }
    const uint32_t spill = max_stack > (uint32_t)kShortStack ? max_stack - kShortStack : 1u;
    const uint64_t lanes = (uint64_t)b->cus * 8u * 256u;  // the most blocks_per_cu allows
    if (b->djobs.alloc(sizeof(BatchJob) * n_jobs) != hipSuccess ||
        b->djobs_tune.alloc(sizeof(BatchJob) * n_jobs) != hipSuccess ||
        b->table.alloc(sizeof(FrameConst) * table.size()) != hipSuccess || b->head.alloc(4) != hipSuccess ||
        b->gstack.alloc(sizeof(uint2) * lanes * spill) != hipSuccess)
        return bail(fail(WSRT_E_OOM, "batch buffers"));
    if (hipMemcpyAsync(b->table.p, table.data(), sizeof(FrameConst) * table.size(), hipMemcpyHostToDevice,
                       b->stream) != hipSuccess)
        return bail(fail(WSRT_E_HIP, "frame table upload"));
    b->args.jobs = b->djobs.as<BatchJob>();
    b->args.njobs = n_jobs;
    b->args.table = b->table.as<FrameConst>();
    b->args.head = b->head.as<uint32_t>();
    b->args.gstack = b->gstack.as<uint2>();
    b->args.spill = spill;
    b->args.max_depth = c.max_depth;
    b->args.one_frame = 0;
    b->args.frame = 0;
    b->args.frame_span = 1;
    if (b->abort_buf.alloc(16) != hipSuccess || hipMemsetAsync(b->abort_buf.p, 0, 16, b->stream) != hipSuccess)
        return bail(fail(WSRT_E_OOM, "batch abort counter"));
    void* stop_dev = nullptr;
    if (hipHostMalloc((void**)&b->stop_host, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return bail(fail(WSRT_E_OOM, "batch stop flag"));
    __atomic_store_n(b->stop_host, 0u, __ATOMIC_SEQ_CST);
    if (hipHostGetDevicePointer(&stop_dev, b->stop_host, 0) != hipSuccess)
        return bail(fail(WSRT_E_HIP, "batch stop flag mapping"));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev->ordinal) == hipSuccess && khz > 0)
        b->wall_khz = khz;
    b->args.stop = reinterpret_cast<const volatile uint32_t*>(stop_dev);
    b->args.abort = b->abort_buf.as<uint32_t>();
    b->args.time_limit = 0;
    b->photons = c.enable_photons != 0;
    if (b->photons) {
        rc = batch_alloc_photons(b, max_stack);
        if (rc) return bail(rc);
    }
    rc = wsrt_batch_set_tuning(b, &b->tuning);
    if (rc) return bail(rc);
    *out = b;
    return WSRT_OK;
}

extern "C" int wsrt_batch_destroy(wsrt_batch* b) {
    if (!b) return WSRT_OK;
    batch_release(b);
    return WSRT_OK;
}

extern "C" int wsrt_batch_render(wsrt_batch* b) {
    CHECK_ARG(b);
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    const uint32_t blocks = (uint32_t)b->cus * b->tuning.blocks_per_cu;
    __atomic_store_n(b->stop_host, 0u, __ATOMIC_SEQ_CST);
    // the render starts from empty images (a stopped render leaves partial ones)
    for (size_t j = 0; j < b->jobs.size(); ++j)
        HIP_OR_FAIL(hipMemsetAsync(b->accum[j].p, 0, b->accum[j].bytes, b->stream));
    if (b->photons) {
        // window by window of photon_frames frames: every job's photon
        // pass and grids, then one persistent launch over all jobs' pixels;
        // each window one captured graph
        const uint32_t windows = (b->max_spp + b->photon_frames - 1) / b->photon_frames;
        if (b->frame_graph.empty()) b->frame_graph.assign(windows, {nullptr, nullptr});
        for (uint32_t w = 0; w < windows; ++w) {
            const uint32_t f = w * b->photon_frames;
            auto& fg = b->frame_graph[w];
            if (!fg.second && b->cfg.frame_graphs &&
                hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                batch_enqueue_frames(b, f, b->args, blocks);
                if (hipStreamEndCapture(b->stream, &fg.first) != hipSuccess ||
                    hipGraphInstantiate(&fg.second, fg.first, nullptr, nullptr, 0) != hipSuccess) {
                    if (fg.first) (void)hipGraphDestroy(fg.first);
                    fg = {nullptr, nullptr};
                }
                (void)hipGetLastError();
            }
            if (fg.second) HIP_OR_FAIL(hipGraphLaunch(fg.second, b->stream));
            else batch_enqueue_frames(b, f, b->args, blocks);
        }
        HIP_OR_FAIL(hipGetLastError());
        return WSRT_OK;
    }
    if (!b->exec) {
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            launch_batch(b->args, blocks, b->stream, true);
            if (hipStreamEndCapture(b->stream, &b->graph) != hipSuccess ||
                hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
                batch_drop_graph(b);
        }
        (void)hipGetLastError();
    }
    if (b->exec) HIP_OR_FAIL(hipGraphLaunch(b->exec, b->stream));
    else launch_batch(b->args, blocks, b->stream, true);
    HIP_OR_FAIL(hipGetLastError());
    return WSRT_OK;
}

extern "C" int wsrt_batch_sync(wsrt_batch* b) {
    CHECK_ARG(b);
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    uint32_t dropped = 0;
    HIP_OR_FAIL(hipMemcpy(&dropped, b->abort_buf.p, sizeof dropped, hipMemcpyDeviceToHost));
    if (dropped)
        return fail(WSRT_E_ABORTED, "batch render stopped (%s): %u work items not traced",
                    __atomic_load_n(b->stop_host, __ATOMIC_SEQ_CST) ? "canceled" : "time limit", dropped);
    return WSRT_OK;
}

extern "C" int wsrt_batch_cancel(wsrt_batch* b) {
    CHECK_ARG(b && b->stop_host);
    __atomic_store_n(b->stop_host, 1u, __ATOMIC_SEQ_CST);
    return WSRT_OK;
}

extern "C" int wsrt_batch_set_time_limit(wsrt_batch* b, uint64_t microseconds) {
    CHECK_ARG(b);
    if (microseconds > (1ull << 40)) return fail(WSRT_E_INVALID, "time limit %llu us too large",
                                                 (unsigned long long)microseconds);
    const uint64_t ticks = microseconds * (uint64_t)b->wall_khz / 1000u;
    b->args.time_limit = microseconds ? std::max<uint64_t>(ticks, 1u) : 0u;
    batch_drop_graph(b);  // captured launches hold the old limit
    return WSRT_OK;
}

extern "C" int wsrt_batch_autotune(wsrt_batch* b, wsrt_batch_tuning* best, float* best_ms) {
    CHECK_ARG(b);
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_OR_FAIL(hipEventCreate(&e0));
    HIP_OR_FAIL(hipEventCreate(&e1));
    wsrt_batch_tuning bt = b->tuning;
    float bms = 0.0f;
    int rc = WSRT_OK;
    for (uint32_t mode = 0; mode < 2 && rc == WSRT_OK; ++mode) {
        for (uint32_t bpc = 1; bpc <= 8 && rc == WSRT_OK; bpc *= 2) {
            const wsrt_batch_tuning t{mode, bpc};
            rc = wsrt_batch_set_tuning(b, &t);
            if (rc) break;
            BatchArgs a = b->args;
            a.jobs = b->djobs_tune.as<BatchJob>();  // one sample per pixel
            a.stop = nullptr;                       // tuning passes always finish
            a.time_limit = 0;
            const uint32_t blocks = (uint32_t)b->cus * bpc;
            if (b->photons) {
                // the first window's grids, then frame 0 of every job
                batch_enqueue_frames(b, 0, a, blocks);
                a.one_frame = 1u;
                a.frame = 0u;
                a.frame_span = 1u;
            }
            launch_batch(a, blocks, b->stream, true);  // warm
            (void)hipEventRecord(e0, b->stream);
            launch_batch(a, blocks, b->stream, true);
            (void)hipEventRecord(e1, b->stream);
            float ms = 0.0f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(WSRT_E_HIP, "batch autotune timing");
                break;
            }
            if (bms == 0.0f || ms < bms) {
                bms = ms;
                bt = t;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    rc = wsrt_batch_set_tuning(b, &bt);
    if (rc) return rc;
    // the tuning passes wrote frame-0 images: clear them for the real render
    for (size_t j = 0; j < b->jobs.size(); ++j)
        HIP_OR_FAIL(hipMemsetAsync(b->accum[j].p, 0, b->accum[j].bytes, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return WSRT_OK;
}

extern "C" int wsrt_batch_read_accum(wsrt_batch* b, uint32_t job, float* out) {
    CHECK_ARG(b && out && job < b->jobs.size());
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(out, b->accum[job].p, b->accum[job].bytes, hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return WSRT_OK;
}

extern "C" int wsrt_batch_read_rgba8(wsrt_batch* b, uint32_t job, uint8_t* out) {
    CHECK_ARG(b && out && job < b->jobs.size());
    HIP_OR_FAIL(hipSetDevice(b->dev->ordinal));
    HIP_OR_FAIL(hipMemcpyAsync(out, b->rgba[job].p, b->rgba[job].bytes, hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return WSRT_OK;
}
