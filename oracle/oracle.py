"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP path. It restates the
reference's shaders (src/renderer/shaders/*.wgsl) and scene code (src/scene)
on the CPU; nothing in the product library calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_LIB64_PATH = os.path.join(_HERE, "liboracle64.so")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
f32p = _f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compiles liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class CameraUniform(C.Structure):
    _fields_ = [("view_inverse", C.c_float * 16), ("proj_inverse", C.c_float * 16),
                ("view", C.c_float * 16), ("proj", C.c_float * 16),
                ("frame_count", C.c_uint32), ("pad", C.c_uint32 * 3)]


class Config(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_depth", C.c_uint32),
                ("photon_max_depth", C.c_uint32), ("photons_per_pass", C.c_uint32),
                ("hash_size", C.c_uint32), ("cell_size", C.c_float),
                ("gather_radius", C.c_float), ("gather_radius_exponent", C.c_float),
                ("spectral_radius", C.c_float), ("gather_cap", C.c_uint32),
                ("rr_after", C.c_uint32), ("enable_photons", C.c_uint32),
                ("photon_norm", C.c_uint32), ("guiding", C.c_uint32),
                ("guide_mix", C.c_float), ("guide_lr", C.c_float),
                ("guide_train_samples", C.c_uint32), ("env_map", C.c_uint32),
                ("frame_graphs", C.c_uint32), ("frames_in_flight", C.c_uint32),
                ("tail_paths", C.c_uint32), ("quirks", C.c_uint32), ("trace_rounds", C.c_uint32), ("guide_train_frames", C.c_uint32),
                ("extend_refill", C.c_uint32), ("ordered_queues", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


def default_config(**kw) -> Config:
    """The reference's compile-time constants (renderer/mod.rs:18-19,
    compute_pass.rs:26, build_grid.wgsl:23, shader.wgsl:274-342,696)."""
    c = Config(C.sizeof(Config), 12, 5, 1 << 20, 1 << 22, 0.02, 0.02, -0.15, 20.0, 64, 3, 1,
               1 << 20, 0, 0.5, 1e-3, 1 << 18, 0, 1, 8, 1 << 18)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


PHOTON_DTYPE = np.dtype([("position", np.float32, 3), ("wavelength", np.float32),
                         ("direction", np.float32, 3), ("energy", np.float32)])
VERTEX_DTYPE = np.dtype([("pos", np.float32, 4), ("normal", np.float32, 4)])
MATERIAL_DTYPE = np.dtype([("color", np.float32, 4), ("extra", np.float32, 4)])
LIGHT_DTYPE = np.dtype([("v0", np.float32, 4), ("v1", np.float32, 4), ("v2", np.float32, 4),
                        ("params", np.float32, 4)])


def _load(fp64: bool = False):
    """liboracle.so (f32), or with fp64 liboracle64.so: the same source built
    with double arithmetic, whose float arrays and scalars are double."""
    path = _LIB64_PATH if fp64 else os.environ.get("WSRT_ORACLE_LIB", _LIB_PATH)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    f32p = f64p if fp64 else _f32p
    C_float = C.c_double if fp64 else C.c_float

    class _C:  # ctypes with c_float replaced where the build uses double
        c_float = C_float
        c_uint32 = C.c_uint32
        c_int = C.c_int
        POINTER = staticmethod(C.POINTER)
    sig = {
        "or_pcg_hash": (C.c_uint32, [C.c_uint32]),
        "or_rand_stream": (None, [C.c_uint32, C.c_uint32, f32p]),
        "or_cie_xyz_f": (None, [_C.c_float, f32p]),
        "or_rgb_to_spectrum_f": (_C.c_float, [f32p, _C.c_float]),
        "or_blackbody_f": (_C.c_float, [_C.c_float, _C.c_float]),
        "or_thin_film_f": (_C.c_float, [_C.c_float] * 6),
        "or_schlick_f": (_C.c_float, [_C.c_float, _C.c_float]),
        "or_cauchy_constants": (None, [C.POINTER(_C.c_float)] * 3),
        "or_cauchy_sample_f": (_C.c_float, [_C.c_float]),
        "or_grid_jitter_f": (None, [C.c_uint32, _C.c_float, f32p]),
        "or_hash_position_f": (C.c_uint32, [f32p, C.c_uint32, _C.c_float, C.c_uint32]),
        "or_hero_wavelengths": (None, [_C.c_float, _C.c_float, f32p]),
        "or_geometry_plane": (C.c_int, [vp, vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "or_geometry_cube": (C.c_int, [vp, vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "or_geometry_sphere": (C.c_int, [C.c_uint32, vp, vp, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32)]),
        "or_mat4_identity": (None, [f32p]),
        "or_mat4_mul": (None, [f32p, f32p, f32p]),
        "or_mat4_translation": (None, [_C.c_float] * 3 + [f32p]),
        "or_mat4_scale": (None, [_C.c_float] * 3 + [f32p]),
        "or_mat4_rotation_x": (None, [_C.c_float, f32p]),
        "or_mat4_rotation_y": (None, [_C.c_float, f32p]),
        "or_mat4_rotation_z": (None, [_C.c_float, f32p]),
        "or_camera_uniform": (None, [f32p, _C.c_float, _C.c_float, _C.c_float, C.c_uint32,
                                     C.POINTER(CameraUniform)]),
        "or_scene_create": (vp, []),
        "or_scene_destroy": (None, [vp]),
        "or_scene_add_material": (C.c_uint32, [vp, vp]),
        "or_scene_add_mesh": (C.c_uint32, [vp, vp, C.c_uint32, u32p, C.c_uint32]),
        "or_scene_add_instance": (None, [vp, C.c_uint32, C.c_uint32, f32p]),
        "or_scene_add_light_instance": (None, [vp, C.c_uint32, C.c_uint32, f32p, f32p]),
        "or_scene_add_point_light": (None, [vp, f32p, f32p]),
        "or_scene_set_environment": (None, [vp, f32p, C.c_uint32, C.c_uint32]),
        "or_scene_build": (C.c_int, [vp]),
        "or_scene_num_triangles": (C.c_uint32, [vp]),
        "or_scene_num_lights": (C.c_uint32, [vp]),
        "or_scene_get_lights": (None, [vp, vp]),
        "or_scene_get_world": (None, [vp, vp, vp]),
        "or_scene_cornell_box": (None, [vp]),
        "or_closest_hit": (None, [vp, f32p, C.c_uint32, C.c_int, u32p, u32p]),
        "or_any_hit": (None, [vp, f32p, C.c_uint32, C.c_int, u32p]),
        "or_bvh8_closest": (None, [vp, C.c_uint32, vp, f32p, C.c_uint32, vp, vp, vp, vp]),
        "or_morton_codes": (None, [vp, u32p]),
        "or_photon_trace": (C.c_uint32, [vp, C.POINTER(Config), C.c_uint32, C.c_uint32,
                                         C.c_uint32, vp]),
        "or_grid_counting": (C.c_uint32, [vp, C.c_uint32, C.c_uint32, _C.c_float, C.c_uint32,
                                          u32p, u32p, u32p, u32p]),
        "or_grid_linked": (None, [vp, C.c_uint32, C.c_uint32, _C.c_float, C.c_uint32, u32p,
                                  u32p]),
        "or_frame_create": (vp, [vp, C.POINTER(Config), C.POINTER(CameraUniform), C.c_uint32,
                                 C.c_uint32]),
        "or_frame_create_with_photons": (vp, [vp, C.POINTER(Config), C.POINTER(CameraUniform),
                                              C.c_uint32, C.c_uint32, vp, C.c_uint32]),
        "or_frame_destroy": (None, [vp]),
        "or_frame_photons_stored": (C.c_uint32, [vp]),
        "or_frame_trace_pixels": (None, [vp, u32p, C.c_uint32, vp, vp, vp]),
        "or_frame_render": (None, [vp, f32p, vp]),
        "or_frame_gather": (None, [vp, f32p, f32p, f32p, f32p]),
        "or_unit_vectors_f": (None, [C.c_uint32, C.c_uint32, f32p]),
        "or_sample_bsdf_f": (C.c_int, [f32p, f32p, vp, C.c_int, f32p, f32p, C.c_uint32, C.c_uint32, f32p,
                                       f32p, C.POINTER(C.c_uint32)]),
        "or_sample_light_f": (C.c_int, [vp, f32p, f32p, C.c_uint32, f32p, C.POINTER(_C.c_float),
                                        C.POINTER(_C.c_float), f32p, C.POINTER(C.c_uint32)]),
        "or_photon_emit_f": (None, [vp, C.c_uint32, C.c_uint32, f32p, f32p, C.POINTER(_C.c_float),
                                    C.POINTER(_C.c_float)]),
        "or_ray_color_f": (None, [vp, f32p, f32p, f32p, C.c_uint32, f32p]),
        "or_tone_map_f": (None, [f32p, C.c_uint32, u8p]),
        "or_roulette_f": (C.c_int, [f32p, C.c_uint32, C.c_uint32, f32p, C.POINTER(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
_lib64 = None


def lib64():
    """The double-precision build (loaded on first use)."""
    global _lib64
    if _lib64 is None:
        _lib64 = _load(True)
    return _lib64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pcg_hash(x: int) -> int:
    return lib.or_pcg_hash(x)


def rand_stream(seed_input: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib.or_rand_stream(seed_input, n, out)
    return out


def geometry(kind: str, subdiv: int = 0):
    nv, ni = C.c_uint32(), C.c_uint32()
    fn = {"plane": lib.or_geometry_plane, "cube": lib.or_geometry_cube}.get(kind)
    if kind == "sphere":
        lib.or_geometry_sphere(subdiv, None, None, C.byref(nv), C.byref(ni))
    else:
        fn(None, None, C.byref(nv), C.byref(ni))
    v = np.zeros(nv.value, VERTEX_DTYPE)
    i = np.zeros(ni.value, np.uint32)
    if kind == "sphere":
        rc = lib.or_geometry_sphere(subdiv, _ptr(v), _ptr(i), C.byref(nv), C.byref(ni))
    else:
        rc = fn(_ptr(v), _ptr(i), C.byref(nv), C.byref(ni))
    assert rc == 0, rc
    return v, i


def mat4(kind: str, *args) -> np.ndarray:
    m = np.zeros(16, np.float32)
    getattr(lib, "or_mat4_" + kind)(*args, m)
    return m


def mat4_mul(a, b) -> np.ndarray:
    out = np.zeros(16, np.float32)
    lib.or_mat4_mul(np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32), out)
    return out


def camera_uniform(pos=(0.0, 0.0, 3.0), yaw=None, pitch=0.0, aspect=1.0, frame=0):
    """camera.rs:144-194 defaults: position (0,0,3), yaw -90 deg."""
    if yaw is None:
        yaw = float(np.float32(-90.0) * np.float32(np.pi / 180.0))
    cam = CameraUniform()
    lib.or_camera_uniform(np.array(pos, np.float32), yaw, pitch, aspect, frame, C.byref(cam))
    return cam


class Scene:
    """Oracle scene mirroring SceneBuilder (builder.rs:47-266)."""

    def __init__(self, fp64: bool = False):
        self.lib = lib64() if fp64 else lib
        self.fdt = np.float64 if fp64 else np.float32
        self.h = self.lib.or_scene_create()
        self.built = False

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_scene_destroy(self.h)
            self.h = None

    def add_material(self, color, extra) -> int:
        m = np.zeros(1, MATERIAL_DTYPE)
        m["color"] = color
        m["extra"] = extra
        return self.lib.or_scene_add_material(self.h, _ptr(m))

    def add_mesh(self, verts, idx) -> int:
        verts = np.ascontiguousarray(verts)
        idx = np.ascontiguousarray(idx, np.uint32)
        return self.lib.or_scene_add_mesh(self.h, _ptr(verts), len(verts), idx, len(idx))

    def add_instance(self, mesh, mat, m16):
        self.lib.or_scene_add_instance(self.h, mesh, mat, np.ascontiguousarray(m16, self.fdt))

    def add_light_instance(self, mesh, mat, m16, params):
        self.lib.or_scene_add_light_instance(self.h, mesh, mat, np.ascontiguousarray(m16, self.fdt),
                                        np.asarray(params, self.fdt))

    def set_environment(self, rgb):
        rgb = np.ascontiguousarray(rgb, self.fdt)
        h, w = rgb.shape[:2]
        self.lib.or_scene_set_environment(self.h, rgb, w, h)

    def add_point_light(self, pos, params):
        self.lib.or_scene_add_point_light(self.h, np.asarray(pos, self.fdt),
                                     np.asarray(params, self.fdt))

    def build(self):
        assert self.lib.or_scene_build(self.h) == 0
        self.built = True
        return self

    @property
    def n_triangles(self) -> int:
        return self.lib.or_scene_num_triangles(self.h)

    def lights(self) -> np.ndarray:
        out = np.zeros(self.lib.or_scene_num_lights(self.h), LIGHT_DTYPE)
        self.lib.or_scene_get_lights(self.h, _ptr(out))
        return out

    def world(self):
        n = self.n_triangles
        tris = np.zeros((n, 9), self.fdt)
        ids = np.zeros((n, 3), np.uint32)
        self.lib.or_scene_get_world(self.h, _ptr(tris), _ptr(ids))
        return tris, ids

    def closest_hit(self, rays: np.ndarray, brute: bool = False):
        rays = np.ascontiguousarray(rays, self.fdt).reshape(-1, 8)
        n = len(rays)
        hits = np.zeros((n, 4), np.uint32)
        flags = np.zeros(n, np.uint32)
        self.lib.or_closest_hit(self.h, rays, n, int(brute), hits, flags)
        return hits, flags

    def any_hit(self, rays: np.ndarray, brute: bool = False):
        rays = np.ascontiguousarray(rays, self.fdt).reshape(-1, 8)
        occ = np.zeros(len(rays), np.uint32)
        self.lib.or_any_hit(self.h, rays, len(rays), int(brute), occ)
        return occ

    def morton(self) -> np.ndarray:
        out = np.zeros(self.n_triangles, np.uint32)
        self.lib.or_morton_codes(self.h, out)
        return out

    def sample_light(self, hp, wl, rng):
        """One call of sample_random_light: (delta, direction, distance, pdf,
        radiance, rng after)."""
        d = np.zeros(3, self.fdt)
        rad = np.zeros(4, self.fdt)
        F = C.c_double if self.fdt == np.float64 else C.c_float
        dist, pdf, r = F(), F(), C.c_uint32()
        delta = self.lib.or_sample_light_f(self.h, np.asarray(hp, self.fdt), np.asarray(wl, self.fdt), rng, d,
                                           C.byref(dist), C.byref(pdf), rad, C.byref(r))
        return bool(delta), d, dist.value, pdf.value, rad, r.value

    def photon_emit(self, frame: int, pid: int):
        """Emission of photon path pid: (position, direction, wavelength, energy)."""
        p = np.zeros(3, self.fdt)
        d = np.zeros(3, self.fdt)
        F = C.c_double if self.fdt == np.float64 else C.c_float
        wl, en = F(), F()
        self.lib.or_photon_emit_f(self.h, frame, pid, p, d, C.byref(wl), C.byref(en))
        return p, d, wl.value, en.value

    def photon_trace(self, cfg: Config, frame: int, first: int, n: int) -> np.ndarray:
        out = np.zeros(n, PHOTON_DTYPE)
        self.lib.or_photon_trace(self.h, C.byref(cfg), frame, first, n, _ptr(out))
        return out


def unit_vectors(seed: int, n: int) -> np.ndarray:
    """n directions of random_unit_vector from the stream seeded with pcg_hash(seed)."""
    out = np.zeros((n, 3), np.float32)
    lib.or_unit_vectors_f(seed, n, out)
    return out


def material(color, extra) -> np.ndarray:
    m = np.zeros(1, MATERIAL_DTYPE)
    m["color"] = color
    m["extra"] = extra
    return m


FIX_EMISSION_MIS, FIX_RR_CLAMP, FIX_ONB_SIGN = 1, 2, 4  # wsrt_config.quirks bits


def sample_bsdf(rdir, ff, mat, front, wl, tp, rng, quirks=0):
    """One call of sample_bsdf: (absorbed, direction, attenuation, rng after)."""
    f = np.float32
    d = np.zeros(3, f)
    a = np.zeros(4, f)
    r = C.c_uint32()
    ab = lib.or_sample_bsdf_f(np.asarray(rdir, f), np.asarray(ff, f), _ptr(mat), int(front), np.asarray(wl, f),
                              np.asarray(tp, f), rng, quirks, d, a, C.byref(r))
    return bool(ab), d, a, r.value


def tone_map(accum) -> np.ndarray:
    """RGBA8 of accumulated pixels (n x 4 floats), shader.wgsl:809-822."""
    a = np.ascontiguousarray(accum, np.float32).reshape(-1, 4)
    out = np.zeros((len(a), 4), np.uint8)
    lib.or_tone_map_f(a, len(a), out)
    return out


def roulette(tp, rng, quirks=0):
    """One Russian roulette step: (survived, throughput after, rng after)."""
    out = np.zeros(4, np.float32)
    r = C.c_uint32()
    alive = lib.or_roulette_f(np.asarray(tp, np.float32), rng, quirks, out, C.byref(r))
    return bool(alive), out, r.value


def cornell_box(fp64: bool = False) -> Scene:
    s = Scene(fp64)
    s.lib.or_scene_cornell_box(s.h)
    return s.build()


def bvh8_closest(nodes: np.ndarray, tris: np.ndarray, rays: np.ndarray):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = len(rays)
    hits = np.zeros((n, 4), np.uint32)
    flags = np.zeros(n, np.uint32)
    nv = np.zeros(n, np.uint32)
    nt = np.zeros(n, np.uint32)
    n_nodes = len(nodes) // 80
    lib.or_bvh8_closest(_ptr(nodes), n_nodes, _ptr(tris), rays, n, _ptr(hits), _ptr(flags),
                        _ptr(nv), _ptr(nt))
    return hits, flags, nv, nt


def grid_counting(photons: np.ndarray, frame: int, cell: float, hash_size: int):
    n = len(photons)
    cell_id = np.zeros(n, np.uint32)
    count = np.zeros(hash_size, np.uint32)
    offset = np.zeros(hash_size, np.uint32)
    sorted_slot = np.zeros(max(n, 1), np.uint32)
    valid = lib.or_grid_counting(_ptr(photons), n, frame, cell, hash_size, cell_id, count,
                                 offset, sorted_slot)
    return cell_id, count, offset, sorted_slot[:valid]


def grid_linked(photons: np.ndarray, frame: int, cell: float, hash_size: int):
    n = len(photons)
    head = np.zeros(hash_size, np.uint32)
    nxt = np.zeros(max(n, 1), np.uint32)
    lib.or_grid_linked(_ptr(photons), n, frame, cell, hash_size, head, nxt)
    return head, nxt[:n]


class Frame:
    """One frame of the render loop on the CPU: photon pass, grid, camera paths."""

    def __init__(self, scene: Scene, cfg: Config, cam: CameraUniform, width: int, height: int,
                 photons: np.ndarray | None = None):
        self.scene, self.W, self.H = scene, width, height
        self.lib, self.fdt = scene.lib, scene.fdt
        if photons is None:
            self.h = self.lib.or_frame_create(scene.h, C.byref(cfg), C.byref(cam), width, height)
        else:
            photons = np.ascontiguousarray(photons, PHOTON_DTYPE)
            self.h = self.lib.or_frame_create_with_photons(scene.h, C.byref(cfg), C.byref(cam), width,
                                                      height, _ptr(photons), len(photons))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_frame_destroy(self.h)
            self.h = None

    @property
    def photons_stored(self) -> int:
        return self.lib.or_frame_photons_stored(self.h)

    def trace_pixels(self, pixels):
        px = np.ascontiguousarray(pixels, np.uint32)
        n = len(px)
        wl = np.zeros((n, 4), self.fdt)
        rad = np.zeros((n, 4), self.fdt)
        rgb = np.zeros((n, 3), self.fdt)
        self.lib.or_frame_trace_pixels(self.h, px, n, _ptr(wl), _ptr(rad), _ptr(rgb))
        return wl, rad, rgb

    def render(self, accum: np.ndarray | None = None, want_rgba8: bool = False):
        if accum is None:
            accum = np.zeros((self.H * self.W, 4), self.fdt)
        rgba = np.zeros((self.H * self.W, 4), np.uint8) if want_rgba8 else None
        self.lib.or_frame_render(self.h, accum, _ptr(rgba))
        return accum, rgba

    def ray_color(self, o, d, wl, rng):
        """ray_color of one ray: spectral radiance at the 4 wavelengths."""
        out = np.zeros(4, self.fdt)
        self.lib.or_ray_color_f(self.h, np.asarray(o, self.fdt), np.asarray(d, self.fdt),
                                np.asarray(wl, self.fdt), rng, out)
        return out

    def gather(self, pos, n, wl):
        out = np.zeros(4, self.fdt)
        self.lib.or_frame_gather(self.h, np.asarray(pos, self.fdt), np.asarray(n, self.fdt),
                            np.asarray(wl, self.fdt), out)
        return out


# ---- presented frame: blit pass and photon overlay ---------------------------
def _srgb_oetf8(v: np.ndarray) -> np.ndarray:
    """Linear [0, 1] -> 8-bit sRGB (IEC 61966-2-1 OETF), in float64."""
    v = np.clip(np.asarray(v, np.float64), 0.0, 1.0)
    e = np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(v, 1.0 / 2.4) - 0.055)
    return np.floor(np.clip(e, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def _srgb_eotf(c: np.ndarray) -> np.ndarray:
    """8-bit sRGB -> linear, in float64."""
    e = np.asarray(c, np.float64) / 255.0
    return np.where(e <= 0.04045, e / 12.92, np.power((e + 0.055) / 1.055, 2.4))


def present(image: np.ndarray, win_w: int, win_h: int, photons=None, view=None, proj=None) -> np.ndarray:
    """The presented frame, [win_h, win_w, 4] uint8, rows top to bottom.

    image: the storage texture, [H, W, 4] uint8 with row 0 the image bottom
    (shader.wgsl:721-722). The blit (blit_pass.rs:114-180, blit.wgsl:1-34)
    draws a quad scaled by (sx, sy) -- sx = target/window aspect when the
    window is wider, else sy = window/target, in f32 -- over a blue clear
    (blit_pass.rs:160); a fragment at the pixel centre samples the texel
    floor(uv * size) (wgpu's default sampler: nearest, clamp to edge), where
    uv = (ndc / scale + 1) / 2 and uv = (0, 0) is texture row 0 at the window
    bottom. The surface is sRGB, so written values pass through the OETF.
    With photons (32-byte records, energy 0 = skipped) and the camera's view
    and proj (column-major), each photon is drawn as a one-pixel point
    (debug_photons.wgsl:28-46, debug_photons_pass.rs:84-96): clip = proj *
    view * p, clipped to the view volume, colour (1, 0.6, 0.1) with alpha
    0.3 added to the linear destination and re-encoded, point by point;
    alpha is replaced by the f32 0.3 (unorm8 77)."""
    H, W = image.shape[:2]
    f32 = np.float32
    target, window = f32(W) / f32(H), f32(win_w) / f32(win_h)
    sx, sy = f32(1.0), f32(1.0)
    if window > target:
        sx = target / window
    else:
        sy = window / target
    xs = np.arange(win_w, dtype=np.float32)
    ys = np.arange(win_h, dtype=np.float32)
    ndx = (xs + f32(0.5)) / f32(win_w) * f32(2.0) - f32(1.0)
    ndy = f32(1.0) - (ys + f32(0.5)) / f32(win_h) * f32(2.0)
    inx = (ndx >= -sx) & (ndx < sx)
    iny = (ndy > -sy) & (ndy <= sy)
    u = (ndx / sx + f32(1.0)) * f32(0.5)
    v = (ndy / sy + f32(1.0)) * f32(0.5)
    tx = np.minimum(np.floor(u * f32(W)).clip(0).astype(np.int64), W - 1)
    ty = np.minimum(np.floor(v * f32(H)).clip(0).astype(np.int64), H - 1)
    enc = _srgb_oetf8(np.arange(256) / 255.0)
    out = np.zeros((win_h, win_w, 4), np.uint8)
    out[..., 2] = 255
    out[..., 3] = 255
    tex = image[ty[:, None], tx[None, :]]
    inside = iny[:, None] & inx[None, :]
    out[inside, 0] = enc[tex[inside, 0]]
    out[inside, 1] = enc[tex[inside, 1]]
    out[inside, 2] = enc[tex[inside, 2]]
    out[inside, 3] = tex[inside, 3]
    if photons is None:
        return out
    pos = np.ascontiguousarray(photons["position"], np.float32)
    keep = photons["energy"] != 0
    pos = pos[keep]
    V = np.asarray(view, np.float32).reshape(4, 4)  # V[c, r]: column c
    P = np.asarray(proj, np.float32).reshape(4, 4)
    p4 = [pos[:, 0], pos[:, 1], pos[:, 2], np.ones(len(pos), np.float32)]
    vv = [((V[0, r] * p4[0] + V[1, r] * p4[1]) + V[2, r] * p4[2]) + V[3, r] * p4[3] for r in range(4)]
    cc = [((P[0, r] * vv[0] + P[1, r] * vv[1]) + P[2, r] * vv[2]) + P[3, r] * vv[3] for r in range(4)]
    w = cc[3]
    ok = (w > 0) & (cc[0] >= -w) & (cc[0] <= w) & (cc[1] >= -w) & (cc[1] <= w) & (cc[2] >= 0) & (cc[2] <= w)
    with np.errstate(divide="ignore", invalid="ignore"):
        nx, ny = cc[0] / w, cc[1] / w
        fx = (nx + f32(1.0)) * f32(0.5) * f32(win_w)
        fy = (f32(1.0) - ny) * f32(0.5) * f32(win_h)
    px, py = np.floor(fx[ok]).astype(np.int64), np.floor(fy[ok]).astype(np.int64)
    on = (px >= 0) & (py >= 0) & (px < win_w) & (py < win_h)
    count = np.zeros((win_h, win_w), np.int64)
    np.add.at(count, (py[on], px[on]), 1)
    add = (0.3 * 1.0, 0.3 * 0.6, 0.3 * 0.1)
    steps = [_srgb_oetf8(_srgb_eotf(np.arange(256)) + a) for a in add]
    hit = count > 0
    for ch in range(3):
        c = out[..., ch][hit].astype(np.int64)
        k = np.minimum(count[hit], 256)
        for _ in range(int(k.max(initial=0))):
            c = np.where(k > 0, steps[ch][c], c)
            k = k - 1
        out[..., ch][hit] = c
    out[..., 3][hit] = int(np.floor(float(np.float32(0.3) * np.float32(255.0)) + 0.5))  # f32 alpha 0.3 as unorm8
    return out
