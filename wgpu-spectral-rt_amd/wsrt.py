"""Python binding of libwsrt over its C ABI (include/wsrt.h).

Mirrors the reference's host interface: ``SceneBuilder`` (src/scene/builder.rs),
the ``geometry`` generators (src/scene/geometry.rs), ``create_cornell_box``
(src/scene/cornell_box.rs), ``CameraController.build_uniform``
(src/scene/camera.rs) and ``Renderer`` (src/renderer/mod.rs) with its public
``frame_count``. Every call goes to the HIP library; there is no fallback: a
missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WSRT_LIB: another build of the library (the sanitized host build)
LIB_PATH = os.environ.get("WSRT_LIB", os.path.join(_HERE, "libwsrt.so"))


class WsrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wsrt error {code}: {msg}")
        self.code = code


# ---- ABI structs (byte layouts of include/wsrt.h) ---------------------------
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


# wsrt_config.quirks bits (WSRT_FIX_*): each replaces one behaviour of the
# reference's shader by the textbook estimator
FIX_EMISSION_MIS, FIX_RR_CLAMP, FIX_ONB_SIGN = 1, 2, 4


class AutotuneResult(C.Structure):
    _fields_ = [("n_candidates", C.c_uint32), ("frames_in_flight", C.c_uint32 * 12),
                ("ms_per_frame", C.c_float * 12), ("device_bytes", C.c_uint64 * 12), ("best", C.c_uint32),
                ("n_stage2", C.c_uint32), ("tail_paths", C.c_uint32 * 8), ("trace_rounds", C.c_uint32 * 8),
                ("stage2_ms_per_frame", C.c_float * 8), ("best_tail_paths", C.c_uint32),
                ("best_trace_rounds", C.c_uint32), ("n_stage3", C.c_uint32), ("extend_refill", C.c_uint32 * 8),
                ("ordered_queues", C.c_uint32 * 8), ("stage3_ms_per_frame", C.c_float * 8),
                ("best_extend_refill", C.c_uint32), ("best_ordered_queues", C.c_uint32)]


class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", CameraUniform), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("pad", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_mode", C.c_uint32), ("blocks_per_cu", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("n_triangles", C.c_uint32), ("n_instances", C.c_uint32),
                ("n_lights", C.c_uint32), ("n_materials", C.c_uint32), ("n_nodes", C.c_uint32),
                ("max_depth", C.c_uint32), ("stack_size", C.c_uint32),
                ("node_bytes", C.c_uint32), ("tri_bytes", C.c_uint32), ("build_ms", C.c_float),
                ("bounds", C.c_float * 6)]


class FrameStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("frames", "camera_paths", "extend_rays", "shadow_rays",
                                           "photon_paths", "photon_rays", "photons_stored",
                                           "guide_samples")] + \
        [("trav_nodes", C.c_uint64 * 3), ("trav_tris", C.c_uint64 * 3), ("trav_rays", C.c_uint64 * 3),
         ("kernel_rays", C.c_uint64 * 6), ("kernel_nodes", C.c_uint64 * 6), ("kernel_tris", C.c_uint64 * 6),
         ("kernel_counted_rays", C.c_uint64 * 6), ("extend_wave_nodes", C.c_uint64),
         ("ray_max_nodes", C.c_uint64 * 6)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double),
                ("items", C.c_uint64)]


VERTEX_DTYPE = np.dtype([("pos", np.float32, 4), ("normal", np.float32, 4)])
MATERIAL_DTYPE = np.dtype([("color", np.float32, 4), ("extra", np.float32, 4)])
PHOTON_DTYPE = np.dtype([("position", np.float32, 3), ("wavelength", np.float32),
                         ("direction", np.float32, 3), ("energy", np.float32)])

vp = C.c_void_p
u32 = C.c_uint32
P = C.POINTER

_SIGS = {
    "wsrt_last_error": (C.c_char_p, []),
    "wsrt_abi_version": (C.c_int, []),
    "wsrt_config_default": (C.c_int, [P(Config)]),
    "wsrt_device_count": (C.c_int, []),
    "wsrt_device_create": (C.c_int, [C.c_int, P(vp)]),
    "wsrt_device_destroy": (C.c_int, [vp]),
    "wsrt_device_set_memory_limit": (C.c_int, [vp, C.c_uint64]),
    "wsrt_device_get_memory_used": (C.c_int, [vp, P(C.c_uint64)]),
    "wsrt_geometry_plane": (C.c_int, [vp, vp, P(u32), P(u32)]),
    "wsrt_geometry_cube": (C.c_int, [vp, vp, P(u32), P(u32)]),
    "wsrt_geometry_sphere": (C.c_int, [u32, vp, vp, P(u32), P(u32)]),
    "wsrt_scene_builder_create": (C.c_int, [P(vp)]),
    "wsrt_scene_builder_destroy": (C.c_int, [vp]),
    "wsrt_scene_add_material": (C.c_int, [vp, vp, P(u32)]),
    "wsrt_scene_add_mesh": (C.c_int, [vp, vp, C.c_size_t, vp, C.c_size_t, P(u32)]),
    "wsrt_scene_add_instance": (C.c_int, [vp, u32, u32, vp]),
    "wsrt_scene_add_light_instance": (C.c_int, [vp, u32, u32, vp, vp]),
    "wsrt_scene_add_point_light": (C.c_int, [vp, vp, vp]),
    "wsrt_scene_set_environment": (C.c_int, [vp, vp, u32, u32]),
    "wsrt_scene_build": (C.c_int, [vp, vp, P(vp)]),
    "wsrt_scene_destroy": (C.c_int, [vp]),
    "wsrt_scene_builder_cornell_box": (C.c_int, [P(vp)]),
    "wsrt_scene_get_info": (C.c_int, [vp, P(SceneInfo)]),
    "wsrt_camera_build_uniform": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, u32,
                                            P(CameraUniform)]),
    "wsrt_renderer_create": (C.c_int, [vp, vp, u32, u32, P(Config), P(vp)]),
    "wsrt_renderer_destroy": (C.c_int, [vp]),
    "wsrt_renderer_set_camera": (C.c_int, [vp, P(CameraUniform)]),
    "wsrt_renderer_render": (C.c_int, [vp]),
    "wsrt_renderer_render_frames": (C.c_int, [vp, u32]),
    "wsrt_renderer_get_frame_count": (C.c_int, [vp, P(u32)]),
    "wsrt_renderer_set_frame_count": (C.c_int, [vp, u32]),
    "wsrt_renderer_resize": (C.c_int, [vp, u32, u32]),
    "wsrt_renderer_sync": (C.c_int, [vp]),
    "wsrt_renderer_read_accum": (C.c_int, [vp, vp]),
    "wsrt_renderer_read_rgba8": (C.c_int, [vp, vp]),
    "wsrt_renderer_write_accum": (C.c_int, [vp, vp]),
    "wsrt_renderer_read_screenshot": (C.c_int, [vp, vp]),
    "wsrt_renderer_present": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_int, vp]),
    "wsrt_renderer_get_stats": (C.c_int, [vp, P(FrameStats)]),
    "wsrt_renderer_reset_stats": (C.c_int, [vp]),
    "wsrt_renderer_set_traversal_counting": (C.c_int, [vp, C.c_int]),
    "wsrt_renderer_set_profiling": (C.c_int, [vp, C.c_int]),
    "wsrt_renderer_get_profile": (C.c_int, [vp, vp, u32, P(u32)]),
    "wsrt_renderer_reset_profile": (C.c_int, [vp]),
    "wsrt_debug_read_photons": (C.c_int, [vp, vp, u32, P(u32)]),
    "wsrt_debug_read_grid": (C.c_int, [vp, vp, vp, vp, vp]),
    "wsrt_debug_read_paths": (C.c_int, [vp, vp, vp]),
    "wsrt_debug_trace_paths": (C.c_int, [vp, vp, C.c_uint32, vp, vp]),
    "wsrt_photon_grid_build": (C.c_int, [vp, vp, u32, u32, C.c_float, u32, vp, vp, vp, vp]),
    "wsrt_photon_grid_build_noalloc": (C.c_int, [vp, vp, u32, u32, C.c_float, u32, vp, vp, vp, vp, vp]),
    "wsrt_debug_closest_hit": (C.c_int, [vp, vp, u32, vp, vp]),
    "wsrt_debug_any_hit": (C.c_int, [vp, vp, u32, vp]),
    "wsrt_debug_traversal_counts": (C.c_int, [vp, vp, u32, vp, vp]),
    "wsrt_debug_read_morton": (C.c_int, [vp, vp, vp]),
    "wsrt_debug_read_bvh": (C.c_int, [vp, vp, vp]),
    "wsrt_debug_read_world": (C.c_int, [vp, vp, vp]),
    "wsrt_guide_create": (C.c_int, [vp, u32, u32, P(vp)]),
    "wsrt_guide_destroy": (C.c_int, [vp]),
    "wsrt_guide_get_params": (C.c_int, [vp, vp]),
    "wsrt_guide_set_params": (C.c_int, [vp, vp]),
    "wsrt_guide_forward": (C.c_int, [vp, vp, u32, vp]),
    "wsrt_guide_gradients": (C.c_int, [vp, vp, vp, vp, u32, vp]),
    "wsrt_guide_train_step": (C.c_int, [vp, vp, vp, vp, u32, C.c_float]),
    "wsrt_renderer_get_guide": (C.c_int, [vp, P(vp)]),
    "wsrt_guide_get_state": (C.c_int, [vp, vp, vp, vp, P(u32)]),
    "wsrt_guide_set_state": (C.c_int, [vp, vp, vp, vp, u32]),
    "wsrt_guide_get_average": (C.c_int, [vp, vp]),
    "wsrt_guide_set_average": (C.c_int, [vp, vp]),
    "wsrt_renderer_get_guide_frames_trained": (C.c_int, [vp, P(C.c_uint64)]),
    "wsrt_renderer_set_guide_frames_trained": (C.c_int, [vp, C.c_uint64]),
    "wsrt_autotune": (C.c_int, [vp, vp, u32, u32, vp, P(Config), C.c_uint64, P(AutotuneResult)]),
    "wsrt_batch_create": (C.c_int, [vp, vp, u32, P(Config), P(vp)]),
    "wsrt_batch_destroy": (C.c_int, [vp]),
    "wsrt_batch_set_tuning": (C.c_int, [vp, P(BatchTuning)]),
    "wsrt_batch_autotune": (C.c_int, [vp, P(BatchTuning), P(C.c_float)]),
    "wsrt_batch_cancel": (C.c_int, [vp]),
    "wsrt_batch_set_time_limit": (C.c_int, [vp, C.c_uint64]),
    "wsrt_batch_render": (C.c_int, [vp]),
    "wsrt_batch_sync": (C.c_int, [vp]),
    "wsrt_batch_read_accum": (C.c_int, [vp, u32, vp]),
    "wsrt_batch_read_rgba8": (C.c_int, [vp, u32, vp]),
}

GUIDE_PARAMS = 27952
GUIDE_OUTPUTS = 40

_lib = None


def load(path: str = LIB_PATH):
    """Loads libwsrt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise WsrtError(-4, f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _check(rc: int):
    if rc != 0:
        raise WsrtError(rc, load().wsrt_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def default_config(**kw) -> Config:
    c = Config()
    _check(load().wsrt_config_default(C.byref(c)))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# ---- geometry (src/scene/geometry.rs) ---------------------------------------
def _geometry(fn, *pre):
    nv, ni = u32(), u32()
    _check(fn(*pre, None, None, C.byref(nv), C.byref(ni)))
    v = np.zeros(nv.value, VERTEX_DTYPE)
    i = np.zeros(ni.value, np.uint32)
    _check(fn(*pre, _p(v), _p(i), C.byref(nv), C.byref(ni)))
    return v, i


def create_plane():
    return _geometry(load().wsrt_geometry_plane)


def create_cube():
    return _geometry(load().wsrt_geometry_cube)


def create_sphere(subdivisions: int):
    return _geometry(load().wsrt_geometry_sphere, subdivisions)


# ---- glam-style matrices (column-major) used by scene definitions ------------
def mat4_mul(a, b) -> np.ndarray:
    """glam Mat4 * Mat4 in f32 with glam's mul_vec4 summation order."""
    a = np.asarray(a, np.float32).reshape(4, 4)  # rows = columns of the matrix
    b = np.asarray(b, np.float32).reshape(4, 4)
    out = np.zeros((4, 4), np.float32)
    for c in range(4):
        s = a[0] * b[c][0]
        s = s + a[1] * b[c][1]
        s = s + a[2] * b[c][2]
        s = s + a[3] * b[c][3]
        out[c] = s
    return out.reshape(16)


def mat4_translation(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)
    return m.reshape(16)


def mat4_scale(x, y, z):
    return np.diag(np.array([x, y, z, 1.0], np.float32)).reshape(16)


def mat4_rotation(axis: str, angle: float):
    a = np.float32(angle)
    s, c = np.float32(math.sin(a)), np.float32(math.cos(a))
    m = np.eye(4, dtype=np.float32)
    if axis == "x":
        m[1, 1], m[1, 2], m[2, 1], m[2, 2] = c, s, -s, c
    elif axis == "y":
        m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    else:
        m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, s, -s, c
    return m.reshape(16)


# ---- device ------------------------------------------------------------------
class Device:
    def __init__(self, ordinal: int = 0):
        h = vp()
        _check(load().wsrt_device_create(ordinal, C.byref(h)))
        self.h = h

    def set_memory_limit(self, nbytes: int):
        """Caps the device memory the library's objects hold (0: no cap)."""
        _check(load().wsrt_device_set_memory_limit(self.h, nbytes))

    @property
    def memory_used(self) -> int:
        out = C.c_uint64()
        _check(load().wsrt_device_get_memory_used(self.h, C.byref(out)))
        return out.value

    def close(self):
        """Destroys the device; a no-op while scenes or renderers created on it
        are alive (the library refuses, and they keep this object alive)."""
        if getattr(self, "h", None) and load().wsrt_device_destroy(self.h) == 0:
            self.h = None

    def __del__(self):
        self.close()


def device_count() -> int:
    return load().wsrt_device_count()


# ---- scene (src/scene/builder.rs) ---------------------------------------------
class SceneBuilder:
    def __init__(self, _handle=None):
        if _handle is None:
            _handle = vp()
            _check(load().wsrt_scene_builder_create(C.byref(_handle)))
        self.h = _handle

    def __del__(self):
        if getattr(self, "h", None):
            load().wsrt_scene_builder_destroy(self.h)
            self.h = None

    def add_material(self, color, extra) -> int:
        m = np.zeros(1, MATERIAL_DTYPE)
        m["color"] = color
        m["extra"] = extra
        i = u32()
        _check(load().wsrt_scene_add_material(self.h, _p(m), C.byref(i)))
        return i.value

    def add_mesh(self, vertices, indices) -> int:
        v = np.ascontiguousarray(vertices)
        i = np.ascontiguousarray(indices, np.uint32)
        out = u32()
        _check(load().wsrt_scene_add_mesh(self.h, _p(v), len(v), _p(i), len(i), C.byref(out)))
        return out.value

    def add_instance(self, mesh_id: int, mat_id: int, transform):
        t = np.ascontiguousarray(transform, np.float32)
        _check(load().wsrt_scene_add_instance(self.h, mesh_id, mat_id, _p(t)))

    def add_light_instance(self, mesh_id: int, mat_id: int, transform, params):
        t = np.ascontiguousarray(transform, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        _check(load().wsrt_scene_add_light_instance(self.h, mesh_id, mat_id, _p(t), _p(p)))

    def add_point_light(self, position, params):
        pos = np.ascontiguousarray(position, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        _check(load().wsrt_scene_add_point_light(self.h, _p(pos), _p(p)))

    def set_environment(self, rgb: np.ndarray):
        rgb = np.ascontiguousarray(rgb, np.float32)
        h, w = rgb.shape[:2]
        _check(load().wsrt_scene_set_environment(self.h, _p(rgb), w, h))

    def build(self, device: Device) -> "Scene":
        out = vp()
        _check(load().wsrt_scene_build(self.h, device.h, C.byref(out)))
        return Scene(out, device)


class Scene:
    """SceneResources (builder.rs:30-39): device buffers plus the BVH8."""

    def __init__(self, handle, device: Device):
        self.h = handle
        self.device = device

    def close(self):
        if getattr(self, "h", None) and load().wsrt_scene_destroy(self.h) == 0:
            self.h = None

    def __del__(self):
        self.close()

    @property
    def info(self) -> SceneInfo:
        i = SceneInfo()
        _check(load().wsrt_scene_get_info(self.h, C.byref(i)))
        return i

    def closest_hit(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        hits = np.zeros((n, 4), np.uint32)
        flags = np.zeros(n, np.uint32)
        _check(load().wsrt_debug_closest_hit(self.h, _p(rays), n, _p(hits), _p(flags)))
        return hits, flags

    def any_hit(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(len(rays), np.uint32)
        _check(load().wsrt_debug_any_hit(self.h, _p(rays), len(rays), _p(occ)))
        return occ

    def traversal_counts(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        nodes = np.zeros(n, np.uint32)
        tris = np.zeros(n, np.uint32)
        _check(load().wsrt_debug_traversal_counts(self.h, _p(rays), n, _p(nodes), _p(tris)))
        return nodes, tris

    def morton(self):
        n = self.info.n_triangles
        codes = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        _check(load().wsrt_debug_read_morton(self.h, _p(codes), _p(order)))
        return codes, order

    def bvh(self):
        info = self.info
        nodes = np.zeros(info.n_nodes * info.node_bytes, np.uint8)
        tris = np.zeros(info.n_triangles * info.tri_bytes, np.uint8)
        _check(load().wsrt_debug_read_bvh(self.h, _p(nodes), _p(tris)))
        return nodes, tris

    def world(self):
        n = self.info.n_triangles
        tris = np.zeros((n, 9), np.float32)
        ids = np.zeros((n, 3), np.uint32)
        _check(load().wsrt_debug_read_world(self.h, _p(tris), _p(ids)))
        return tris, ids


def cornell_box_builder() -> SceneBuilder:
    h = vp()
    _check(load().wsrt_scene_builder_cornell_box(C.byref(h)))
    return SceneBuilder(h)


def create_cornell_box(device: Device) -> Scene:
    """cornell_box.rs:6-110"""
    return cornell_box_builder().build(device)


# ---- camera (src/scene/camera.rs) ---------------------------------------------
class CameraController:
    """The controller's state and build_uniform (camera.rs:125-194); input
    handling is not part of the render loop."""

    def __init__(self):
        self.position = (0.0, 0.0, 3.0)
        self.yaw = float(np.float32(-90.0) * np.float32(3.14159265359 / 180.0))
        self.pitch = 0.0

    def build_uniform(self, aspect: float, frame_count: int) -> CameraUniform:
        cam = CameraUniform()
        pos = np.asarray(self.position, np.float32)
        _check(load().wsrt_camera_build_uniform(_p(pos), self.yaw, self.pitch, aspect, frame_count,
                                                C.byref(cam)))
        return cam


# ---- renderer (src/renderer/mod.rs) -----------------------------------------------
class Guide:
    """The path-guiding network (wsrt_guide_*): 7 input scalars in [0, 1] ->
    40 outputs (8 vMF lobes of mu xyz, log kappa, logit)."""

    def __init__(self, device: Device | None = None, max_samples: int = 1 << 16, seed: int = 1,
                 _handle=None, _owner=None):
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            return
        h = vp()
        _check(load().wsrt_guide_create(device.h, max_samples, seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self._owner is None:
            load().wsrt_guide_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(GUIDE_PARAMS, np.float32)
        _check(load().wsrt_guide_get_params(self.h, _p(out)))
        return out

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (GUIDE_PARAMS,)
        _check(load().wsrt_guide_set_params(self.h, _p(w)))

    def state(self) -> dict:
        """Parameters, Adam moments and step, and the running average of the
        parameters the tracer samples with (a training checkpoint)."""
        p, m1, m2, avg = (np.zeros(GUIDE_PARAMS, np.float32) for _ in range(4))
        step = u32()
        _check(load().wsrt_guide_get_state(self.h, _p(p), _p(m1), _p(m2), C.byref(step)))
        _check(load().wsrt_guide_get_average(self.h, _p(avg)))
        return {"params": p, "m1": m1, "m2": m2, "step": step.value, "average": avg}

    def set_state(self, st: dict):
        p, m1, m2 = (np.ascontiguousarray(st[k], np.float32) for k in ("params", "m1", "m2"))
        assert p.size == m1.size == m2.size == GUIDE_PARAMS
        _check(load().wsrt_guide_set_state(self.h, _p(p), _p(m1), _p(m2), int(st["step"])))
        if "average" in st:  # else the average restarts at the parameters
            avg = np.ascontiguousarray(st["average"], np.float32)
            assert avg.size == GUIDE_PARAMS
            _check(load().wsrt_guide_set_average(self.h, _p(avg)))

    @property
    def average(self) -> np.ndarray:
        """The running average of the parameters (the network the tracer samples with)."""
        out = np.zeros(GUIDE_PARAMS, np.float32)
        _check(load().wsrt_guide_get_average(self.h, _p(out)))
        return out

    def forward(self, scalars) -> np.ndarray:
        x = np.ascontiguousarray(scalars, np.float32).reshape(-1, 7)
        out = np.zeros((len(x), GUIDE_OUTPUTS), np.float32)
        _check(load().wsrt_guide_forward(self.h, _p(x), len(x), _p(out)))
        return out

    def gradients(self, scalars, dirs, weights) -> np.ndarray:
        x = np.ascontiguousarray(scalars, np.float32).reshape(-1, 7)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        _check(load().wsrt_guide_gradients(self.h, _p(x), _p(d), _p(w), len(x), _p(g)))
        return g

    def train_step(self, scalars, dirs, weights, lr: float = 1e-3):
        x = np.ascontiguousarray(scalars, np.float32).reshape(-1, 7)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        _check(load().wsrt_guide_train_step(self.h, _p(x), _p(d), _p(w), len(x), lr))


class Renderer:
    def __init__(self, device: Device, scene: Scene, target_width: int, target_height: int,
                 config: Config | None = None):
        h = vp()
        cfg = config if config is not None else default_config()
        _check(load().wsrt_renderer_create(device.h, scene.h, target_width, target_height,
                                           C.byref(cfg), C.byref(h)))
        self.h = h
        self.scene = scene
        self.config = cfg
        self.target_width = target_width
        self.target_height = target_height

    def close(self):
        if getattr(self, "h", None):
            load().wsrt_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def frame_count(self) -> int:
        v = u32()
        _check(load().wsrt_renderer_get_frame_count(self.h, C.byref(v)))
        return v.value

    @frame_count.setter
    def frame_count(self, value: int):
        _check(load().wsrt_renderer_set_frame_count(self.h, value))

    def set_camera(self, cam: CameraUniform):
        _check(load().wsrt_renderer_set_camera(self.h, C.byref(cam)))

    @property
    def guide_frames_trained(self) -> int:
        out = C.c_uint64()
        _check(load().wsrt_renderer_get_guide_frames_trained(self.h, C.byref(out)))
        return out.value

    @guide_frames_trained.setter
    def guide_frames_trained(self, n: int):
        _check(load().wsrt_renderer_set_guide_frames_trained(self.h, n))

    def guide(self) -> Guide:
        """The guiding network this renderer trains (config.guiding = 1)."""
        h = vp()
        _check(load().wsrt_renderer_get_guide(self.h, C.byref(h)))
        return Guide(_handle=h, _owner=self)

    def render(self):
        _check(load().wsrt_renderer_render(self.h))

    def render_frames(self, n: int):
        _check(load().wsrt_renderer_render_frames(self.h, n))

    def resize(self, w: int, h: int):
        _check(load().wsrt_renderer_resize(self.h, w, h))

    def sync(self):
        _check(load().wsrt_renderer_sync(self.h))

    def read_accum(self) -> np.ndarray:
        out = np.zeros((self.target_height * self.target_width, 4), np.float32)
        _check(load().wsrt_renderer_read_accum(self.h, _p(out)))
        return out

    def write_accum(self, acc: np.ndarray):
        acc = np.ascontiguousarray(acc, np.float32)
        _check(load().wsrt_renderer_write_accum(self.h, _p(acc)))

    def present(self, window_width: int, window_height: int, overlay_photons: bool = False) -> np.ndarray:
        """[window_height, window_width, 4] uint8: the presented frame (the
        blit pass into an sRGB window surface, blit_pass.rs / blit.wgsl), with
        the photon overlay of debug_photons.wgsl drawn over it on request."""
        out = np.zeros((window_height, window_width, 4), np.uint8)
        _check(load().wsrt_renderer_present(self.h, window_width, window_height, int(overlay_photons), _p(out)))
        return out

    def screenshot(self) -> np.ndarray:
        """[H, W, 4] uint8: the screenshot image (screenshot.rs:47-73)."""
        out = np.zeros((self.target_height, self.target_width, 4), np.uint8)
        _check(load().wsrt_renderer_read_screenshot(self.h, _p(out)))
        return out

    def save_png(self, path: str):
        """Writes the screenshot image as an RGBA8 PNG."""
        save_png(path, self.screenshot())

    def read_rgba8(self) -> np.ndarray:
        out = np.zeros((self.target_height * self.target_width, 4), np.uint8)
        _check(load().wsrt_renderer_read_rgba8(self.h, _p(out)))
        return out

    def stats(self) -> dict:
        s = FrameStats()
        _check(load().wsrt_renderer_get_stats(self.h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in FrameStats._fields_[:8]}
        for k in ("trav_nodes", "trav_tris", "trav_rays", "kernel_rays", "kernel_nodes", "kernel_tris",
                  "kernel_counted_rays", "ray_max_nodes"):
            out[k] = list(getattr(s, k))
        out["extend_wave_nodes"] = s.extend_wave_nodes
        return out

    def set_traversal_counting(self, on: bool):
        _check(load().wsrt_renderer_set_traversal_counting(self.h, int(on)))

    def reset_stats(self):
        _check(load().wsrt_renderer_reset_stats(self.h))

    def set_profiling(self, on: bool):
        _check(load().wsrt_renderer_set_profiling(self.h, int(on)))

    def profile(self) -> dict:
        n = u32()
        _check(load().wsrt_renderer_get_profile(self.h, None, 0, C.byref(n)))
        arr = (KernelStat * n.value)()
        _check(load().wsrt_renderer_get_profile(self.h, C.cast(arr, vp), n.value, C.byref(n)))
        return {a.name.decode(): {"launches": a.launches, "total_ms": a.total_ms} for a in arr}

    def reset_profile(self):
        _check(load().wsrt_renderer_reset_profile(self.h))

    def read_photons(self):
        n = self.config.photons_per_pass if self.config.enable_photons else 0
        out = np.zeros(max(n, 1), PHOTON_DTYPE)
        stored = u32()
        _check(load().wsrt_debug_read_photons(self.h, _p(out), len(out), C.byref(stored)))
        return out[:n], stored.value

    def read_grid(self):
        n = self.config.photons_per_pass if self.config.enable_photons else 0
        hs = self.config.hash_size
        cell = np.zeros(max(n, 1), np.uint32)
        count = np.zeros(hs, np.uint32)
        offset = np.zeros(hs, np.uint32)
        slot = np.zeros(max(n, 1), np.uint32)
        _check(load().wsrt_debug_read_grid(self.h, _p(cell), _p(count), _p(offset), _p(slot)))
        return cell[:n], count, offset, slot[:n]

    def read_paths(self):
        npx = self.target_width * self.target_height
        wl = np.zeros((npx, 4), np.float32)
        L = np.zeros((npx, 4), np.float32)
        _check(load().wsrt_debug_read_paths(self.h, _p(wl), _p(L)))
        return wl, L

    def trace_paths(self, pixel_ids):
        """Hero wavelengths and spectral radiance (n x 4 each) of the chosen
        pixels' paths in the last frame."""
        ids = np.ascontiguousarray(pixel_ids, np.uint32).reshape(-1)
        wl = np.zeros((max(len(ids), 1), 4), np.float32)
        L = np.zeros((max(len(ids), 1), 4), np.float32)
        _check(load().wsrt_debug_trace_paths(self.h, _p(ids), len(ids), _p(wl), _p(L)))
        return wl[: len(ids)], L[: len(ids)]


def photon_grid_build(device: Device, photons: np.ndarray, frame_count: int, cell_size: float,
                      hash_size: int):
    """build_grid pass on a fixed photon buffer; returns cell_id, count,
    offset and the slots of the valid photons in cell order."""
    ph = np.ascontiguousarray(photons, PHOTON_DTYPE)
    n = len(ph)
    cell = np.zeros(max(n, 1), np.uint32)
    count = np.zeros(hash_size, np.uint32)
    offset = np.zeros(hash_size, np.uint32)
    slot = np.zeros(max(n, 1), np.uint32)
    _check(load().wsrt_photon_grid_build(device.h, _p(ph), n, frame_count, cell_size, hash_size,
                                         _p(cell), _p(count), _p(offset), _p(slot)))
    valid = int(count.sum())
    return cell[:n], count, offset, slot[:valid]


def save_png(path: str, rgba: np.ndarray):
    """RGBA8 PNG writer (zlib-compressed scanlines, filter 0), [H, W, 4] uint8."""
    import struct
    import zlib

    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w = rgba.shape[:2]

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(raw, 1)) + chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)


def load_png(path: str) -> np.ndarray:
    """Reads back an 8-bit RGBA PNG with filter type 0 (what save_png writes)."""
    import struct
    import zlib

    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        assert zlib.crc32(tag + body) & 0xFFFFFFFF == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        if tag == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4).copy()


def autotune(device: "Device", scene: "Scene", width: int, height: int, config: Config,
             camera: CameraUniform | None = None, mem_budget: int = 0) -> dict:
    """Picks config.frames_in_flight, then tail_paths and trace_rounds, then
    extend_refill and ordered_queues, by timing candidates (wsrt_autotune) for the given camera; updates config in
    place and returns the timings."""
    res = AutotuneResult()
    cam = C.cast(C.byref(camera), vp) if camera is not None else None
    _check(load().wsrt_autotune(device.h, scene.h, width, height, cam, C.byref(config), mem_budget,
                                C.byref(res)))
    n = res.n_candidates
    return {"best": res.best,
            "candidates": [{"frames_in_flight": res.frames_in_flight[i],
                            "ms_per_frame": round(res.ms_per_frame[i], 4),
                            "device_mib": round(res.device_bytes[i] / 2**20, 1)} for i in range(n)],
            "best_tail_paths": res.best_tail_paths, "best_trace_rounds": res.best_trace_rounds,
            "stage2": [{"tail_paths": res.tail_paths[i], "trace_rounds": res.trace_rounds[i],
                        "ms_per_frame": round(res.stage2_ms_per_frame[i], 4)} for i in range(res.n_stage2)],
            "best_extend_refill": res.best_extend_refill, "best_ordered_queues": res.best_ordered_queues,
            "stage3": [{"extend_refill": res.extend_refill[i], "ordered_queues": res.ordered_queues[i],
                        "ms_per_frame": round(res.stage3_ms_per_frame[i], 4)} for i in range(res.n_stage3)]}


class Batch:
    """Jobs (scene, camera, width, height, spp) rendered by one persistent
    launch (wsrt_batch_*); photon pass and guiding off."""

    def __init__(self, device: "Device", jobs, config: Config | None = None):
        arr = (BatchJob * len(jobs))()
        self.sizes = []
        self._scenes = []
        for k, (scene, cam, w, h, spp) in enumerate(jobs):
            arr[k].scene = scene.h
            arr[k].camera = cam
            arr[k].width, arr[k].height, arr[k].spp = w, h, spp
            self.sizes.append((w, h))
            self._scenes.append(scene)
        cfg = config if config is not None else default_config(enable_photons=0)
        h = vp()
        _check(load().wsrt_batch_create(device.h, C.cast(arr, vp), len(jobs), C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().wsrt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_tuning(self, tile_mode: int, blocks_per_cu: int):
        t = BatchTuning(tile_mode, blocks_per_cu)
        _check(load().wsrt_batch_set_tuning(self.h, C.byref(t)))

    def cancel(self):
        """Stops the render in flight; its sync then raises WsrtError (-6)."""
        _check(load().wsrt_batch_cancel(self.h))

    def set_time_limit(self, microseconds: int):
        """Wall-clock budget of each render (0: none)."""
        _check(load().wsrt_batch_set_time_limit(self.h, microseconds))

    def autotune(self) -> dict:
        t, ms = BatchTuning(), C.c_float()
        _check(load().wsrt_batch_autotune(self.h, C.byref(t), C.byref(ms)))
        return {"tile_mode": t.tile_mode, "blocks_per_cu": t.blocks_per_cu, "one_spp_ms": ms.value}

    def render(self):
        _check(load().wsrt_batch_render(self.h))

    def sync(self):
        _check(load().wsrt_batch_sync(self.h))

    def read_accum(self, job: int) -> np.ndarray:
        w, h = self.sizes[job]
        out = np.zeros((w * h, 4), np.float32)
        _check(load().wsrt_batch_read_accum(self.h, job, _p(out)))
        return out

    def read_rgba8(self, job: int) -> np.ndarray:
        w, h = self.sizes[job]
        out = np.zeros((w * h, 4), np.uint8)
        _check(load().wsrt_batch_read_rgba8(self.h, job, _p(out)))
        return out
