"""Synthetic scenes of the benchmark configurations, written against the
reference's SceneBuilder API (add_material / add_mesh / add_instance /
add_light_instance, src/scene/builder.rs:70-129) so the same definition can
drive any builder with that interface.

  cornell       the reference's default scene (src/scene/cornell_box.rs)
  c1_box        34-triangle Cornell-style box: 5 wall quads and two 12-triangle
                boxes rotated about Y by seeded angles in [0, 45] deg, one of
                them GGX (alpha 0.3); a 0.25 x 0.25 blackbody 6500 K area light
  c2_caustic    glass icosphere (subdivision 4, radius 0.5, Cauchy A = 1.5,
                B = 0.0042 um^2) on a 2-triangle diffuse plane, point light
                2 units above the sphere centre

Materials use the reference's layout (builder.rs:5-10): extra.x = type
(0 Lambert, 1 GGX, 2 dielectric, 3 emitter, 4 thin film).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

import wsrt


@dataclass
class View:
    position: tuple
    yaw: float
    pitch: float
    width: int
    height: int


def _seeded_angles(seed: int, n: int) -> list[float]:
    """Rotation angles in [0, 45] deg from the reference's PCG hash."""
    out, s = [], seed
    for _ in range(n):
        s = _pcg(s)
        out.append((s / 4294967296.0) * math.radians(45.0))
    return out


def _pcg(x: int) -> int:
    s = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def _trs(t, axis=None, angle=0.0, s=(1.0, 1.0, 1.0)):
    m = wsrt.mat4_translation(*t)
    if axis is not None:
        m = wsrt.mat4_mul(m, wsrt.mat4_rotation(axis, angle))
    return wsrt.mat4_mul(m, wsrt.mat4_scale(*s))


PI = float(np.float32(math.pi))
HALF_PI = float(np.float32(math.pi / 2))


def _walls(b, white, red, green):
    plane = b.add_mesh(*wsrt.create_plane())
    b.add_instance(plane, white, _trs((0, -1, 0), s=(2, 2, 2)))
    b.add_instance(plane, white, _trs((0, 1, 0), "x", PI, (2, 2, 2)))
    b.add_instance(plane, white, _trs((0, 0, -1), "x", HALF_PI, (2, 2, 2)))
    b.add_instance(plane, red, _trs((-1, 0, 0), "z", -HALF_PI, (2, 2, 2)))
    b.add_instance(plane, green, _trs((1, 0, 0), "z", HALF_PI, (2, 2, 2)))
    return plane


def build_c1_box(b, seed: int = 0):
    """34 triangles + a 2-triangle area light (config C1)."""
    light = b.add_material([6500.0, 1.0, 0.0, 1.0], [3, 0, 0, 0])
    red = b.add_material([0.65, 0.05, 0.05, 1.0], [0, 0, 0, 0])
    green = b.add_material([0.12, 0.45, 0.15, 1.0], [0, 0, 0, 0])
    white = b.add_material([0.73, 0.73, 0.73, 1.0], [0, 0, 0, 0])
    metal = b.add_material([0.9, 0.8, 0.6, 1.0], [1, 0.3, 0, 0])
    plane = _walls(b, white, red, green)
    cube = b.add_mesh(*wsrt.create_cube())
    a0, a1 = _seeded_angles(seed, 2)
    b.add_instance(cube, white, _trs((-0.35, -0.4, -0.3), "y", a0, (0.6, 1.2, 0.6)))
    b.add_instance(cube, metal, _trs((0.35, -0.7, 0.3), "y", a1, (0.6, 0.6, 0.6)))
    b.add_light_instance(plane, light, _trs((0, 0.99, 0), "x", PI, (0.25, 0.25, 0.25)),
                         [0.0, 6500.0, 1.0, 0.0])
    return b


def build_c2_caustic(b):
    """Glass icosphere over a diffuse plane under a point light (config C2)."""
    floor = b.add_material([0.8, 0.8, 0.8, 1.0], [0, 0, 0, 0])
    glass = b.add_material([1.0, 1.0, 1.0, 1.0], [2, 0, 1.5, 0.0042])
    plane = b.add_mesh(*wsrt.create_plane())
    sphere = b.add_mesh(*wsrt.create_sphere(4))
    b.add_instance(plane, floor, _trs((0, 0, 0), s=(4, 4, 4)))
    b.add_instance(sphere, glass, _trs((0, 0.5, 0)))
    b.add_point_light([0.0, 2.5, 0.0], [1.0, 6500.0, 5.0, 0.0])
    return b


C2_VIEW = View((0.0, 2.0, 3.0), -HALF_PI, -math.atan2(1.5, 3.0), 512, 512)
C1_VIEW = View((0.0, 0.0, 3.0), -HALF_PI, 0.0, 256, 256)


def c2_config(**kw):
    """2^20 photons per pass, 2^22 cells, cell size = gather radius = 0.01."""
    c = wsrt.default_config(cell_size=0.01, gather_radius=0.01, gather_radius_exponent=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def c1_config(**kw):
    c = wsrt.default_config(max_depth=8)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def camera(view: View, frame: int = 0, builder=None):
    """The camera uniform of a view (camera.rs:178-194)."""
    ctl = wsrt.CameraController()
    ctl.position, ctl.yaw, ctl.pitch = view.position, view.yaw, view.pitch
    return ctl.build_uniform(view.width / view.height, frame)
