"""Synthetic scenes of the benchmark configurations, written against the
reference's SceneBuilder API (add_material / add_mesh / add_instance /
add_light_instance, src/scene/builder.rs:70-129) so the same definition can
drive any builder with that interface.

  cornell       the reference's default scene (src/scene/cornell_box.rs)
  small_box     34-triangle Cornell-style box: 5 wall quads and two 12-triangle
                boxes rotated about Y by seeded angles in [0, 45] deg, one of
                them GGX (alpha 0.3); a 0.25 x 0.25 blackbody 6500 K area light
  caustic       glass icosphere (subdivision 4, radius 0.5, Cauchy A = 1.5,
                B = 0.0042 um^2) on a 2-triangle diffuse plane, point light
                2 units above the sphere centre
  heightfield   1024 x 1024-vertex heightfield (2,093,058 triangles) displaced
                by seeded 6-octave Perlin noise of amplitude 0.1, 16 x 16
                tiles of thin-film-coated GGX (alpha U[0.05, 0.6], film
                200..800 nm of index 1.33 on a 1.5 substrate), lit by a
                seeded 512 x 256 sky-gradient environment
  custom        the reference's second scene (src/scene/custom.rs): a dark
                room, a vertical slit light, a high-dispersion glass prism
                (Cauchy B = 0.12), a thin-film bubble and a white board
  batch_scene   one of the procedurally generated scenes of the batch
                workload: seeded displaced icospheres and boxes on a floor
                under an area light, triangle count log-uniform in
                [1e3, 2e6], materials drawn from Lambert, GGX, dielectric
                and thin film
  occluded_room two rooms split by a wall with a 0.05-wide full-height door
                gap; the only light is on the ceiling of the far room, so the
                near room is lit through the gap; both rooms are cluttered
                with seeded boxes (174,762 boxes for the 2M-triangle room)

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


def build_small_box(b, seed: int = 0):
    """34 triangles + a 2-triangle area light."""
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


def build_caustic(b):
    """Glass icosphere over a diffuse plane under a point light."""
    floor = b.add_material([0.8, 0.8, 0.8, 1.0], [0, 0, 0, 0])
    glass = b.add_material([1.0, 1.0, 1.0, 1.0], [2, 0, 1.5, 0.0042])
    plane = b.add_mesh(*wsrt.create_plane())
    sphere = b.add_mesh(*wsrt.create_sphere(4))
    b.add_instance(plane, floor, _trs((0, 0, 0), s=(4, 4, 4)))
    b.add_instance(sphere, glass, _trs((0, 0.5, 0)))
    b.add_point_light([0.0, 2.5, 0.0], [1.0, 6500.0, 5.0, 0.0])
    return b


CAUSTIC_VIEW = View((0.0, 2.0, 3.0), -HALF_PI, -math.atan2(1.5, 3.0), 512, 512)
SMALL_BOX_VIEW = View((0.0, 0.0, 3.0), -HALF_PI, 0.0, 256, 256)


def caustic_config(**kw):
    """2^20 photons per pass, 2^22 cells, cell size = gather radius = 0.01;
    closest-hit rays on the lane-refill kernel (refill at 24 idle lanes: at
    256 frames in flight 0.1064 ms per frame against 0.1097 at 16, 0.1074 at
    32 and 0.1129 without refill)."""
    c = wsrt.default_config(cell_size=0.01, gather_radius=0.01, gather_radius_exponent=0.0, extend_refill=24)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def small_box_config(**kw):
    c = wsrt.default_config(max_depth=8)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def camera(view: View, frame: int = 0, builder=None):
    """The camera uniform of a view (camera.rs:178-194)."""
    ctl = wsrt.CameraController()
    ctl.position, ctl.yaw, ctl.pitch = view.position, view.yaw, view.pitch
    return ctl.build_uniform(view.width / view.height, frame)


def _clutter_mesh(n_boxes: int, seed: int, region):
    """n_boxes cubes merged into one mesh: edge lengths uniform in
    [0.02, 0.15] (height up to 0.3), rotated about Y by a uniform angle,
    standing on the floor at uniform positions inside region
    ((x0, x1), (z0, z1)). Seeded numpy PCG64."""
    cv, ci = wsrt.create_cube()
    rng = np.random.default_rng(seed)
    sx = rng.uniform(0.02, 0.15, n_boxes)
    sy = rng.uniform(0.02, 0.3, n_boxes)
    sz = rng.uniform(0.02, 0.15, n_boxes)
    ang = rng.uniform(0.0, 2.0 * math.pi, n_boxes)
    (x0, x1), (z0, z1) = region
    tx = rng.uniform(x0, x1, n_boxes)
    tz = rng.uniform(z0, z1, n_boxes)
    c, sn = np.cos(ang), np.sin(ang)
    p = cv["pos"][:, :3].astype(np.float64)
    nrm = cv["normal"][:, :3].astype(np.float64)
    # scale, rotate about Y, translate (box bottom on y = 0)
    px = p[None, :, 0] * sx[:, None]
    py = (p[None, :, 1] + 0.5) * sy[:, None]
    pz = p[None, :, 2] * sz[:, None]
    wx = c[:, None] * px + sn[:, None] * pz + tx[:, None]
    wz = -sn[:, None] * px + c[:, None] * pz + tz[:, None]
    nx = c[:, None] * nrm[None, :, 0] + sn[:, None] * nrm[None, :, 2]
    nz = -sn[:, None] * nrm[None, :, 0] + c[:, None] * nrm[None, :, 2]
    ny = np.broadcast_to(nrm[None, :, 1], nx.shape)
    v = np.zeros(n_boxes * len(cv), wsrt.VERTEX_DTYPE)
    v["pos"][:, 0] = wx.reshape(-1)
    v["pos"][:, 1] = py.reshape(-1)
    v["pos"][:, 2] = wz.reshape(-1)
    v["pos"][:, 3] = 1.0
    v["normal"][:, 0] = nx.reshape(-1)
    v["normal"][:, 1] = ny.reshape(-1)
    v["normal"][:, 2] = nz.reshape(-1)
    idx = (ci[None, :].astype(np.int64) + (np.arange(n_boxes) * len(cv))[:, None]).astype(np.uint32)
    return v, idx.reshape(-1)


OCCLUDED_BOXES = 174762  # x 12 triangles + 24 room triangles = 2,097,168 triangles


def build_occluded_room(b, n_boxes: int = OCCLUDED_BOXES, seed: int = 5, light_kind: str = "ceiling"):
    """Near room z in [0, 2], far room z in [-2, 0], x in [-2, 2], y in [0, 2].
    The partition at z = 0 is two 0.05-thick slabs leaving a door gap
    x in [-0.025, 0.025] over the full height. The far room's ceiling is a
    3.8 x 1.8 blackbody (6500 K, intensity 1) panel facing down, so the near
    room is lit only through the gap."""
    white = b.add_material([0.75, 0.75, 0.75, 1.0], [0, 0, 0, 0])
    warm = b.add_material([0.7, 0.55, 0.4, 1.0], [0, 0, 0, 0])
    cool = b.add_material([0.35, 0.45, 0.6, 1.0], [0, 0, 0, 0])
    light = b.add_material([6500.0, 1.0, 0.0, 1.0], [3, 0, 0, 0])
    plane = b.add_mesh(*wsrt.create_plane())
    cube = b.add_mesh(*wsrt.create_cube())
    b.add_instance(plane, white, _trs((0, 0, 0), s=(4, 1, 4)))                 # floor
    b.add_instance(plane, white, _trs((0, 2, 0), "x", PI, (4, 1, 4)))          # ceiling
    b.add_instance(plane, white, _trs((0, 1, -2), "x", HALF_PI, (4, 1, 2)))    # far wall
    b.add_instance(plane, white, _trs((0, 1, 2), "x", -HALF_PI, (4, 1, 2)))    # near wall
    b.add_instance(plane, warm, _trs((-2, 1, 0), "z", -HALF_PI, (2, 1, 4)))    # left
    b.add_instance(plane, cool, _trs((2, 1, 0), "z", HALF_PI, (2, 1, 4)))      # right
    half = (2.0 - 0.025) / 2.0
    b.add_instance(cube, white, _trs((-0.025 - half, 1.0, 0.0), s=(2 * half, 2.0, 0.05)))
    b.add_instance(cube, white, _trs((0.025 + half, 1.0, 0.0), s=(2 * half, 2.0, 0.05)))
    if light_kind == "ceiling":
        b.add_light_instance(plane, light, _trs((0, 1.99, -1.0), "x", PI, (3.8, 1, 1.8)),
                             [0.0, 6500.0, 1.0, 0.0])
    else:  # a 0.5 x 0.5 panel on the far wall facing the gap
        lamp = b.add_material([6500.0, 8.0, 0.0, 1.0], [3, 0, 0, 0])
        b.add_light_instance(plane, lamp, _trs((0, 1.0, -1.98), "x", HALF_PI, (0.5, 1, 0.5)),
                             [0.0, 6500.0, 8.0, 0.0])
    if n_boxes:
        nf = n_boxes // 2
        va, ia = _clutter_mesh(nf, seed, ((-1.9, 1.9), (-1.9, -0.15)))
        vb, ib = _clutter_mesh(n_boxes - nf, seed + 1, ((-1.9, 1.9), (0.15, 1.9)))
        b.add_instance(b.add_mesh(va, ia), warm, _trs((0, 0, 0)))
        b.add_instance(b.add_mesh(vb, ib), white, _trs((0, 0, 0)))
    return b


OCCLUDED_VIEW = View((0.0, 1.0, 1.9), -HALF_PI, -0.6, 1024, 1024)


def occluded_config(**kw):
    """Plain tracer settings of the occluded-light room (no photon pass, depth 8).
    The closest-hit kernel refills a wave's lanes once 16 idle (ray lengths in
    the room differ widely: a few rays through the door gap, most stopped by
    the walls or the clutter) and rays are queued in path order; measured on
    a 1024^2 sample at 128 frames in flight: 8.70 ms without either, 6.54
    with the refill, 6.14 with both."""
    c = wsrt.default_config(max_depth=8, enable_photons=0, extend_refill=16, ordered_queues=1)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def guided_config(**kw):
    """The guided tracer: 4 x 64 MLP, 8 vMF lobes, Adam lr 1e-3, 2^18 training
    samples per frame, 50 % one-sample MIS with BSDF sampling."""
    c = occluded_config(guiding=1, guide_mix=0.5, guide_lr=1e-3, guide_train_samples=1 << 18)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def build_custom(b):
    """create_custom_scene (src/scene/custom.rs:6-116) through the builder API:
    same materials, meshes and instance transforms (translation * rotation *
    scale, glam order)."""
    light = b.add_material([6500.0, 5.0, 0.0, 1.0], [3, 0, 0, 0])
    wall = b.add_material([0.01, 0.01, 0.01, 1.0], [0, 0, 0, 0])
    board = b.add_material([0.95, 0.95, 0.95, 1.0], [0, 0, 0, 0])
    prism = b.add_material([1.0, 1.0, 1.0, 1.0], [2, 0.0, 1.6, 0.12])
    bubble = b.add_material([1.0, 1.0, 1.0, 1.0], [4, 500.0, 1.33, 1.0])
    plane = b.add_mesh(*wsrt.create_plane())
    cube = b.add_mesh(*wsrt.create_cube())
    sphere = b.add_mesh(*wsrt.create_sphere(4))
    room = 4.0
    b.add_instance(plane, wall, _trs((0, -1, 0), s=(room * 2, room * 2, room * 2)))
    b.add_instance(plane, wall, _trs((0, room - 1.0, 0), "x", PI, (room * 2, room * 2, room * 2)))
    b.add_instance(plane, wall, _trs((0, room * 0.5 - 1.0, -room), "x", HALF_PI, (room * 2, room * 2, room * 2)))
    b.add_instance(cube, board, _trs((0.5, -0.55, 0.0), s=(5.0, 0.1, 3.0)))
    tilt = float(np.float32(-np.float32(math.pi / 2) + np.float32(0.02)))
    b.add_light_instance(plane, light, _trs((-2.0, -0.2, 0.0), "z", tilt, (0.5, 1.0, 0.02)),
                         [0.0, 6500.0, 5.0, 0.0])
    b.add_instance(cube, prism, _trs((-1.1, -0.2, 0.0), "y", 0.7, (0.5, 0.8, 0.5)))
    b.add_instance(sphere, bubble, _trs((1.0, -0.1, -0.5), s=(0.6, 0.6, 0.6)))
    return b


DEFAULT_VIEW = View((0.0, 0.0, 3.0), -HALF_PI, 0.0, 256, 256)  # CameraController::new (camera.rs:35-52)


def _perlin_grads(rng):
    a = rng.uniform(0.0, 2.0 * math.pi, 256)
    return np.stack([np.cos(a), np.sin(a)], 1), rng.permutation(256)


def perlin_fbm(x, z, seed: int, octaves: int = 6) -> np.ndarray:
    """Seeded 2D gradient (Perlin) noise, fractal sum of `octaves` octaves with
    lacunarity 2 and gain 1/2, normalised to [-1, 1] by the sum of gains."""
    rng = np.random.default_rng(seed)
    out = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, 4.0, 0.0
    for _ in range(octaves):
        g, perm = _perlin_grads(rng)
        px, pz = x * freq, z * freq
        x0, z0 = np.floor(px).astype(np.int64), np.floor(pz).astype(np.int64)
        fx, fz = px - x0, pz - z0

        def grad(ix, iz, dx, dz):
            h = perm[(perm[ix & 255] + iz) & 255]
            return g[h, 0] * dx + g[h, 1] * dz

        def fade(t):
            return t * t * t * (t * (t * 6 - 15) + 10)

        n00 = grad(x0, z0, fx, fz)
        n10 = grad(x0 + 1, z0, fx - 1, fz)
        n01 = grad(x0, z0 + 1, fx, fz - 1)
        n11 = grad(x0 + 1, z0 + 1, fx - 1, fz - 1)
        u, v = fade(fx), fade(fz)
        nx0 = n00 + u * (n10 - n00)
        nx1 = n01 + u * (n11 - n01)
        out += amp * (nx0 + v * (nx1 - nx0)) * math.sqrt(2.0)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


def sky_environment(seed: int = 7, w: int = 512, h: int = 256) -> np.ndarray:
    """Seeded sky gradient, [h, w, 3] float32 RGB radiance, row 0 at +Y: a
    zenith-to-horizon blue-to-white ramp, a warm sun disc at a seeded
    azimuth and elevation, and a dim ground below the horizon."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h                     # 0 at +Y
    u = (np.arange(w) + 0.5) / w
    theta = v * math.pi
    y = np.cos(theta)[:, None]
    phi = (u - 0.5) * 2.0 * math.pi
    zen = np.array(rng.uniform([0.15, 0.3, 0.7], [0.25, 0.45, 0.9]))
    hor = np.array([0.9, 0.9, 0.95])
    t = np.clip(y, 0.0, 1.0) ** 0.5
    sky = hor[None, None] * (1 - t[..., None]) + zen[None, None] * t[..., None]
    ground = np.array([0.08, 0.07, 0.06])
    img = np.where((y > 0)[..., None], sky, ground[None, None]) * np.ones((h, w, 1))
    sun_phi, sun_el = rng.uniform(-math.pi, math.pi), rng.uniform(0.3, 0.8)
    sd = np.array([math.cos(sun_el) * math.cos(sun_phi), math.sin(sun_el), math.cos(sun_el) * math.sin(sun_phi)])
    dx = np.sin(theta)[:, None] * np.cos(phi)[None, :]
    dz = np.sin(theta)[:, None] * np.sin(phi)[None, :]
    dy = np.cos(theta)[:, None] * np.ones((1, w))
    cosang = dx * sd[0] + dy * sd[1] + dz * sd[2]
    img = img + (cosang > math.cos(0.05))[..., None] * np.array([60.0, 50.0, 35.0])[None, None]
    return img.astype(np.float32)


HEIGHTFIELD_N = 1024  # vertices per side


def build_heightfield(b, n: int = HEIGHTFIELD_N, tiles: int = 16, seed: int = 3, sky_seed: int = 7):
    """n x n vertices over [-1, 1]^2 in x, z, y = 0.1 * fbm(x, z); split into
    tiles x tiles meshes (shared border vertices duplicated), each with its
    own thin-film-coated GGX material (type 5)."""
    rng = np.random.default_rng(seed + 1000)
    xs = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(xs, xs, indexing="xy")  # [row = z][col = x]
    Y = 0.1 * perlin_fbm(X, Z, seed)
    # normals from central differences of the height
    dydx = np.gradient(Y, xs, axis=1)
    dydz = np.gradient(Y, xs, axis=0)
    N = np.stack([-dydx, np.ones_like(Y), -dydz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    q = n - 1  # quads per side
    edges = np.linspace(0, q, tiles + 1).round().astype(int)
    for ti in range(tiles):
        for tj in range(tiles):
            r0, r1, c0, c1 = edges[ti], edges[ti + 1], edges[tj], edges[tj + 1]
            vr, vc = r1 - r0 + 1, c1 - c0 + 1
            v = np.zeros(vr * vc, wsrt.VERTEX_DTYPE)
            sl = (slice(r0, r1 + 1), slice(c0, c1 + 1))
            v["pos"][:, 0] = X[sl].reshape(-1)
            v["pos"][:, 1] = Y[sl].reshape(-1)
            v["pos"][:, 2] = Z[sl].reshape(-1)
            v["pos"][:, 3] = 1.0
            v["normal"][:, :3] = N[sl].reshape(-1, 3)
            rr, cc = np.meshgrid(np.arange(vr - 1), np.arange(vc - 1), indexing="ij")
            a = (rr * vc + cc).reshape(-1)
            # counter-clockwise seen from +Y: (a, a+vc, a+1), (a+1, a+vc, a+vc+1)
            idx = np.stack([a, a + vc, a + 1, a + 1, a + vc, a + vc + 1], 1).reshape(-1).astype(np.uint32)
            alpha = rng.uniform(0.05, 0.6)
            film = rng.uniform(200.0, 800.0)
            mat = b.add_material([film, 0.0, 0.0, 1.0], [5, alpha, 1.33, 1.5])
            b.add_instance(b.add_mesh(v, idx), mat, _trs((0, 0, 0)))
    b.set_environment(sky_environment(sky_seed))
    return b


HEIGHTFIELD_VIEW = View((0.0, 0.9, 1.6), -HALF_PI, -0.55, 1920, 1080)


def heightfield_config(**kw):
    """Depth 6, environment light on, no photon pass (no caustic casters)."""
    c = wsrt.default_config(max_depth=6, enable_photons=0, env_map=1)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


_SPHERE_CACHE: dict = {}


def _sphere(sub: int):
    if sub not in _SPHERE_CACHE:
        _SPHERE_CACHE[sub] = wsrt.create_sphere(sub)
    return _SPHERE_CACHE[sub]


def _displaced_sphere(sub: int, rng, amp: float):
    """Icosphere(sub) with radial displacement 1 + amp * sum of 3 seeded
    sinusoidal lobes of the direction; normals left radial."""
    v, i = _sphere(sub)
    v = v.copy()
    p = v["pos"][:, :3].astype(np.float64)
    d = p / np.linalg.norm(p, axis=1, keepdims=True)
    disp = np.ones(len(p))
    for _ in range(3):
        k = rng.normal(size=3)
        k *= rng.uniform(2.0, 6.0) / np.linalg.norm(k)
        disp += amp * np.sin(d @ k + rng.uniform(0, 2 * math.pi)) / 3.0
    v["pos"][:, :3] = (p * disp[:, None]).astype(np.float32)
    return v, i


def _random_material(b, rng):
    kind = rng.integers(0, 4)
    col = list(rng.uniform(0.2, 0.9, 3)) + [1.0]
    if kind == 0:
        return b.add_material(col, [0, 0, 0, 0])
    if kind == 1:
        return b.add_material(col, [1, float(rng.uniform(0.05, 0.8)), 0, 0])
    if kind == 2:
        return b.add_material([1.0, 1.0, 1.0, 1.0], [2, 0, float(rng.uniform(1.3, 1.8)), 0.0])
    return b.add_material([1.0, 1.0, 1.0, 1.0], [4, float(rng.uniform(200, 800)), 1.33, 1.0])


BATCH_RESOLUTIONS = (128, 256, 512, 1024)


def build_batch_scene(b, index: int, seed: int = 11):
    """Scene `index` of the batch workload. Returns (builder, view, spp)."""
    rng = np.random.default_rng([seed, index])
    target = float(np.exp(rng.uniform(math.log(1e3), math.log(2e6))))
    floor = b.add_material([0.7, 0.7, 0.7, 1.0], [0, 0, 0, 0])
    light = b.add_material([6500.0, 4.0, 0.0, 1.0], [3, 0, 0, 0])
    plane = b.add_mesh(*wsrt.create_plane())
    cube = b.add_mesh(*wsrt.create_cube())
    b.add_instance(plane, floor, _trs((0, 0, 0), s=(8, 1, 8)))
    b.add_light_instance(plane, light, _trs((0, 3.0, 0), "x", PI, (1.5, 1, 1.5)), [0.0, 6500.0, 4.0, 0.0])
    tris, objects = 4, 0
    while tris + 12 <= target and objects < 48:
        left = target - tris
        sub = int(min(8, max(0, math.floor(math.log(max(left / 20.0, 1.0), 4)))))
        if 20 * 4**sub > left or rng.random() < 0.3:  # a box
            r = float(rng.uniform(0.1, 0.5))
            pos = (float(rng.uniform(-2.5, 2.5)), r, float(rng.uniform(-2.5, 2.5)))
            b.add_instance(cube, _random_material(b, rng),
                           _trs(pos, "y", float(rng.uniform(0, 2 * math.pi)), (2 * r, 2 * r, 2 * r)))
            tris += 12
        else:
            v, i = _displaced_sphere(sub, rng, float(rng.uniform(0.0, 0.25)))
            r = float(rng.uniform(0.2, 0.8))
            pos = (float(rng.uniform(-2.5, 2.5)), r * 1.25, float(rng.uniform(-2.5, 2.5)))
            b.add_instance(b.add_mesh(v, i), _random_material(b, rng), _trs(pos, s=(r, r, r)))
            tris += 20 * 4**sub
        objects += 1
    res = int(BATCH_RESOLUTIONS[rng.integers(0, 4)])
    spp = int(rng.integers(8, 65))
    yaw = float(rng.uniform(-math.pi, math.pi))
    dist = float(rng.uniform(4.0, 7.0))
    pos = (-dist * math.cos(yaw), float(rng.uniform(1.5, 3.5)), -dist * math.sin(yaw))
    pitch = -math.atan2(pos[1] - 0.5, dist)
    return b, View(pos, yaw, pitch, res, res), spp
