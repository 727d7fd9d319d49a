"""Pins the CPU oracle before it is trusted as the checker.

The reference ships no tests, fixtures or golden vectors (its validation is
visual), so every oracle function is pinned here by what the mathematics fixes:
integer closed forms restated independently in Python, textbook limits and
special cases, brute force on small inputs, invariants and statistical
identities. Each check is chosen so a plausible slip (wrong constant, swapped
operand, dropped term, wrong index) fails it.

Reference functions pinned (src/renderer/shaders/*.wgsl, src/scene/*.rs):
  pcg_hash / init_rng / rand ......... shader.wgsl:144-160   integer KATs + Python restatement
  grid jitter / hash_position ........ build_grid.wgsl:25-52 KAT + numpy f32 restatement
  hero wavelengths ................... shader.wgsl:734-797   closed forms, unbiasedness
  cie / rgb->spectrum / blackbody .... shader.wgsl:85-139    f64 formulas, special cases, Wien
  thin film, Schlick ................. shader.wgsl:233-268   limits that reduce to Fresnel
  geometry generators ................ geometry.rs:12-232    counts, radius, winding
  camera uniform ..................... camera.rs:178-194     centre-ray KAT, inverse identity
  counting-sort grid ................. build_grid.wgsl:54-73 linked-list equivalence
  closest / any hit .................. (driver)              BVH == brute force, plane closed form
"""
import math

import numpy as np
import pytest

import oracle as O


def pcg_py(x: int) -> int:
    s = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def test_pcg_hash_kats():
    assert [O.pcg_hash(i) for i in range(4)] == [0x07BB2FE2, 0xA8BEEA3C, 0x7A7ECC88, 0x7F0EF6BC]


def test_pcg_hash_matches_python_restatement():
    rng = np.random.default_rng(1)
    for x in list(rng.integers(0, 2**32, 2000, dtype=np.uint64)) + [0, 1, 2**31, 2**32 - 1]:
        assert O.pcg_hash(int(x)) == pcg_py(int(x))


def test_rand_stream_first_values():
    # pixel (0,0), frame 0: seed = pcg(0); first rand() = pcg(pcg(0)) / 2^32
    r = O.rand_stream(0, 3)
    assert r[0] == np.float32(0.19039936)
    s = pcg_py(0)
    for k in range(3):
        s = pcg_py(s)
        assert r[k] == np.float32(np.float32(s) / np.float32(4294967296.0))


def test_rand_reaches_one():
    # f32(0xFFFFFFFF) rounds to 2^32, so rand() == 1.0 is reachable (Q12)
    assert np.float32(np.uint32(0xFFFFFFFF)) / np.float32(4294967295.0) == np.float32(1.0)


def test_grid_jitter_kat_frame0():
    j = np.zeros(3, np.float32)
    O.lib.or_grid_jitter_f(0, 0.02, j)
    np.testing.assert_array_equal(j, np.array([0.0077254903, -0.0063137254, 0.004666667], np.float32))


def _jitter_np(frame: int, cell: float) -> np.ndarray:
    seed = pcg_py((frame * 114514) & 0xFFFFFFFF)
    r = np.array([seed & 0xFF, (seed >> 8) & 0xFF, (seed >> 16) & 0xFF], np.float32) / np.float32(255.0)
    return (r - np.float32(0.5)) * np.float32(cell)


@pytest.mark.parametrize("frame", [0, 1, 7, 63, 1000])
def test_hash_position_matches_numpy_restatement(frame):
    rng = np.random.default_rng(frame)
    cell, hs = np.float32(0.02), 1 << 22
    pts = rng.uniform(-3, 3, (3000, 3)).astype(np.float32)
    j = _jitter_np(frame, cell)
    g = np.floor((pts + j) / cell).astype(np.int64).astype(np.uint32).astype(np.uint64)
    h = ((g[:, 0] * 73856093) ^ (g[:, 1] * 19349663) ^ (g[:, 2] * 83492791)) & (hs - 1)
    got = np.array([O.lib.or_hash_position_f(np.ascontiguousarray(p), frame, cell, hs) for p in pts])
    np.testing.assert_array_equal(got, h.astype(np.uint32))


def test_cauchy_constants():
    cmin, crange, pnorm = (np.ctypeslib.ctypes.c_float() for _ in range(3))
    O.lib.or_cauchy_constants(cmin, crange, pnorm)
    assert cmin.value == pytest.approx(0.0715278, abs=2e-7)
    assert cmin.value + crange.value == pytest.approx(0.9439967, abs=2e-7)
    assert pnorm.value == pytest.approx(0.00912095, rel=1e-6)
    # f64 closed form of the truncated Cauchy CDF bounds
    assert cmin.value == pytest.approx(math.atan(-175 / 40) / math.pi + 0.5, abs=1e-7)
    # the pdf integrates to 1 over [380, 780]
    lam = np.linspace(380, 780, 400001)
    pdf = pnorm.value / (1 + ((lam - 555) / 40) ** 2)
    assert np.trapezoid(pdf, lam) == pytest.approx(1.0, rel=1e-5)


def test_cauchy_sample_inverse_cdf():
    # sampling at u maps to the quantile of the truncated distribution
    for u in [0.0, 0.1, 0.5, 0.9, 1.0]:
        lam = O.lib.or_cauchy_sample_f(u)
        cmin = math.atan(-175 / 40) / math.pi + 0.5
        cmax = math.atan(225 / 40) / math.pi + 0.5
        want = 555 + 40 * math.tan(math.pi * (cmin + u * (cmax - cmin) - 0.5))
        assert lam == pytest.approx(min(max(want, 380), 780), abs=2e-3)


def test_hero_wavelengths_strata_and_rotation():
    out = np.zeros(4, np.float32)
    for u in np.linspace(0, 1, 37):
        for rot in [0.1, 0.3, 0.6, 0.9]:
            O.lib.or_hero_wavelengths(float(u), rot, out)
            assert np.all((out >= 380) & (out < 780))
            s = np.sort(out)
            np.testing.assert_allclose(np.diff(s), 100.0, atol=1e-3)
    O.lib.or_hero_wavelengths(0.5, 0.1, out)
    assert out[0] == pytest.approx(O.lib.or_cauchy_sample_f(0.5), abs=1e-4)  # rot < 0.25 keeps hero first
    O.lib.or_hero_wavelengths(0.5, 0.3, out)
    assert out[3] == pytest.approx(O.lib.or_cauchy_sample_f(0.5), abs=1e-4)  # .yzwx puts hero last


def test_hero_estimator_is_unbiased():
    # sum_i f(l_i) / sum_j p(l_j) estimates the integral of f over [380, 780]
    rng = np.random.default_rng(3)
    cmin, crange, pnorm = (np.ctypeslib.ctypes.c_float() for _ in range(3))
    O.lib.or_cauchy_constants(cmin, crange, pnorm)
    out = np.zeros(4, np.float32)
    f = lambda l: np.exp(-((l - 600) / 50) ** 2)  # noqa: E731
    est = []
    for u, rot in rng.random((20000, 2)):
        O.lib.or_hero_wavelengths(float(u), float(rot), out)
        p = pnorm.value / (1 + ((out.astype(np.float64) - 555) / 40) ** 2)
        est.append(f(out.astype(np.float64)).sum() / p.sum())
    truth = 50 * math.sqrt(math.pi) * 0.5 * (math.erf(180 / 50) + math.erf(220 / 50))
    assert np.mean(est) == pytest.approx(truth, rel=0.01)


def test_cie_matches_f64():
    out = np.zeros(3, np.float32)
    for lam in [380.0, 442.0, 500.0, 555.0, 599.8, 700.0, 780.0]:
        O.lib.or_cie_xyz_f(lam, out)
        g = lambda mu, s: math.exp(-((lam - mu) / s) ** 2)  # noqa: E731
        x = 1.056 * g(599.8, 37.9) + 0.362 * g(442.0, 16.0) - 0.065 * g(501.1, 26.7)
        y = 0.821 * g(568.8, 46.9) + 0.286 * g(530.9, 16.3)
        z = 1.217 * g(437.0, 11.8) + 0.681 * g(459.0, 26.0)
        np.testing.assert_allclose(out, [max(x, 0), max(y, 0), max(z, 0)], rtol=2e-6, atol=1e-7)


def test_rgb_to_spectrum_special_cases():
    for g in [0.0, 0.3, 0.73, 0.98]:
        vals = [O.lib.or_rgb_to_spectrum_f(np.array([g] * 3, np.float32), l) for l in (400, 555, 700)]
        assert vals == pytest.approx([g] * 3, abs=1e-7)  # gray is flat
    assert O.lib.or_rgb_to_spectrum_f(np.ones(3, np.float32), 555) == pytest.approx(0.99)  # clamp
    red = np.array([0.65, 0.05, 0.05], np.float32)
    assert O.lib.or_rgb_to_spectrum_f(red, 630) == pytest.approx(0.65, abs=1e-6)  # peak = white + c.r
    assert O.lib.or_rgb_to_spectrum_f(red, 630 + 20) == pytest.approx(0.05 + 0.6 / math.e, rel=1e-5)


def test_blackbody_planck_and_wien():
    h, c, k = 6.62607015e-34, 299792458.0, 1.380649e-23
    for T in [3000.0, 6500.0]:
        for lam in [380.0, 500.0, 780.0]:
            lm = lam * 1e-9
            want = 2 * h * c * c / (lm**5 * (math.exp(h * c / (k * T * lm)) - 1)) * 1e-13
            assert O.lib.or_blackbody_f(lam, T) == pytest.approx(want, rel=2e-5)
    lam = np.arange(300.0, 900.0, 0.5)
    vals = np.array([O.lib.or_blackbody_f(float(x), 6500.0) for x in lam])
    assert lam[np.argmax(vals)] == pytest.approx(2.897771955e6 / 6500.0, abs=1.0)


def test_thin_film_reduces_to_fresnel():
    # film index equal to the outside: one interface n1|n3 at normal incidence
    for n3 in [1.0, 1.33, 1.5]:
        r = O.lib.or_thin_film_f(1.0, 1.0, 1.0, n3, 500.0, 550.0)
        assert r == pytest.approx(((1 - n3) / (1 + n3)) ** 2, abs=1e-6)
    # soap film in air at normal incidence, closed-form Airy reflectance
    n2, d, lam = 1.33, 500.0, 600.0
    r = ((1 - n2) / (1 + n2))
    delta = 4 * math.pi * n2 * d / lam
    want = (2 * r * r - 2 * r * r * math.cos(delta)) / (1 + r**4 - 2 * r * r * math.cos(delta))
    assert O.lib.or_thin_film_f(1.0, 1.0, n2, 1.0, d, lam) == pytest.approx(want, rel=1e-4)
    # periodic in thickness with period lambda / (2 n2 cos theta2)
    a = O.lib.or_thin_film_f(1.0, 1.0, 1.33, 1.5, 300.0, 600.0)
    b = O.lib.or_thin_film_f(1.0, 1.0, 1.33, 1.5, 300.0 + 600.0 / (2 * 1.33), 600.0)
    assert a == pytest.approx(b, abs=2e-5)
    # total internal reflection into a rarer substrate
    assert O.lib.or_thin_film_f(0.1, 1.0, 1.0, 0.5, 400.0, 500.0) == 1.0


def test_schlick():
    for n in [1.2, 1.5, 2.0]:
        assert O.lib.or_schlick_f(1.0, n) == pytest.approx(((1 - n) / (1 + n)) ** 2, rel=1e-6)
        assert O.lib.or_schlick_f(0.0, n) == pytest.approx(1.0, rel=1e-6)


def test_geometry_counts_and_winding():
    for kind, nv, ni in [("plane", 4, 6), ("cube", 24, 36)]:
        v, i = O.geometry(kind)
        assert (len(v), len(i)) == (nv, ni)
    for s, (nv, nf) in {0: (12, 20), 1: (42, 80), 2: (162, 320), 4: (2562, 5120)}.items():
        v, i = O.geometry("sphere", s)
        assert (len(v), len(i) // 3) == (nv, nf)
        np.testing.assert_allclose(np.linalg.norm(v["pos"][:, :3], axis=1), 0.5, rtol=1e-6)
        np.testing.assert_allclose(np.linalg.norm(v["normal"][:, :3], axis=1), 1.0, rtol=1e-6)
    # winding: cross(v1 - v0, v2 - v0) points along the stored normal (front-face rule)
    for kind, sub in [("plane", 0), ("cube", 0), ("sphere", 3)]:
        v, i = O.geometry(kind, sub)
        p = v["pos"][:, :3].astype(np.float64)
        t = i.reshape(-1, 3)
        n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
        assert np.all(np.einsum("ij,ij->i", n, v["normal"][t[:, 0], :3]) > 0)


def test_cornell_scene_layout():
    s = O.cornell_box()
    assert s.n_triangles == 5 * 2 + 2 + 12 + 5120
    L = s.lights()
    assert len(L) == 2
    e1 = L["v1"][:, :3] - L["v0"][:, :3]
    e2 = L["v2"][:, :3] - L["v0"][:, :3]
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    np.testing.assert_allclose(area, 0.125, rtol=1e-6)
    n = np.cross(e1, e2)
    assert np.all(n[:, 1] < 0)  # the light faces down into the box
    np.testing.assert_allclose(L["v0"][:, 1], 0.99, rtol=1e-6)
    np.testing.assert_array_equal(L["params"][0], [0.0, 6500.0, 1.0, 0.0])


def test_camera_centre_ray_and_inverse():
    cam = O.camera_uniform(aspect=16 / 9)
    vi = np.array(cam.view_inverse, np.float64).reshape(4, 4).T
    pi = np.array(cam.proj_inverse, np.float64).reshape(4, 4).T
    v = np.array(cam.view, np.float64).reshape(4, 4).T
    p = np.array(cam.proj, np.float64).reshape(4, 4).T
    np.testing.assert_allclose(vi @ v, np.eye(4), atol=1e-6)
    np.testing.assert_allclose(pi @ p, np.eye(4), atol=1e-6)
    np.testing.assert_allclose(vi[:3, 3], [0, 0, 3], atol=1e-6)
    t = vi @ pi @ np.array([0, 0, 1, 1.0])
    d = t[:3] / t[3] - vi[:3, 3]
    np.testing.assert_allclose(d / np.linalg.norm(d), [0, 0, -1], atol=1e-6)
    # vertical field of view of 45 degrees
    t = vi @ pi @ np.array([0, 1, 1, 1.0])
    d = t[:3] / t[3] - vi[:3, 3]
    assert math.degrees(math.atan2(d[1], -d[2])) == pytest.approx(22.5, abs=1e-4)


def _rays_in_box(n, seed, tmin=1e-4, tmax=100.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.95, 0.95, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = o, tmin, d, tmax
    return r


def test_bvh_equals_brute_force():
    s = O.cornell_box()
    rays = _rays_in_box(20000, 5)
    hb, fb = s.closest_hit(rays, brute=True)
    hv, fv = s.closest_hit(rays, brute=False)
    np.testing.assert_array_equal(hb, hv)
    np.testing.assert_array_equal(fb, fv)
    # the box is open towards +z only: rays leaving through z = +1 miss
    d = rays[:, 4:7]
    assert np.all((fb & 1)[d[:, 2] < 0])
    assert 0.5 < (fb & 1).mean() < 1.0
    rays[:, 7] = 0.3
    np.testing.assert_array_equal(s.any_hit(rays, True), s.any_hit(rays, False))
    # any hit <=> a closest hit exists within (tmin, tmax)
    h, f = s.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(s.any_hit(rays, True), f & 1)


def test_closest_hit_matches_plane_closed_form():
    s = O.cornell_box()
    n = 4000
    rng = np.random.default_rng(9)
    o = np.column_stack([rng.uniform(-0.9, 0.9, n), np.full(n, 0.5), rng.uniform(0.2, 0.9, n)])
    d = np.column_stack([rng.uniform(-0.1, 0.1, n), np.full(n, -1.0), rng.uniform(-0.1, 0.1, n)])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, 100
    hits, flags = s.closest_hit(rays)
    t = hits[:, 0].view(np.float32)
    tris, ids = s.world()
    hit = (flags & 1).astype(bool)
    floor = np.zeros(n, bool)
    floor[hit] = ids[hits[hit, 1], 0] == 0  # instance 0 is the floor quad
    rays64 = rays.astype(np.float64)
    want = (-1.0 - rays64[:, 1]) / rays64[:, 5]
    sel = floor & (flags & 1).astype(bool)
    assert sel.mean() > 0.5
    np.testing.assert_allclose(t[sel], want[sel], rtol=1e-5)
    assert np.all(flags[sel] & 2)  # seen from above the floor's front face is hit


def test_counting_sort_equals_linked_list():
    rng = np.random.default_rng(11)
    n, hs = 50000, 1 << 12  # small table: many collisions and long chains
    ph = np.zeros(n, O.PHOTON_DTYPE)
    ph["position"] = rng.normal(0, 0.05, (n, 3))
    ph["energy"] = rng.uniform(0.1, 2, n)
    ph["energy"][rng.random(n) < 0.2] = 0.0  # skipped photons
    for frame in [0, 5]:
        cell, count, offset, slots = O.grid_counting(ph, frame, 0.02, hs)
        head, nxt = O.grid_linked(ph, frame, 0.02, hs)
        valid = ph["energy"] != 0
        assert count.sum() == valid.sum() == len(slots)
        assert np.all(cell[~valid] == 0xFFFFFFFF)
        np.testing.assert_array_equal(offset, np.concatenate([[0], np.cumsum(count)[:-1]]))
        assert sorted(slots.tolist()) == np.flatnonzero(valid).tolist()
        for c in np.flatnonzero(count)[:500]:
            seg = slots[offset[c]:offset[c] + count[c]]
            chain, p = [], head[c]
            while p != 0xFFFFFFFF:
                chain.append(p)
                p = nxt[p]
            assert list(seg) == sorted(chain)  # same set; the segment is in slot order
            assert np.all(cell[seg] == c)


# ---- integrator closed forms ---------------------------------------------------
PLANCK_H, PLANCK_C, PLANCK_K = 6.62607015e-34, 299792458.0, 1.380649e-23


def planck_1e13(lam_nm, t):
    lam = np.asarray(lam_nm, np.float64) * 1e-9
    return 2 * PLANCK_H * PLANCK_C**2 / (lam**5 * np.expm1(PLANCK_H * PLANCK_C / (lam * PLANCK_K * t))) * 1e-13


def _cauchy_pnorm():
    cmin, crange, pnorm = (np.ctypeslib.ctypes.c_float() for _ in range(3))
    O.lib.or_cauchy_constants(cmin, crange, pnorm)
    return pnorm.value


def _cie(lam):
    out = np.zeros(3, np.float32)
    O.lib.or_cie_xyz_f(float(lam), out)
    return out.astype(np.float64)


def test_first_hit_emission_and_accumulation_closed_form():
    """A camera facing an emitter wall: each path's spectral radiance is
    blackbody(lambda, T) * intensity at its four hero wavelengths, and the
    pixel colour is XYZ (CIE fit) over the summed Cauchy pdf, to linear sRGB
    (shader.wgsl:586-631, 771-799)."""
    import wsrt

    s = O.Scene()
    light = s.add_material([6500.0, 2.5, 0.0, 1.0], [3, 0, 0, 0])
    plane = s.add_mesh(*wsrt.create_plane())
    m = wsrt.mat4_mul(wsrt.mat4_translation(0, 0, -1), wsrt.mat4_mul(wsrt.mat4_rotation("x", float(np.float32(np.pi / 2))),
                                                                     wsrt.mat4_scale(40, 40, 40)))
    s.add_light_instance(plane, light, m, [0.0, 6500.0, 2.5, 0.0])
    sc = s.build()
    cfg = O.default_config(enable_photons=0, max_depth=1)
    f = O.Frame(sc, cfg, O.camera_uniform(), 16, 16)
    wl, L, rgb = f.trace_pixels(np.arange(256, dtype=np.uint32))
    want = planck_1e13(wl, 6500.0) * 2.5
    np.testing.assert_allclose(L, want, rtol=2e-5)
    # accumulation: sum_i cie(l_i) L_i / sum_i p(l_i), XYZ -> linear sRGB
    pdf = _cauchy_pnorm() / (1 + ((wl.astype(np.float64) - 555.0) / 40.0) ** 2)
    xyz = np.zeros((len(wl), 3))
    for i in range(4):
        xyz += np.array([_cie(lam) for lam in wl[:, i]]) * L[:, i:i + 1]
    xyz /= pdf.sum(1, keepdims=True)
    M = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    np.testing.assert_allclose(rgb, xyz @ M.T, rtol=2e-4, atol=1e-6)


def test_point_light_direct_lighting_closed_form():
    """Lambert floor (gray 0.5: a flat spectrum) lit by a point light at the
    camera, one bounce: L = (0.5 / pi) B(lambda) I cos^3(phi) / h^2 with phi
    the angle of the centre pixel's ray from the vertical (NEE of a delta
    light has weight 1 / pdf = 1 / d^2, shader.wgsl:650-671)."""
    import wsrt

    s = O.Scene()
    gray = s.add_material([0.5, 0.5, 0.5, 1.0], [0, 0, 0, 0])
    s.add_instance(s.add_mesh(*wsrt.create_plane()), gray, wsrt.mat4_scale(50, 50, 50))
    h, I = 2.0, 3.0
    s.add_point_light([0.0, h, 0.0], [1.0, 6500.0, I, 0.0])
    sc = s.build()
    cfg = O.default_config(enable_photons=0, max_depth=1)
    pitch = float(np.float32(-89.0 * np.pi / 180))
    cam = O.camera_uniform(pos=(0.0, h, 0.0), pitch=pitch)
    W = 101
    f = O.Frame(sc, cfg, cam, W, W)
    wl, L, _ = f.trace_pixels(np.array([50 + 50 * W], np.uint32))
    phi = np.radians(1.0)
    want = 0.5 / np.pi * planck_1e13(wl, 6500.0) * I * np.cos(phi) ** 3 / h**2
    np.testing.assert_allclose(L, want, rtol=3e-4)


def test_photon_gather_density_closed_form():
    """Density estimate of shader.wgsl:297-342 for hand-placed photons in one
    cell: E (1 - d/r) max(1 - |dl|/20, 0) 3/(pi r^2) / 20 / 2^20, photons
    facing away from the normal or beyond the radius excluded."""
    sc = O.cornell_box()
    cfg = O.default_config()
    ph = np.zeros(4, O.PHOTON_DTYPE)
    p = np.array([0.005, 0.005, 0.005], np.float32)
    ph[0] = (p, 550.0, (0.0, 1.0, 0.0), 2.0)                      # counted
    ph[1] = (p + np.float32([0.0, 0.0, 0.004]), 565.0, (0.0, 1.0, 0.0), 3.0)  # counted
    ph[2] = (p, 550.0, (0.0, -1.0, 0.0), 5.0)                     # faces away
    ph[3] = (p + np.float32([0.0, 0.0, 0.0199]), 550.0, (0.0, 1.0, 0.0), 7.0)  # outside r once offset
    f = O.Frame(sc, cfg, O.camera_uniform(), 8, 8, photons=ph)
    q = p + np.float32([0.003, 0.0, 0.0])
    wl = np.array([550.0, 560.0, 575.0, 590.0], np.float32)
    got = f.gather(q, np.array([0.0, 1.0, 0.0], np.float32), wl)
    r = 0.02
    norm = 3.0 / (np.pi * r * r) / 20.0 / 2**20
    want = np.zeros(4)
    for pos, lam, e in ((ph[0]["position"], 550.0, 2.0), (ph[1]["position"], 565.0, 3.0)):
        d = np.linalg.norm(q.astype(np.float64) - pos)
        assert d < r
        want += e * (1 - d / r) * np.maximum(1 - np.abs(wl - lam) / 20.0, 0) * norm
    assert np.linalg.norm(q.astype(np.float64) - ph[3]["position"]) > r
    np.testing.assert_allclose(got, want, rtol=1e-5)
