"""Parity at the benchmark workloads' own sizes and at the shipped settings.

  * caustic photon mapping at 2^20 photons per pass and 256 frames in flight
    (the autotuner's choice: 2^28 photon slots, 2^30 grid cells) equals
    frame-by-frame rendering bit for bit, images, photons and grid;
  * the frame-constant table wraps (frame 4095 -> 0) inside a batch;
  * the 2,093,058-triangle heightfield and the 2,097,182-triangle occluded
    room: closest and any hit against the oracle's own BVH, and camera paths
    at the workload's resolution against the oracle, pixel by pixel;
  * the guided tracer's mean image equals the plain tracer's on the full room;
  * the batch workload's procedural scenes (every material type, one of 1.5 M
    triangles) at their own resolutions and sample counts: the persistent
    launch equals each scene's renderer bit for bit, and the renderers' paths
    equal the oracle's.

Tolerances are those of test_gpu_render.py: hits bit-exact; paths 1e-4
relative for >= 99 % of paths (a few ulp of the GPU's exp/pow/sin/cos can
flip a branch such as rand() < F).
"""
import numpy as np
import pytest

import oracle as O
import scenes
import wsrt

pytestmark = pytest.mark.gpu


def rel_close(a, b, rel, floor):
    return np.all(np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + floor, axis=-1)


def cfg_copy(cfg, **kw):
    c = wsrt.Config.from_buffer_copy(bytes(cfg))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_256_frames_in_flight_equal_frame_by_frame(device):
    """The shipped setting of the caustic workload: 256 frames in one
    wavefront with 2^20 photon paths each."""
    print("caustic, 2^20 photons, 256 frames in flight vs 1", flush=True)
    g = scenes.build_caustic(wsrt.SceneBuilder()).build(device)
    view = scenes.View(scenes.CAUSTIC_VIEW.position, scenes.CAUSTIC_VIEW.yaw, scenes.CAUSTIC_VIEW.pitch, 32, 32)
    out = {}
    for S in (256, 1):
        r = wsrt.Renderer(device, g, 32, 32, scenes.caustic_config(frames_in_flight=S))
        r.set_camera(scenes.camera(view, 0))
        r.render_frames(256)
        r.sync()
        ph, stored = r.read_photons()
        out[S] = (r.read_accum(), ph, stored, r.read_grid(), r.read_paths())
        r.close()
    a, b = out[256], out[1]
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[0][:, 3], 256)
    assert a[2] == b[2] > 100
    # photons of the last frame: stored ones bit for bit
    keep = b[1]["energy"] != 0
    np.testing.assert_array_equal(a[1]["energy"] != 0, keep)
    np.testing.assert_array_equal(a[1][keep], b[1][keep])
    for x, y in zip(a[3], b[3]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a[4][1], b[4][1])


def test_frame_table_wraps_inside_a_batch(device, cornell_pair):
    """Frames 4093 .. 4100 span the end of the 4096-entry frame-constant
    table; a batch of 4 crossing it equals frames one at a time."""
    g, _ = cornell_pair
    accs = []
    for S in (4, 1):
        r = wsrt.Renderer(device, g, 24, 24, wsrt.default_config(photons_per_pass=1 << 12, frames_in_flight=S))
        r.frame_count = 4093
        r.render_frames(8)
        accs.append((r.read_accum(), r.read_paths()[1]))
        assert r.frame_count == 4101
    np.testing.assert_array_equal(accs[0][0], accs[1][0])
    np.testing.assert_array_equal(accs[0][1], accs[1][1])
    np.testing.assert_array_equal(accs[0][0][:, 3], 8)


def random_rays(bounds, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(bounds[:3]), np.array(bounds[3:])
    c, ext = (lo + hi) / 2, (hi - lo) / 2 * 1.2
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = c + rng.uniform(-1, 1, (n, 3)) * ext
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 1e-4, 100.0
    return rays


def check_hits(g, o, seed):
    rays = random_rays(list(g.info.bounds), 20000, seed)
    gh, gf = g.closest_hit(rays)
    oh, of = o.closest_hit(rays)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gh, oh)
    assert (gf & 1).mean() > 0.2
    rays[:, 3] = 1e-3
    rays[:, 7] = np.random.default_rng(seed + 1).uniform(0.01, 1.0, len(rays))
    occ = g.any_hit(rays)
    np.testing.assert_array_equal(occ, o.any_hit(rays))
    assert 0.05 < occ.mean() < 0.95


def check_paths(device, g, o, cfg, view, frame, n_pixels, seed, min_lit=0.2):
    W, H = view.width, view.height
    r = wsrt.Renderer(device, g, W, H, cfg)
    cam = scenes.camera(view, frame)
    r.set_camera(cam)
    r.render()
    ids = np.sort(np.random.default_rng(seed).choice(W * H, n_pixels, replace=False)).astype(np.uint32)
    wl, L = r.trace_paths(ids)
    f = O.Frame(o, O.Config.from_buffer_copy(bytes(cfg)), O.CameraUniform.from_buffer_copy(bytes(cam)), W, H)
    owl, oL, _ = f.trace_pixels(ids)
    assert rel_close(wl, owl, 1e-4, 0.0).mean() >= 0.999
    floor = 1e-6 * max(float(np.abs(oL).mean()), 1e-12)
    same = rel_close(L, oL, 1e-4, floor)
    print(f"  {n_pixels} paths of {W}x{H}: within 1e-4 {same.mean():.5f}; nonzero {np.mean(np.abs(oL).sum(1) > 0):.3f}",
          flush=True)
    assert same.mean() >= 0.99
    assert np.mean(np.abs(oL).sum(1) > 0) > min_lit
    return r


def test_heightfield_at_full_size(device):
    print("heightfield 1024x1024 vertices, 1920x1080", flush=True)
    g = scenes.build_heightfield(wsrt.SceneBuilder()).build(device)
    o = scenes.build_heightfield(O.Scene()).build()
    assert g.info.n_triangles == o.n_triangles == 2093058
    check_hits(g, o, 31)
    check_paths(device, g, o, scenes.heightfield_config(), scenes.HEIGHTFIELD_VIEW, 2, 4096, 32)


@pytest.fixture(scope="module")
def full_room(device):
    g = scenes.build_occluded_room(wsrt.SceneBuilder()).build(device)
    o = scenes.build_occluded_room(O.Scene()).build()
    return g, o


def test_occluded_room_at_full_size(device, full_room):
    print("occluded room, 2.1M triangles, 1024x1024", flush=True)
    g, o = full_room
    assert g.info.n_triangles == o.n_triangles == 2097182
    check_hits(g, o, 41)
    # lit only through the door gap: about 5 % of one-sample paths carry light
    check_paths(device, g, o, scenes.occluded_config(frames_in_flight=4), scenes.OCCLUDED_VIEW, 5, 16384, 42,
                min_lit=0.02)


@pytest.mark.parametrize("quirks", [0, wsrt.FIX_EMISSION_MIS])
def test_guided_mean_equals_plain_on_the_full_room(device, full_room, quirks):
    """One-sample MIS keeps the guided estimator unbiased on the workload's
    own scene: per-frame image means of guided and plain tracers agree
    within their Monte Carlo error (z < 4.5 per channel) at 256 x 256, with
    the reference's emission weighting and with the balance heuristic."""
    print(f"guided vs plain mean, full room, 256x256, quirks {quirks}", flush=True)
    g, _ = full_room
    v = scenes.OCCLUDED_VIEW
    view = scenes.View(v.position, v.yaw, v.pitch, 256, 256)
    means = {}
    for name, cfg in (("plain", scenes.occluded_config(frames_in_flight=1, quirks=quirks)),
                      ("guided", scenes.guided_config(frames_in_flight=1, guide_train_samples=1 << 14,
                                                      quirks=quirks))):
        r = wsrt.Renderer(device, g, 256, 256, cfg)
        r.set_camera(scenes.camera(view, 0))
        out, prev = [], np.zeros(3)
        for _ in range(64):
            r.render()
            tot = r.read_accum()[:, :3].sum(0, dtype=np.float64)
            out.append((tot - prev) / (256 * 256))
            prev = tot
        means[name] = np.array(out)
        assert np.all(np.isfinite(means[name]))
        r.close()
    for c in range(3):
        a, b = means["plain"][:, c], means["guided"][:, c]
        z = (a.mean() - b.mean()) / np.sqrt(a.var() / len(a) + b.var() / len(b))
        print(f"  channel {c}: plain {a.mean():.4e} guided {b.mean():.4e} z {z:+.2f}", flush=True)
        assert abs(z) < 4.5


class _Materials:
    """Records the material types a scene definition adds."""

    def __init__(self, b):
        self.b, self.types = b, set()

    def add_material(self, color, extra):
        self.types.add(int(extra[0]))
        return self.b.add_material(color, extra)

    def __getattr__(self, name):
        return getattr(self.b, name)


@pytest.mark.parametrize("photons", [0, 1])
def test_batch_of_procedural_scenes(device, photons):
    """Scenes 0..7 of the batch workload (1.5 M triangles at most) at their
    own resolutions and sample counts in one persistent launch (with the
    photon pass: 2^20 photon paths per job and frame, run for a window of
    frames at once, then one persistent launch per window)."""
    print(f"batch: 8 procedural scenes, photons {photons}", flush=True)
    cfg = wsrt.default_config(enable_photons=photons, max_depth=8)
    jobs, refs, types, tris = [], [], set(), []
    for k in range(8):
        rec = _Materials(wsrt.SceneBuilder())
        b, v, spp = scenes.build_batch_scene(rec, k)
        g = b.build(device)
        types |= rec.types
        tris.append(g.info.n_triangles)
        cam = scenes.camera(v, 0)
        r = wsrt.Renderer(device, g, v.width, v.height, cfg_copy(cfg, frames_in_flight=8))
        r.set_camera(cam)
        r.render_frames(spp)
        refs.append(r.read_accum())
        jobs.append((g, cam, v.width, v.height, spp))
        if k in (1, 3):  # paths of the last frame against the oracle
            o = scenes.build_batch_scene(O.Scene(), k)[0].build()
            ids = np.sort(np.random.default_rng(k).choice(v.width * v.height, 2048, replace=False)).astype(np.uint32)
            wl, L = r.trace_paths(ids)
            lc = scenes.camera(v, spp - 1)
            ph = r.read_photons()[0] if photons else None
            f = O.Frame(o, O.Config.from_buffer_copy(bytes(cfg)), O.CameraUniform.from_buffer_copy(bytes(lc)),
                        v.width, v.height, photons=ph)
            owl, oL, _ = f.trace_pixels(ids)
            floor = 1e-6 * max(float(np.abs(oL).mean()), 1e-12)
            same = rel_close(L, oL, 1e-4, floor).mean()
            print(f"  scene {k}: {tris[-1]} triangles, paths within 1e-4 {same:.5f}", flush=True)
            assert rel_close(wl, owl, 1e-4, 0.0).mean() >= 0.999 and same >= 0.99
        r.close()
    assert types == {0, 1, 2, 3, 4} and max(tris) > 1_000_000
    batch = wsrt.Batch(device, jobs, cfg)
    batch.autotune()
    batch.render()
    batch.sync()
    for k in range(len(jobs)):
        np.testing.assert_array_equal(batch.read_accum(k), refs[k])
