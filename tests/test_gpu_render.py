"""The render loop on the GPU against the oracle (photon pass, grid, camera
paths, accumulation), on the reference's Cornell box and on the caustic scene.

Tolerances (f32 on both sides; the GPU's exp/pow/tan/cos/sin differ from the
host libm by a few ulp, which can flip a threshold test such as rand() < F
and send a path down another branch):
  * photons: the stored/not-stored decision agrees for >= 99% of photon
    paths; positions of photons stored by both agree to 1e-3 for >= 99%;
  * grid: bit-exact on the GPU's own photon buffer;
  * camera paths (oracle fed the GPU's photon buffer): hero wavelengths agree
    to 1e-4 relative and the spectral radiance of a path to 1e-4 relative
    (absolute floor 1e-6 of the frame's mean) for >= 99% of paths;
  * accumulated image: relMSE = mean (I - R)^2 / (R^2 + 1e-2) <= 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import scenes
import wsrt

pytestmark = pytest.mark.gpu


def to_oracle_cam(cam):
    return O.CameraUniform.from_buffer_copy(bytes(cam))


def to_oracle_cfg(cfg):
    return O.Config.from_buffer_copy(bytes(cfg))


def setup(device, name, W, H, photons_per_pass):
    if name == "cornell":
        g = wsrt.create_cornell_box(device)
        o = O.cornell_box()
        cfg = wsrt.default_config(photons_per_pass=photons_per_pass)
        view = scenes.View((0.0, 0.0, 3.0), scenes.HALF_PI * -1, 0.0, W, H)
    elif name == "small_box":
        g = scenes.build_small_box(wsrt.SceneBuilder()).build(device)
        o = scenes.build_small_box(O.Scene()).build()
        cfg = scenes.small_box_config(photons_per_pass=photons_per_pass)
        v = scenes.SMALL_BOX_VIEW
        view = scenes.View(v.position, v.yaw, v.pitch, W, H)
    elif name == "heightfield":
        g = scenes.build_heightfield(wsrt.SceneBuilder(), n=65, tiles=4).build(device)
        o = scenes.build_heightfield(O.Scene(), n=65, tiles=4).build()
        cfg = scenes.heightfield_config(photons_per_pass=photons_per_pass)
        v = scenes.HEIGHTFIELD_VIEW
        view = scenes.View(v.position, v.yaw, v.pitch, W, H)
    elif name == "custom":
        g = scenes.build_custom(wsrt.SceneBuilder()).build(device)
        o = scenes.build_custom(O.Scene()).build()
        cfg = wsrt.default_config(photons_per_pass=photons_per_pass)
        view = scenes.View(scenes.DEFAULT_VIEW.position, scenes.DEFAULT_VIEW.yaw, scenes.DEFAULT_VIEW.pitch, W, H)
    else:
        g = scenes.build_caustic(wsrt.SceneBuilder()).build(device)
        o = scenes.build_caustic(O.Scene()).build()
        cfg = scenes.caustic_config(photons_per_pass=photons_per_pass)
        view = scenes.View(scenes.CAUSTIC_VIEW.position, scenes.CAUSTIC_VIEW.yaw, scenes.CAUSTIC_VIEW.pitch, W, H)
    return g, o, cfg, view


def rel_close(a, b, rel, floor):
    return np.all(np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + floor, axis=-1)


# the caustic case at 2^20 photons is the caustic workload's full photon pass
# (2^22 cells): its ~14k stored photons all come through the emission's cone
# and root-box filters, so a filter that dropped a path would show here
@pytest.mark.parametrize("name,frame,photons", [("cornell", 0, 1 << 16), ("cornell", 5, 1 << 16),
                                                ("caustic", 2, 1 << 16), ("custom", 3, 1 << 16),
                                                ("caustic", 7, 1 << 20)])
def test_photon_pass_and_grid(device, name, frame, photons):
    print(f"photons {name} frame {frame} photons {photons}", flush=True)
    g, o, cfg, view = setup(device, name, 64, 64, photons)
    r = wsrt.Renderer(device, g, view.width, view.height, cfg)
    r.set_camera(scenes.camera(view, frame))
    r.render()
    ph, stored = r.read_photons()
    ref = o.photon_trace(to_oracle_cfg(cfg), frame, 0, cfg.photons_per_pass)
    gs, os_ = ph["energy"] != 0, ref["energy"] != 0
    print(f"  stored {gs.sum()} oracle {os_.sum()} same decision {(gs == os_).mean():.6f}", flush=True)
    assert stored == gs.sum() and stored > 100
    # only paths that diverge numerically may differ (none did on these
    # cases); the emission's specular-root filter must drop no stored photon
    assert (gs == os_).mean() >= 0.999
    both = gs & os_
    ok = np.all(np.abs(ph["position"][both] - ref["position"][both]) <= 1e-3, axis=1)
    assert ok.mean() >= 0.99
    wl_ok = np.abs(ph["wavelength"][both] - ref["wavelength"][both]) <= 1e-3
    assert wl_ok.mean() >= 0.99
    # the rest of each record: direction (the reversed incoming ray) and
    # energy (flux of photon_emit.wgsl:184-208; the GPU's exp differs from
    # libm by a few ulp)
    dir_ok = np.all(np.abs(ph["direction"][both] - ref["direction"][both]) <= 1e-3, axis=1)
    assert dir_ok.mean() >= 0.99
    e_ok = np.abs(ph["energy"][both] - ref["energy"][both]) <= 1e-4 * np.abs(ref["energy"][both])
    print(f"  direction within 1e-3 {dir_ok.mean():.5f}, energy within 1e-4 {e_ok.mean():.5f}", flush=True)
    assert e_ok.mean() >= 0.99
    # grid of the frame: bit-exact with the oracle's counting sort of the GPU photons
    cell, count, offset, slot = r.read_grid()
    oc, ocount, ooff, oslot = O.grid_counting(ph, frame, cfg.cell_size, cfg.hash_size)
    np.testing.assert_array_equal(cell, oc)
    np.testing.assert_array_equal(count, ocount)
    np.testing.assert_array_equal(offset, ooff)
    np.testing.assert_array_equal(slot[: len(oslot)], oslot)


def test_photon_pass_without_specular_geometry(device):
    """With no metal, glass or thin-film instance the specular BVH is empty:
    no photon can become caustic, so none is stored (as in the oracle) and
    the grid is empty."""
    print("photons, diffuse-only scene", flush=True)
    g = scenes.build_occluded_room(wsrt.SceneBuilder(), n_boxes=0).build(device)
    o = scenes.build_occluded_room(O.Scene(), n_boxes=0).build()
    cfg = scenes.occluded_config(enable_photons=1, photons_per_pass=1 << 14, frames_in_flight=2)
    v = scenes.OCCLUDED_VIEW
    view = scenes.View(v.position, v.yaw, v.pitch, 32, 32)
    r = wsrt.Renderer(device, g, 32, 32, cfg)
    r.set_camera(scenes.camera(view, 3))
    r.render()
    ph, stored = r.read_photons()
    ref = o.photon_trace(to_oracle_cfg(cfg), 3, 0, cfg.photons_per_pass)
    assert stored == 0 and np.count_nonzero(ref["energy"]) == 0
    assert not np.any(ph["energy"])
    _, count, _, _ = r.read_grid()
    assert count.sum() == 0


@pytest.mark.parametrize("name,frame,photons", [("cornell", 0, 1 << 16), ("cornell", 9, 1 << 16),
                                                ("caustic", 1, 1 << 18), ("custom", 4, 1 << 17),
                                                ("heightfield", 2, 1 << 10), ("small_box", 3, 1 << 16)])
def test_camera_paths_match_oracle(device, name, frame, photons):
    print(f"paths {name} frame {frame}", flush=True)
    W = H = 96
    g, o, cfg, view = setup(device, name, W, H, photons)
    r = wsrt.Renderer(device, g, W, H, cfg)
    cam = scenes.camera(view, frame)
    r.set_camera(cam)
    r.render()
    wl, L = r.read_paths()
    ph, _ = r.read_photons()
    f = O.Frame(o, to_oracle_cfg(cfg), to_oracle_cam(cam), W, H, photons=ph)
    owl, oL, orgb = f.trace_pixels(np.arange(W * H, dtype=np.uint32))
    assert rel_close(wl, owl, 1e-4, 0.0).mean() >= 0.999
    floor = 1e-6 * max(float(np.abs(oL).mean()), 1e-12)
    same = rel_close(L, oL, 1e-4, floor)
    print(f"  paths within 1e-4: {same.mean():.5f}; nonzero {np.mean(np.abs(oL).sum(1) > 0):.3f}")
    assert same.mean() >= 0.99
    assert np.abs(oL).sum() > 0
    # the per-pixel tap returns the same records for chosen pixels, and
    # rejects a pixel outside the image
    ids = np.array([W * H - 1, 0, 7, W * H // 2], np.uint32)
    twl, tL = r.trace_paths(ids)
    np.testing.assert_array_equal(twl, wl[ids])
    np.testing.assert_array_equal(tL, L[ids])
    with pytest.raises(wsrt.WsrtError):
        r.trace_paths([W * H])


@pytest.mark.parametrize("name", ["cornell", "caustic", "custom", "heightfield", "small_box"])
def test_accumulated_image_matches_oracle(device, name):
    print(f"image {name}", flush=True)
    W = H = 64
    frames = 8
    g, o, cfg, view = setup(device, name, W, H, 1 << 16)
    r = wsrt.Renderer(device, g, W, H, cfg)
    r.set_camera(scenes.camera(view, 0))
    r.render_frames(frames)
    acc = r.read_accum()
    assert r.frame_count == frames
    ref = np.zeros((W * H, 4), np.float32)
    ocfg = to_oracle_cfg(cfg)
    for k in range(frames):
        cam = to_oracle_cam(scenes.camera(view, k))
        O.Frame(o, ocfg, cam, W, H).render(ref)
    np.testing.assert_array_equal(acc[:, 3], frames)
    img, want = acc[:, :3] / frames, ref[:, :3] / frames
    relmse = float(np.mean((img - want) ** 2 / (want**2 + 1e-2)))
    print(f"  relMSE {relmse:.3e}")
    assert relmse <= 1e-4
    rgba = r.read_rgba8()
    assert rgba[:, 3].min() == 255 and rgba[:, :3].max() > 0
    # the storage image (shader.wgsl:809-822) of the GPU's own accumulation
    np.testing.assert_array_equal(rgba, O.tone_map(acc))


def test_frame_count_semantics(device, cornell_pair):
    g, _ = cornell_pair
    r = wsrt.Renderer(device, g, 32, 32, wsrt.default_config(photons_per_pass=1 << 12))
    assert r.frame_count == 0
    r.render_frames(3)
    assert r.frame_count == 3
    np.testing.assert_array_equal(r.read_accum()[:, 3], 3)
    r.frame_count = 0  # camera moved (engine.rs:154-156): accumulation restarts
    r.render_frames(2)
    np.testing.assert_array_equal(r.read_accum()[:, 3], 2)
    # resume: exported accumulation written back is continued
    acc = r.read_accum()
    r2 = wsrt.Renderer(device, g, 32, 32, wsrt.default_config(photons_per_pass=1 << 12))
    r2.write_accum(acc)
    r2.frame_count = 2
    r2.render_frames(1)
    r.render_frames(1)
    np.testing.assert_array_equal(r2.read_accum(), r.read_accum())
    r.resize(1920, 1080)  # refits presentation only
    assert r.frame_count == 3


def test_traversal_counting(device, cornell_pair):
    """Counting node visits leaves the image unchanged; the extend kernel's
    lockstep visits lie between its summed visits and 64 times them."""
    g, _ = cornell_pair
    cfg = wsrt.default_config(photons_per_pass=1 << 12)
    plain = wsrt.Renderer(device, g, 32, 32, cfg)
    counted = wsrt.Renderer(device, g, 32, 32, cfg)
    counted.set_traversal_counting(True)
    plain.render_frames(2)
    counted.render_frames(2)
    np.testing.assert_array_equal(plain.read_accum(), counted.read_accum())
    assert plain.stats()["extend_wave_nodes"] == 0
    s = counted.stats()
    nodes = s["kernel_nodes"][1]
    assert s["kernel_counted_rays"][1] == s["kernel_rays"][1] > 0
    assert 0 < nodes <= s["extend_wave_nodes"] <= 64 * nodes


def test_zero_light_scene_renders_black(device):
    b = wsrt.SceneBuilder()
    m = b.add_material([0.5, 0.5, 0.5, 1], [0, 0, 0, 0])
    b.add_instance(b.add_mesh(*wsrt.create_cube()), m, wsrt.mat4_scale(1, 1, 1))
    s = b.build(device)
    r = wsrt.Renderer(device, s, 16, 16, wsrt.default_config(photons_per_pass=1 << 10))
    r.render()
    assert np.all(r.read_accum()[:, :3] == 0)
    _, stored = r.read_photons()
    assert stored == 0


def test_empty_scene(device):
    b = wsrt.SceneBuilder()
    b.add_material([0.5, 0.5, 0.5, 1], [0, 0, 0, 0])
    s = b.build(device)
    assert s.info.n_triangles == 0 and s.info.n_nodes == 0
    r = wsrt.Renderer(device, s, 8, 8, wsrt.default_config(photons_per_pass=256))
    r.render()
    assert np.all(r.read_accum()[:, :3] == 0)
    hits, flags = s.closest_hit(np.array([[0, 0, 0, 1e-4, 0, 0, -1, 100]], np.float32))
    assert flags[0] == 0


@pytest.mark.parametrize("name", ["cornell", "caustic"])
def test_frame_graph_replays_match_eager(device, name):
    """Each frame is one captured HIP graph; every replay (with eager frames
    of another renderer and synchronizations between replays) must equal the
    eagerly launched frame bit for bit."""
    print(f"graph vs eager {name}", flush=True)
    W = H = 48
    g, _, cfg, view = setup(device, name, W, H, 1 << 15)
    cfg_g = wsrt.Config.from_buffer_copy(bytes(cfg))
    cfg_e = wsrt.Config.from_buffer_copy(bytes(cfg))
    cfg_g.frame_graphs, cfg_e.frame_graphs = 1, 0
    rg = wsrt.Renderer(device, g, W, H, cfg_g)
    re = wsrt.Renderer(device, g, W, H, cfg_e)
    cam = scenes.camera(view, 0)
    rg.set_camera(cam)
    re.set_camera(cam)
    for k in range(4):
        rg.render_frames(1)
        re.render_frames(1)
        rg.sync()
        np.testing.assert_array_equal(rg.read_accum(), re.read_accum())
        np.testing.assert_array_equal(rg.read_paths()[1], re.read_paths()[1])
        np.testing.assert_array_equal(rg.read_grid()[1], re.read_grid()[1])
    rg.render_frames(5)  # several replays back to back
    re.render_frames(5)
    np.testing.assert_array_equal(rg.read_accum(), re.read_accum())
    # a single render() with an explicit camera frame_count also goes through the graph
    cam7 = scenes.camera(view, 7)
    rg.set_camera(cam7)
    re.set_camera(cam7)
    rg.render()
    re.render()
    np.testing.assert_array_equal(rg.read_accum(), re.read_accum())


def srgb8(c):
    """screenshot.rs:52-61 restated: u8 -> sRGB OETF -> u8."""
    cf = np.asarray(c, np.float32) / np.float32(255.0)
    v = np.where(cf <= 0.0031308, cf * np.float32(12.92),
                 np.float32(1.055) * np.power(cf, np.float32(1.0 / 2.4)) - np.float32(0.055))
    return np.round(np.clip(v, 0, 1) * 255).astype(np.uint8)


def test_screenshot_and_png(device, cornell_pair, tmp_path):
    g, _ = cornell_pair
    r = wsrt.Renderer(device, g, 40, 24, wsrt.default_config(photons_per_pass=1 << 12))
    r.render_frames(4)
    lin = r.read_rgba8().reshape(24, 40, 4)
    shot = r.screenshot()
    want = srgb8(lin[::-1])
    want[..., 3] = 255
    np.testing.assert_array_equal(shot, want)
    path = str(tmp_path / "shot.png")
    r.save_png(path)
    np.testing.assert_array_equal(wsrt.load_png(path), shot)


@pytest.mark.parametrize("name", ["cornell", "caustic"])
def test_frames_in_flight_do_not_change_results(device, name):
    """Frames traced together in one wavefront accumulate bit for bit like
    frames traced one at a time."""
    print(f"frames in flight {name}", flush=True)
    W = H = 40
    g, _, cfg, view = setup(device, name, W, H, 1 << 14)
    accs = []
    for S in (1, 3, 4):
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.frames_in_flight = S
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(scenes.camera(view, 0))
        r.render_frames(6)
        accs.append(r.read_accum())
    np.testing.assert_array_equal(accs[0], accs[1])
    np.testing.assert_array_equal(accs[0], accs[2])


@pytest.mark.parametrize("name", ["cornell", "caustic", "heightfield"])
def test_tail_launch_does_not_change_results(device, name):
    """Paths finished by the tail launch (from the bounce where fewer than
    tail_paths remain) accumulate bit for bit like paths traced bounce by
    bounce in the wavefront: tail off, tail from bounce 2 on, and a threshold
    that switches at a later bounce."""
    print(f"tail launch {name}", flush=True)
    W = H = 48
    g, _, cfg, view = setup(device, name, W, H, 1 << 14)
    out = []
    for tail in (0, 0xFFFFFFFF, 300):
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.frames_in_flight = 2
        c.tail_paths = tail
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(scenes.camera(view, 1))
        r.render_frames(4)
        out.append((r.read_accum(), r.read_paths()[1], r.stats()))
    for acc, L, st in out[1:]:
        np.testing.assert_array_equal(acc, out[0][0])
        np.testing.assert_array_equal(L, out[0][1])
        # the same rays, counted by whichever launch traced them
        assert st["extend_rays"] == out[0][2]["extend_rays"]
        assert st["shadow_rays"] == out[0][2]["shadow_rays"]


def test_autotune_picks_a_measured_candidate(device, cornell_pair):
    g, _ = cornell_pair
    cfg = wsrt.default_config(photons_per_pass=1 << 14)
    res = wsrt.autotune(device, g, 64, 64, cfg)
    ms = {c["frames_in_flight"]: c["ms_per_frame"] for c in res["candidates"]}
    assert cfg.frames_in_flight == res["best"] and ms[res["best"]] > 0
    assert ms[res["best"]] <= min(v for v in ms.values() if v > 0) * 1.02 + 1e-9
    # second stage: tail threshold and traversal grid at that setting
    assert len(res["stage2"]) == 7 and all(c["ms_per_frame"] > 0 for c in res["stage2"])
    assert (cfg.tail_paths, cfg.trace_rounds) == (res["best_tail_paths"], res["best_trace_rounds"])
    assert cfg.trace_rounds in (1, 2) and cfg.tail_paths in (0, 1 << 16, 1 << 18, 1 << 20)
    # third stage: lane refill threshold
    assert [(c["extend_refill"], c["ordered_queues"]) for c in res["stage3"]] == \
        [(16, 0), (24, 0), (32, 0), (0, 1), (16, 1), (24, 1), (32, 1)]
    assert all(c["ms_per_frame"] > 0 for c in res["stage3"])
    assert cfg.extend_refill == res["best_extend_refill"] in (0, 16, 24, 32)
    assert cfg.ordered_queues == res["best_ordered_queues"] in (0, 1)


def test_trace_rounds_do_not_change_results(device, cornell_pair):
    g, _ = cornell_pair
    out = []
    for rounds in (1, 2, 4):
        r = wsrt.Renderer(device, g, 40, 40, wsrt.default_config(photons_per_pass=1 << 14, frames_in_flight=2,
                                                                  trace_rounds=rounds))
        r.render_frames(4)
        out.append(r.read_accum())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_environment_requires_a_map(device, cornell_pair):
    g, _ = cornell_pair
    with pytest.raises(wsrt.WsrtError):
        wsrt.Renderer(device, g, 8, 8, wsrt.default_config(env_map=1))


@pytest.mark.parametrize("tile_mode", [0, 1])
def test_batch_equals_single_renderers(device, tile_mode):
    """One persistent launch over jobs of different scenes, sizes and sample
    counts gives each job the image its own renderer gives, bit for bit."""
    print(f"batch tile mode {tile_mode}", flush=True)
    specs = [("cornell", 40, 24, 3), ("caustic", 32, 32, 2), ("custom", 24, 40, 4), ("cornell", 9, 13, 5)]
    jobs, refs = [], []
    for k, (name, W, H, spp) in enumerate(specs):
        g, _, cfg, view = setup(device, name, W, H, 1 << 10)
        cam = scenes.camera(view, 0)
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.enable_photons = 0
        c.frames_in_flight = 1
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(cam)
        r.render_frames(spp)
        refs.append(r.read_accum())
        jobs.append((g, cam, W, H, spp))
    bcfg = wsrt.default_config(enable_photons=0)
    b = wsrt.Batch(device, jobs, bcfg)
    b.set_tuning(tile_mode, 2)
    b.render()
    b.render()  # the graph replay restarts every job at frame 0
    b.sync()
    for k in range(len(specs)):
        np.testing.assert_array_equal(b.read_accum(k), refs[k])
        assert b.read_rgba8(k)[:, 3].min() == 255
    tune = b.autotune()
    assert tune["tile_mode"] in (0, 1) and tune["blocks_per_cu"] in (1, 2, 4, 8)
    b.render()
    b.sync()
    for k in range(len(specs)):
        np.testing.assert_array_equal(b.read_accum(k), refs[k])


def test_batch_rejects_guiding(device, cornell_pair):
    g, _ = cornell_pair
    cam = scenes.camera(scenes.DEFAULT_VIEW, 0)
    with pytest.raises(wsrt.WsrtError):
        wsrt.Batch(device, [(g, cam, 8, 8, 1)], wsrt.default_config(guiding=1))


@pytest.mark.parametrize("graphs", [1, 0])
def test_batch_with_photons_equals_single_renderers(device, graphs):
    """With the photon pass on, a batch renders window by window (each job's
    photon passes and grids for the window's frames, then one persistent
    launch over all pixels of those frames): each
    job's image equals its own renderer's, photon gather included, bit for
    bit, with and without per-frame graphs and after autotuning."""
    print(f"batch with photons, graphs {graphs}", flush=True)
    specs = [("caustic", 32, 32, 3), ("cornell", 40, 24, 5), ("custom", 24, 40, 2), ("caustic", 9, 13, 4)]
    jobs, refs = [], []
    for name, W, H, spp in specs:
        g, _, cfg, view = setup(device, name, W, H, 1 << 14)
        cam = scenes.camera(view, 0)
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.photons_per_pass, c.cell_size, c.gather_radius, c.gather_radius_exponent = 1 << 14, 0.02, 0.02, -0.15
        c.frames_in_flight = 2
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(cam)
        r.render_frames(spp)
        refs.append(r.read_accum())
        jobs.append((g, cam, W, H, spp))
    bcfg = wsrt.default_config(photons_per_pass=1 << 14, frame_graphs=graphs)
    b = wsrt.Batch(device, jobs, bcfg)
    for _ in range(2):  # a second render restarts every job at frame 0
        b.render()
        b.sync()
        for k in range(len(specs)):
            np.testing.assert_array_equal(b.read_accum(k), refs[k])
    b.autotune()
    b.render()
    b.sync()
    for k in range(len(specs)):
        np.testing.assert_array_equal(b.read_accum(k), refs[k])


@pytest.mark.parametrize("name", ["cornell", "small_box"])
def test_quirk_fixes_match_oracle(device, name):
    """wsrt_config.quirks = all three WSRT_FIX_* bits: emission weighted by
    the balance heuristic after a Lambert bounce, roulette clamped at 1, and
    sign(0) = +1 in the GGX basis (the small box's GGX box tops and walls
    have z == 0 normals). Paths match the oracle with the same bits, and
    differ from the reference behaviour."""
    print(f"quirk fixes {name}", flush=True)
    W = H = 64
    g, o, cfg, view = setup(device, name, W, H, 1 << 14)
    out = {}
    for q in (0, 7):
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.quirks = q
        r = wsrt.Renderer(device, g, W, H, c)
        cam = scenes.camera(view, 2)
        r.set_camera(cam)
        r.render()
        wl, L = r.read_paths()
        ph, _ = r.read_photons()
        f = O.Frame(o, to_oracle_cfg(c), to_oracle_cam(cam), W, H, photons=ph)
        owl, oL, _ = f.trace_pixels(np.arange(W * H, dtype=np.uint32))
        floor = 1e-6 * max(float(np.abs(oL).mean()), 1e-12)
        same = rel_close(L, oL, 1e-4, floor).mean()
        print(f"  quirks {q}: paths within 1e-4 {same:.5f}", flush=True)
        assert same >= 0.99
        out[q] = L
    assert np.mean(np.any(out[0] != out[7], axis=1)) > 0.01


def test_quirk_bits_are_validated(device, cornell_pair):
    g, _ = cornell_pair
    with pytest.raises(wsrt.WsrtError):
        wsrt.Renderer(device, g, 8, 8, wsrt.default_config(quirks=8))
    with pytest.raises(wsrt.WsrtError):
        wsrt.Renderer(device, g, 8, 8, wsrt.default_config(quirks=8, guiding=1))


@pytest.mark.parametrize("name", ["cornell", "caustic", "heightfield", "custom"])
def test_extend_refill_does_not_change_results(device, name):
    """The closest-hit kernel that refills lanes as rays end traces the same
    rays to the same hits: images, path radiance and ray counts equal the
    plain extend kernel's bit for bit (the order of paths in the shade
    queues differs; paths are independent). With counting on, both count
    the same node visits."""
    print(f"extend refill {name}", flush=True)
    W = H = 48
    g, _, cfg, view = setup(device, name, W, H, 1 << 14)
    out = []
    for refill in (0, 1, 24):
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.frames_in_flight = 3
        c.extend_refill = refill
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(scenes.camera(view, 1))
        r.render_frames(6)
        acc, L = r.read_accum(), r.read_paths()[1]
        r.reset_stats()
        r.set_traversal_counting(True)
        r.render_frames(3)
        out.append((acc, L, r.stats()))
    a = out[0][2]
    for acc, L, b in out[1:]:
        np.testing.assert_array_equal(out[0][0], acc)
        np.testing.assert_array_equal(out[0][1], L)
        assert a["extend_rays"] == b["extend_rays"] and a["shadow_rays"] == b["shadow_rays"]
        assert a["kernel_nodes"][1] == b["kernel_nodes"][1] and a["kernel_tris"][1] == b["kernel_tris"][1]
        assert b["kernel_nodes"][1] <= b["extend_wave_nodes"] <= 64 * b["kernel_nodes"][1]


@pytest.mark.parametrize("name", ["cornell", "caustic", "heightfield", "guided"])
def test_ordered_queues_do_not_change_results(device, name):
    """Queues built in path order from per-path tags trace the same paths
    as queues in append order: images and path radiance bit for bit, ray
    counts equal (with the tail launch on, and for the guided tracer with
    the balance heuristic, whose network trains on the same records)."""
    print(f"ordered queues {name}", flush=True)
    W = H = 48
    if name == "guided":
        g = scenes.build_occluded_room(wsrt.SceneBuilder(), n_boxes=200).build(device)
        cfg = scenes.guided_config(guide_train_samples=4096, quirks=wsrt.FIX_EMISSION_MIS)
        v = scenes.OCCLUDED_VIEW
        view = scenes.View(v.position, v.yaw, v.pitch, W, H)
    else:
        g, _, cfg, view = setup(device, name, W, H, 1 << 14)
        cfg.tail_paths = 1 << 10
    out = []
    for ordered in (0, 1):
        c = wsrt.Config.from_buffer_copy(bytes(cfg))
        c.frames_in_flight = 3
        c.ordered_queues = ordered
        r = wsrt.Renderer(device, g, W, H, c)
        r.set_camera(scenes.camera(view, 1))
        r.render_frames(6)
        st = r.stats()
        out.append((r.read_accum(), r.read_paths()[1], st["extend_rays"], st["shadow_rays"]))
        r.close()
    a, b = out
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2:] == b[2:]
    assert np.abs(a[0][:, :3]).sum() > 0


def test_device_memory_limit_forces_out_of_memory(device, cornell_pair):
    """An allocation cap makes renderer creation fail with WSRT_E_OOM, frees
    what the failing call had allocated, and lifting it lets the same call
    succeed."""
    g, _ = cornell_pair
    used = device.memory_used
    try:
        device.set_memory_limit(used + (1 << 20))
        with pytest.raises(wsrt.WsrtError) as e:
            wsrt.Renderer(device, g, 256, 256, wsrt.default_config(photons_per_pass=1 << 14))
        assert e.value.code == -2
        assert device.memory_used == used
    finally:
        device.set_memory_limit(0)
    r = wsrt.Renderer(device, g, 256, 256, wsrt.default_config(photons_per_pass=1 << 14))
    assert device.memory_used > used + (1 << 20)
    r.render()
    r.close()
    assert device.memory_used == used


def test_batch_time_limit_and_cancel(device, cornell_pair):
    """A persistent batch launch stops when its wall-clock budget runs out or
    the host sets its stop flag: sync returns WSRT_E_ABORTED (-6), every wave
    still exits normally, and the next render, without the limit, gives each
    job's full image."""
    g, _ = cornell_pair
    cam = wsrt.CameraController().build_uniform(1.0, 0)
    jobs = [(g, cam, 512, 512, 64), (g, cam, 256, 256, 16)]
    cfg = wsrt.default_config(enable_photons=0, max_depth=8)
    ref = wsrt.Renderer(device, g, 256, 256, cfg)
    ref.set_camera(cam)
    ref.render_frames(16)
    want = ref.read_accum()
    ref.close()
    b = wsrt.Batch(device, jobs, cfg)
    b.set_time_limit(1)
    b.render()
    with pytest.raises(wsrt.WsrtError) as e:
        b.sync()
    assert e.value.code == -6 and "time limit" in str(e.value)
    b.set_time_limit(0)
    b.render()
    b.cancel()
    with pytest.raises(wsrt.WsrtError) as e:
        b.sync()
    assert e.value.code == -6 and "canceled" in str(e.value)
    b.render()
    b.sync()
    np.testing.assert_array_equal(b.read_accum(1), want)
    assert np.all(b.read_accum(0)[:, 3] == 64)


def test_nan_normals_render_without_error(device):
    """A mesh whose vertex normals are NaN: the renderer returns no error,
    the image stays finite (its paths end at the NaN-normal surface: their
    next rays are not traversed) and every path equals the oracle's."""
    def build(b):
        white = b.add_material([0.8, 0.8, 0.8, 1.0], [0, 0, 0, 0])
        light = b.add_material([6500.0, 2.0, 0.0, 1.0], [3, 0, 0, 0])
        v, i = wsrt.create_plane()
        bad = v.copy()
        bad["normal"][:, :3] = np.nan
        good = b.add_mesh(v, i)
        nan = b.add_mesh(bad, i)
        b.add_instance(good, white, scenes._trs((0, -1, 0), s=(4, 1, 4)))
        b.add_instance(nan, white, scenes._trs((0.0, 0.0, -1.0), "x", scenes.HALF_PI, (1, 1, 1)))
        b.add_light_instance(good, light, scenes._trs((0, 1.5, 0), "x", np.pi, (1, 1, 1)), [0.0, 6500.0, 2.0, 0.0])
        return b
    g = build(wsrt.SceneBuilder()).build(device)
    o = build(O.Scene()).build()
    cfg = wsrt.default_config(photons_per_pass=1 << 12, max_depth=6)
    W = H = 48
    r = wsrt.Renderer(device, g, W, H, cfg)
    cam = wsrt.CameraController().build_uniform(1.0, 0)
    r.set_camera(cam)
    r.render_frames(4)
    acc = r.read_accum()
    assert np.all(np.isfinite(acc)) and acc[:, :3].sum() > 0
    wl, L = r.read_paths()
    f = O.Frame(o, O.Config.from_buffer_copy(bytes(cfg)), O.CameraUniform.from_buffer_copy(bytes(
        wsrt.CameraController().build_uniform(1.0, 3))), W, H, photons=r.read_photons()[0])
    owl, oL, _ = f.trace_pixels(np.arange(W * H, dtype=np.uint32))
    assert np.all(np.isfinite(L)) and np.all(np.isfinite(oL))
    tol = 1e-4 * np.maximum(np.abs(L), np.abs(oL)) + 1e-6 * float(np.abs(oL).mean())
    assert np.mean(np.all(np.abs(L - oL) <= tol, axis=1)) >= 0.99
