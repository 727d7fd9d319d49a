"""The presented frame: the blit pass (blit_pass.rs:114-180, blit.wgsl) into
an sRGB window surface, and the photon overlay (debug_photons_pass.rs,
debug_photons.wgsl). The oracle's restatement (oracle.present) is pinned on
the CPU by closed forms: a same-size window is the screenshot image, a wider
or taller window adds blue bars of the predicted width, an integer scale
repeats texels, and one overlay point turns a blue-clear pixel into the
hand-computed sRGB code. On the GPU, wsrt_renderer_present must equal the
restatement bit for bit."""
import numpy as np
import pytest

import oracle as O

BLUE = np.array([0, 0, 255, 255], np.uint8)


def srgb8_closed(c):
    """8-bit linear -> 8-bit sRGB (screenshot.rs:52-61), written out."""
    v = np.asarray(c, np.float64) / 255.0
    e = np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055)
    return np.floor(e * 255.0 + 0.5).astype(np.uint8)


def image(W, H, seed=0):
    img = np.random.default_rng(seed).integers(0, 256, (H, W, 4), dtype=np.uint8)
    img[..., 3] = 255
    return img


def test_same_size_window_is_the_screenshot_image():
    img = image(6, 4)
    out = O.present(img, 6, 4)
    np.testing.assert_array_equal(out[..., :3], srgb8_closed(img[::-1, :, :3]))
    np.testing.assert_array_equal(out[..., 3], 255)


def test_wider_window_gets_side_bars():
    W, H = 8, 4
    img = image(W, H, 1)
    out = O.present(img, 2 * W, H)  # aspect 4 against 2: scale_x = 0.5
    assert (out[:, : W // 2] == BLUE).all() and (out[:, W // 2 + W:] == BLUE).all()
    np.testing.assert_array_equal(out[:, W // 2: W // 2 + W, :3], srgb8_closed(img[::-1, :, :3]))


def test_taller_window_gets_top_and_bottom_bars():
    W, H = 4, 6
    img = image(W, H, 2)
    out = O.present(img, W, 2 * H)  # scale_y = 0.5
    assert (out[: H // 2] == BLUE).all() and (out[H // 2 + H:] == BLUE).all()
    np.testing.assert_array_equal(out[H // 2: H // 2 + H, :, :3], srgb8_closed(img[::-1, :, :3]))


def test_integer_scale_repeats_texels():
    img = image(5, 3, 3)
    out = O.present(img, 10, 6)
    want = np.repeat(np.repeat(srgb8_closed(img[::-1, :, :3]), 2, axis=0), 2, axis=1)
    np.testing.assert_array_equal(out[..., :3], want)


def test_overlay_point_closed_form():
    # identity view and projection: clip = (x, y, z, 1), so a photon at
    # (x, y, 0.5) lands on pixel (floor((x + 1) / 2 w), floor((1 - y) / 2 h))
    img = image(4, 4, 4)
    eye = np.eye(4, dtype=np.float32).reshape(-1)
    ph = np.zeros(5, [("position", np.float32, 3), ("wavelength", np.float32),
                      ("direction", np.float32, 3), ("energy", np.float32)])
    ph["energy"] = 1.0
    ph["position"] = [(-0.95, 0.9, 0.5),  # pixel (0, 0) of a 20 x 10 window: a bar
                      (-0.95, 0.9, 0.5),  # the same pixel again
                      (0.9, -0.9, -0.5),  # behind the near plane: clipped
                      (0.0, 0.0, 0.5),    # skipped: energy 0
                      (0.95, 0.9, 0.5)]   # pixel (19, 0): one point
    ph["energy"][3] = 0.0
    base = O.present(img, 20, 10)
    out = O.present(img, 20, 10, ph, eye, eye)
    assert (base[0, 0] == BLUE).all() and (base[0, 19] == BLUE).all()
    # one point on the blue clear: r = 0 + 0.3, g = 0 + 0.18, b stays 1, in
    # linear light; sRGB codes 0.3 -> 149, 0.18 -> 118 by the OETF; alpha is
    # the shader's f32 0.3 as unorm8: 0.3f * 255 = 76.5000030 -> 77
    np.testing.assert_array_equal(out[0, 19], [149, 118, 255, 77])
    # two points: 0.3 -> code 149 -> linear 0.30054 + 0.3 = 0.60054, and
    # 1.055 * 0.60054^(1/2.4) - 0.055 = 0.79809 -> 203.51 -> 204
    assert out[0, 0, 0] == 204 and out[0, 0, 3] == 77
    changed = np.argwhere((out != base).any(axis=2))
    assert sorted(map(tuple, changed)) == [(0, 0), (0, 19)]
    # many points on one pixel saturate red and green at 255
    many = np.repeat(ph[:1], 40)
    sat = O.present(img, 20, 10, many, eye, eye)
    assert tuple(sat[0, 0]) == (255, 255, 255, 77)


@pytest.mark.gpu
@pytest.mark.parametrize("win", [(48, 32), (101, 32), (48, 70), (25, 13), (96, 64)])
def test_present_matches_oracle(device, cornell_pair, win):
    import wsrt

    g, _ = cornell_pair
    W, H = 48, 32
    print(f"present {win}", flush=True)
    r = wsrt.Renderer(device, g, W, H, wsrt.default_config(photons_per_pass=1 << 16))
    cam = wsrt.CameraController().build_uniform(W / H, 0)
    r.set_camera(cam)
    r.render_frames(2)
    img = r.read_rgba8().reshape(H, W, 4)
    ph, stored = r.read_photons()
    assert stored > 100
    np.testing.assert_array_equal(r.present(*win), O.present(img, *win))
    got = r.present(*win, overlay_photons=True)
    want = O.present(img, *win, ph, np.array(cam.view, np.float32), np.array(cam.proj, np.float32))
    np.testing.assert_array_equal(got, want)
    assert (got != r.present(*win)).any()
