"""The guiding network's BF16 MFMA kernels against the CPU reference
(oracle/guide_ref.py, BF16 rounding points emulated), and the guided
renderer against the plain one.

Tolerances: forward outputs 1e-2 relative to the output scale (the GPU's
sin/cos and FP32 sums can flip a BF16 rounding of an encoded feature or an
activation); gradients 3e-2 relative Frobenius error per weight tensor;
Adam from the GPU's own gradients to 1e-6.
"""
import numpy as np
import pytest

import guide_ref as G
import scenes
import wsrt

pytestmark = pytest.mark.gpu


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15 * scale, G.PARAMS)
    p[G.W1:G.B1].reshape(G.HID, G.IN)[:, G.FEAT:] = 0
    p[G.W5:G.B5].reshape(G.OUTPAD, G.HID)[G.OUT:] = 0
    p[G.B5 + G.OUT:] = 0
    return p.astype(np.float32)


def batch(n, seed=1):
    rng = np.random.default_rng(seed)
    s = rng.random((n, 7)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = (rng.random(n) * (rng.random(n) > 0.2)).astype(np.float32)
    return s, d.astype(np.float32), w


TENSORS = [("W1", G.W1, G.B1), ("b1", G.B1, G.W2), ("W2", G.W2, G.B2), ("W3", G.W3, G.B3),
           ("W4", G.W4, G.B4), ("W5", G.W5, G.B5), ("b5", G.B5, G.PARAMS)]


def test_initial_params_and_padding(device):
    g = wsrt.Guide(device, 4096, seed=7)
    p = g.params
    assert np.all(p[G.W1:G.B1].reshape(G.HID, G.IN)[:, G.FEAT:] == 0)
    assert np.all(p[G.W5:G.B5].reshape(G.OUTPAD, G.HID)[G.OUT:] == 0)
    assert np.abs(p[G.W1:G.B1]).max() > 0
    g.params = random_params(1)
    np.testing.assert_array_equal(g.params, random_params(1))


@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_forward_matches_reference(device, n):
    print(f"guide forward n={n}", flush=True)
    g = wsrt.Guide(device, 4096, seed=3)
    p = random_params(2)
    g.params = p
    s, _, _ = batch(n, seed=n)
    out = g.forward(s)
    ref, _ = G.forward(p, s, emulate=True)
    scale = np.abs(ref).max()
    err = np.abs(out - ref)
    print(f"  max err {err.max():.3e} scale {scale:.3e} mean err {err.mean():.3e}")
    assert err.max() <= 1e-2 * scale
    assert err.mean() <= 1e-3 * scale


@pytest.mark.parametrize("n", [256, 5000])
def test_gradients_match_reference(device, n):
    print(f"guide gradients n={n}", flush=True)
    g = wsrt.Guide(device, 8192, seed=3)
    p = random_params(4)
    g.params = p
    s, d, w = batch(n, seed=10 + n)
    gg = g.gradients(s, d, w)
    ref = G.gradients(p, s, d, w, emulate=True)
    for name, a, b in TENSORS:
        rel = np.linalg.norm(gg[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-30)
        print(f"  {name}: rel err {rel:.2e} |g| {np.linalg.norm(ref[a:b]):.3e}")
        assert rel <= 3e-2, name
    # and the emulated reference is itself close to the float64 gradient
    exact = G.gradients(p, s, d, w, emulate=False)
    assert np.linalg.norm(gg - exact) / np.linalg.norm(exact) <= 6e-2


def test_adam_step(device):
    g = wsrt.Guide(device, 4096, seed=3)
    p0 = random_params(6)
    g.params = p0
    s, d, w = batch(2048, seed=6)
    grads = g.gradients(s, d, w)
    g.train_step(s, d, w, lr=1e-2)
    p1 = g.params
    want, _, _ = G.adam(p0.astype(np.float64), grads.astype(np.float64), 0.0, 0.0, 1, 1e-2)
    np.testing.assert_allclose(p1, want, atol=2e-6)
    # a second step continues the moments (t = 2)
    grads2 = g.gradients(s, d, w)
    g.train_step(s, d, w, lr=1e-2)
    m = 0.1 * grads.astype(np.float64)
    v = 0.001 * grads.astype(np.float64) ** 2
    want2, _, _ = G.adam(p1.astype(np.float64), grads2.astype(np.float64), m, v, 2, 1e-2)
    np.testing.assert_allclose(g.params, want2, atol=5e-6)


def test_running_average_of_the_parameters(device):
    """The network the tracer samples with is the running average of the Adam
    iterates: avg_t = b_t avg_(t-1) + (1 - b_t) p_t with b_t = min(0.995,
    (1 + t) / (10 + t)), in float32 as written, restarted at the parameters
    by set_params, and restored by set_average."""
    g = wsrt.Guide(device, 4096, seed=3)
    p0 = random_params(7)
    g.params = p0
    np.testing.assert_array_equal(g.average, g.params)
    s, d, w = batch(2048, seed=8)
    avg = g.params.copy()
    for t in range(1, 4):
        g.train_step(s, d, w, lr=1e-2)
        b = np.float32(min(np.float32(0.995), np.float32(1 + t) / np.float32(10 + t)))
        avg = b * avg + (np.float32(1) - b) * g.params
        np.testing.assert_array_equal(g.average, avg)
    assert not np.array_equal(g.average, g.params)
    st = g.state()
    g.params = p0  # restarts the average
    np.testing.assert_array_equal(g.average, g.params)
    g.set_state(st)
    np.testing.assert_array_equal(g.average, st["average"])


def target_dirs(s, rng):
    """Directions from a vMF (kappa 20) whose mean depends on the inputs."""
    mu = np.stack([np.cos(6 * s[:, 0]), np.sin(6 * s[:, 0]), 2 * s[:, 3] - 1], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    n = len(s)
    return G.vmf_sample_dirs(mu, np.full(n, 20.0), rng.random(n), rng.random(n)).astype(np.float32)


def test_training_learns_an_input_dependent_distribution(device):
    """What the trained network gains: the log-likelihood of held-out
    directions rises from the untrained mixture's to near the target's."""
    print("guide training", flush=True)
    rng = np.random.default_rng(0)
    g = wsrt.Guide(device, 8192, seed=11)
    hs = rng.random((4096, 7)).astype(np.float32)
    hd = target_dirs(hs, rng)

    def loglik():
        return float(np.mean(np.log(G.pdf(g.forward(hs), hd))))

    before = loglik()
    for _ in range(400):
        s = rng.random((8192, 7)).astype(np.float32)
        g.train_step(s, target_dirs(s, rng), np.ones(8192, np.float32), lr=3e-3)
    after = loglik()
    # entropy bound of the target: E[log vMF(kappa 20)] = log(20 / 2 pi) - 1 + ... ~ 0.16
    print(f"  held-out log-likelihood {before:.3f} -> {after:.3f}")
    assert before < -1.5
    assert after > before + 1.5


def occluded_small():
    return scenes.build_occluded_room(wsrt.SceneBuilder(), n_boxes=2000)


def frame_means(r, frames):
    """Per-frame image means (rgb), from the accumulation differences."""
    out, prev = [], np.zeros(3)
    for k in range(frames):
        r.render()
        acc = r.read_accum()
        tot = acc[:, :3].sum(0, dtype=np.float64)
        out.append((tot - prev) / len(acc))
        prev = tot
    return np.array(out)


@pytest.mark.parametrize("quirks", [0, wsrt.FIX_EMISSION_MIS])
def test_guided_estimate_equals_plain(device, quirks):
    """One-sample MIS keeps the estimator unbiased: the mean image of the
    guided tracer equals the plain tracer's within the Monte Carlo error,
    with the reference's emission weighting and with the balance heuristic
    (light samples weighted against the guide's mixture pdf)."""
    print(f"guided vs plain mean, quirks {quirks}", flush=True)
    scene = occluded_small().build(device)
    view = scenes.View(scenes.OCCLUDED_VIEW.position, scenes.OCCLUDED_VIEW.yaw, scenes.OCCLUDED_VIEW.pitch,
                       64, 64)
    means = {}
    for name, cfg in (("plain", scenes.occluded_config(frames_in_flight=1, quirks=quirks)),
                      ("guided", scenes.guided_config(frames_in_flight=1, guide_train_samples=4096,
                                                      quirks=quirks))):
        r = wsrt.Renderer(device, scene, 64, 64, cfg)
        r.set_camera(scenes.camera(view, 0))
        means[name] = frame_means(r, 96)
        assert np.all(np.isfinite(means[name]))
    for c in range(3):
        a, b = means["plain"][:, c], means["guided"][:, c]
        z = (a.mean() - b.mean()) / np.sqrt(a.var() / len(a) + b.var() / len(b))
        print(f"  channel {c}: plain {a.mean():.4e} guided {b.mean():.4e} z {z:+.2f}")
        assert abs(z) < 4.5
    g = r.guide()
    assert np.all(np.isfinite(g.params))


def test_guided_renderer_resumes_bit_for_bit(device):
    """Checkpoint after two batches (accumulation, frame count, the network's
    parameters, Adam moments and step, trained-frame count), restore into a
    new renderer, and the next two batches equal the uninterrupted run's."""
    scene = occluded_small().build(device)
    W = H = 48
    view = scenes.View(scenes.OCCLUDED_VIEW.position, scenes.OCCLUDED_VIEW.yaw, scenes.OCCLUDED_VIEW.pitch, W, H)
    cfg = scenes.guided_config(frames_in_flight=4, guide_train_samples=4096, guide_train_frames=12)
    a = wsrt.Renderer(device, scene, W, H, cfg)
    a.set_camera(scenes.camera(view, 0))
    a.render_frames(8)
    ckpt = {"accum": a.read_accum(), "frame_count": a.frame_count, "guide": a.guide().state(),
            "trained": a.guide_frames_trained}
    assert ckpt["trained"] == 8 and ckpt["guide"]["step"] > 0
    a.render_frames(8)
    want = a.read_accum()
    assert a.guide_frames_trained == 12
    b = wsrt.Renderer(device, scene, W, H, cfg)
    b.set_camera(scenes.camera(view, 0))
    b.write_accum(ckpt["accum"])
    b.frame_count = ckpt["frame_count"]
    b.guide().set_state(ckpt["guide"])
    b.guide_frames_trained = ckpt["trained"]
    b.render_frames(8)
    np.testing.assert_array_equal(b.read_accum(), want)
    sa, sb = a.guide().state(), b.guide().state()
    for k in ("params", "m1", "m2", "average"):
        np.testing.assert_array_equal(sa[k], sb[k])
    assert sa["step"] == sb["step"]
