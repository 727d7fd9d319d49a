"""The CPU reference of the path-guiding network (oracle/guide_ref.py) pinned
by torch autograd in float64, finite differences and closed forms of the
von Mises-Fisher mixture. These tests need no GPU."""
import numpy as np
import pytest
import torch

import guide_ref as G


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(G.PARAMS)
    p[:] = rng.normal(0, 0.15 * scale, G.PARAMS)
    p[G.W1:G.B1].reshape(G.HID, G.IN)[:, G.FEAT:] = 0
    p[G.W5:G.B5].reshape(G.OUTPAD, G.HID)[G.OUT:] = 0
    p[G.B5 + G.OUT:] = 0
    return p


def random_batch(n, seed=1):
    rng = np.random.default_rng(seed)
    s = rng.random((n, 7)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(n) * (rng.random(n) > 0.2)
    return s, d, w


def torch_loss(p, s, d, w):
    """The same network and loss, written with torch ops, float64."""
    x = torch.tensor(G.encode(s))
    P = p
    Ws = [P[G.W1:G.B1].reshape(G.HID, G.IN), P[G.W2:G.B2].reshape(G.HID, G.HID),
          P[G.W3:G.B3].reshape(G.HID, G.HID), P[G.W4:G.B4].reshape(G.HID, G.HID),
          P[G.W5:G.B5].reshape(G.OUTPAD, G.HID)]
    bs = [P[G.B1:G.W2], P[G.B2:G.W3], P[G.B3:G.W4], P[G.B4:G.W5], P[G.B5:G.PARAMS]]
    h = x
    for k in range(4):
        h = torch.relu(h @ Ws[k].T + bs[k])
    o = (h @ Ws[4].T + bs[4])[:, :G.OUT].reshape(-1, G.LOBES, 5)
    m = o[:, :, :3]
    ln = torch.clamp(torch.linalg.norm(m, dim=2), min=1e-6)
    mu = m / ln[:, :, None]
    kap = torch.exp(torch.clamp(G.KAPPA_SCALE * o[:, :, 3], G.LOG_KAPPA_MIN, G.LOG_KAPPA_MAX))
    pi = torch.softmax(G.LOGIT_SCALE * o[:, :, 4], dim=1)
    norm = kap / (2 * np.pi * -torch.expm1(-2 * kap))
    dd = torch.tensor(d).reshape(-1, 1, 3)
    c = (mu * dd).sum(2)
    pdf = (pi * norm * torch.exp(kap * (c - 1))).sum(1)
    return torch.mean(-torch.tensor(w) * torch.log(pdf))


def test_hand_written_gradients_equal_autograd():
    p = random_params(0)
    s, d, w = random_batch(96)
    pt = torch.tensor(p, requires_grad=True)
    torch_loss(pt, s, d, w).backward()
    auto = pt.grad.numpy()
    mine = G.gradients(p, s, d, w, emulate=False)
    mask = np.ones(G.PARAMS, bool)
    mask[G.W1:G.B1].reshape(G.HID, G.IN)[:, G.FEAT:] = False
    mask[G.W5:G.B5].reshape(G.OUTPAD, G.HID)[G.OUT:] = False
    mask[G.B5 + G.OUT:] = False
    np.testing.assert_allclose(mine[mask], auto[mask], rtol=1e-7, atol=1e-12)
    assert float(torch_loss(torch.tensor(p), s, d, w)) == pytest.approx(G.loss(p, s, d, w), rel=1e-12)


def test_gradients_match_finite_differences():
    p = random_params(3)
    s, d, w = random_batch(16, seed=4)
    g = G.gradients(p, s, d, w, emulate=False)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.integers(G.W1, G.B1, 4), rng.integers(G.W3, G.B3, 4),
                          rng.integers(G.W5, G.W5 + G.OUT * G.HID, 6), np.arange(G.B5, G.B5 + G.OUT, 7)])
    for i in idx:
        if not np.any(G.gradients(p, s, d, w, emulate=False)[i]) and (i - G.W1) % G.IN >= G.FEAT:
            continue
        e = np.zeros(G.PARAMS)
        e[i] = 1e-6
        fd = (G.loss(p + e, s, d, w) - G.loss(p - e, s, d, w)) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-9), i


def test_encoding_is_sin_cos_of_frequency_bands():
    s = np.random.default_rng(0).random((50, 7)).astype(np.float32)
    x = G.encode(s)
    for c, (a, band, is_cos) in enumerate(G.feature_table()):
        if a < 0:
            assert np.all(x[:, c] == 0)
            continue
        arg = (2.0 ** band) * np.pi * s[:, a].astype(np.float64)
        want = np.cos(arg) if is_cos else np.sin(arg)
        np.testing.assert_allclose(x[:, c], want, atol=2e-9 * 2 ** band + 1e-12)
    assert sum(1 for t in G.feature_table() if t[0] >= 0) == 168


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5**0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(7)
    out = rng.normal(0, 1, (3, 40))
    out[:, 3::5] = rng.uniform(-2, 3.5, (3, 8)) / G.KAPPA_SCALE  # kappa from 0.14 to 33
    dirs = fib_sphere(400000)
    for o in out:
        p = G.pdf(np.repeat(o[None], len(dirs), 0), dirs)
        assert p.mean() * 4 * np.pi == pytest.approx(1.0, abs=2e-3)


def test_vmf_sampling_matches_the_density():
    rng = np.random.default_rng(11)
    for kappa in (0.3, 4.0, 60.0):
        mu = rng.normal(size=3)
        mu /= np.linalg.norm(mu)
        n = 200000
        d = G.vmf_sample_dirs(np.repeat(mu[None], n, 0), np.full(n, kappa), rng.random(n), rng.random(n))
        c = d @ mu
        # E[cos] = coth(kappa) - 1/kappa; cos theta is distributed ~ exp(kappa c) on [-1, 1]
        assert c.mean() == pytest.approx(1 / np.tanh(kappa) - 1 / kappa, abs=4e-3)
        h, edges = np.histogram(c, bins=20, range=(-1, 1))
        cdf = (np.exp(kappa * (edges - 1)) - np.exp(-2 * kappa)) / -np.expm1(-2 * kappa)
        want = np.diff(cdf) * n
        big = want > 200
        assert np.all(np.abs(h[big] - want[big]) < 6 * np.sqrt(want[big]) + 0.02 * want[big])
        # azimuth uniform around mu
        t = np.cross(mu, [0.0, 0.0, 1.0])
        t /= np.linalg.norm(t)
        b = np.cross(mu, t)
        ph = np.arctan2(d @ b, d @ t)
        hp, _ = np.histogram(ph[c < 0.999], bins=8, range=(-np.pi, np.pi))
        assert hp.min() > 0.9 * hp.mean()


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.3e-7, 65504.0], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0078125)
    assert r[3] == np.float32(1.0078125)
    assert abs(r[4] - x[4]) <= abs(x[4]) * 2**-8
    assert r[5] == 65536.0
