"""CPU reference of the path-guiding network (test infrastructure only).

Only tests/ load this module, as the checker of the HIP kernels in
wgpu-spectral-rt_amd/csrc/guide.hip; the product never imports it. The
reference has no guiding, so nothing here restates reference code: this is
the network the project defines, in numpy.

With ``emulate=True`` the BF16 rounding points of the GPU kernels are
reproduced (encoded inputs, weights, activations and back-propagated deltas
are rounded to BF16; sums are exact in float64 where the GPU accumulates in
FP32). With ``emulate=False`` everything is float64, which is what the torch
autograd check in the tests differentiates.

The hand-written loss gradient (``loss_grad_out``) and the hand-written
backward pass are pinned in the tests against torch autograd in float64 and
against finite differences.
"""
from __future__ import annotations

import numpy as np

SCALARS, BANDS, FEAT, IN, HID, OUT, OUTPAD, LOBES = 7, 12, 168, 192, 64, 40, 48, 8
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -3.0, 7.0
KAPPA_SCALE, LOGIT_SCALE = 8.0, 4.0  # kappa = exp(clamp(8 s)), logit = 4 a

# parameter layout (matches the ABI's WSRT_GUIDE_PARAMS layout)
W1 = 0
B1 = W1 + HID * IN
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * HID
W4 = B3 + HID
B4 = W4 + HID * HID
W5 = B4 + HID
B5 = W5 + OUTPAD * HID
PARAMS = B5 + OUTPAD


def bf16(x) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def feature_table():
    """(scalar, band, is_cos) of each of the 192 encoded columns; scalar -1 = padding.
    Column c = 8 q + j, slot q = 3 a + bq, band = 4 bq + j // 2, cos iff j odd."""
    t = []
    for c in range(IN):
        q, j = divmod(c, 8)
        if q >= 3 * SCALARS:
            t.append((-1, 0, 0))
        else:
            t.append((q // 3, 4 * (q % 3) + j // 2, j & 1))
    return t


def encode(s) -> np.ndarray:
    """sin / cos (2 pi frac(x 2^(band-1))) = sin / cos(2^band pi x) per column."""
    s = np.asarray(s, np.float32).reshape(-1, SCALARS)
    x = np.zeros((len(s), IN), np.float64)
    for c, (a, band, is_cos) in enumerate(feature_table()):
        if a < 0:
            continue
        t = np.ldexp(s[:, a], band - 1).astype(np.float32)
        u = (t - np.floor(t)).astype(np.float64)
        x[:, c] = np.cos(2 * np.pi * u) if is_cos else np.sin(2 * np.pi * u)
    return x


def unpack(p):
    p = np.asarray(p, np.float64)
    return ([p[W1:B1].reshape(HID, IN), p[W2:B2].reshape(HID, HID), p[W3:B3].reshape(HID, HID),
             p[W4:B4].reshape(HID, HID), p[W5:B5].reshape(OUTPAD, HID)],
            [p[B1:W2], p[B2:W3], p[B3:W4], p[B4:W5], p[B5:PARAMS]])


def forward(params, s, emulate: bool = True):
    """Outputs [N, 40] and the per-layer inputs (rows = samples)."""
    Ws, bs = unpack(params)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    h = r(encode(s)).astype(np.float64)
    inputs = [h]
    for layer in range(4):
        z = h @ r(Ws[layer]).astype(np.float64).T + bs[layer]
        h = r(np.maximum(z, 0.0)).astype(np.float64)
        inputs.append(h)
    out = h @ r(Ws[4]).astype(np.float64).T + bs[4]
    return out[:, :OUT], inputs


def mixture(out):
    o = np.asarray(out, np.float64).reshape(-1, LOBES, 5)
    m = o[:, :, 0:3]
    ln = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    mu = m / ln[:, :, None]
    sv = KAPPA_SCALE * o[:, :, 3]
    kap = np.exp(np.clip(sv, LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    lg = LOGIT_SCALE * o[:, :, 4]
    e = np.exp(lg - lg.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    norm = kap / (2 * np.pi * -np.expm1(-2 * kap))
    return mu, kap, pi, norm, ln, sv


def pdf(out, dirs) -> np.ndarray:
    mu, kap, pi, norm, _, _ = mixture(out)
    w = np.asarray(dirs, np.float64).reshape(-1, 1, 3)
    c = (mu * w).sum(axis=2)
    return (pi * norm * np.exp(kap * (c - 1))).sum(axis=1)


def loss_grad_out(out, dirs, weights, norm_factor):
    """d/d out of norm_factor * sum_i -w_i log p(dir_i) (hand-derived), [N, 48]."""
    mu, kap, pi, nrm, ln, sv = mixture(out)
    w = np.asarray(dirs, np.float64).reshape(-1, 1, 3)
    c = (mu * w).sum(axis=2)
    v = nrm * np.exp(kap * (c - 1))
    p = (pi * v).sum(axis=1, keepdims=True)
    wt = np.asarray(weights, np.float64).reshape(-1, 1) * norm_factor
    r = np.where(wt != 0, pi * v / np.maximum(p, 1e-30), 0.0)
    # 2 / (e^(2 kappa) - 1), written so that a large kappa does not overflow
    dlk = 1 / kap - 2 * np.exp(-2 * kap) / -np.expm1(-2 * kap) + (c - 1)
    in_range = (sv > LOG_KAPPA_MIN) & (sv < LOG_KAPPA_MAX)
    gs = np.where(in_range, -wt * r * kap * dlk * KAPPA_SCALE, 0.0)
    gm = (-wt * r * kap / ln)[:, :, None] * (w - mu * c[:, :, None])
    g = np.zeros((len(out), LOBES, 5))
    g[:, :, 0:3] = gm
    g[:, :, 3] = gs
    g[:, :, 4] = LOGIT_SCALE * wt * (pi - r)
    full = np.zeros((len(out), OUTPAD))
    full[:, :OUT] = g.reshape(len(out), OUT)
    return full


def backward(params, inputs, d5, emulate: bool = True) -> np.ndarray:
    """Hand-written backward pass: flat gradient of the parameters."""
    Ws, _ = unpack(params)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    g = np.zeros(PARAMS)
    woff = [W1, W2, W3, W4, W5]
    boff = [B1, B2, B3, B4, B5]
    d = r(d5).astype(np.float64)
    for layer in range(4, -1, -1):
        g[woff[layer]:woff[layer] + Ws[layer].size] = (d.T @ inputs[layer]).reshape(-1)
        g[boff[layer]:boff[layer] + len(Ws[layer])] = d.sum(axis=0)
        if layer == 0:
            break
        dh = d @ r(Ws[layer]).astype(np.float64)
        d = r(np.where(inputs[layer] > 0, dh, 0.0)).astype(np.float64)
    # padding stays zero
    g[W1:B1].reshape(HID, IN)[:, FEAT:] = 0.0
    g[W5:B5].reshape(OUTPAD, HID)[OUT:] = 0.0
    g[B5 + OUT:PARAMS] = 0.0
    return g


def gradients(params, s, dirs, weights, emulate: bool = True) -> np.ndarray:
    out, inputs = forward(params, s, emulate)
    d5 = loss_grad_out(out, dirs, weights, 1.0 / len(out))
    return backward(params, inputs, d5, emulate)


def loss(params, s, dirs, weights) -> float:
    out, _ = forward(params, s, emulate=False)
    w = np.asarray(weights, np.float64)
    return float(np.mean(-w * np.log(np.maximum(pdf(out, dirs), 1e-300))))


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def vmf_sample_dirs(mu, kappa, x1, x2):
    """vMF sampling by the inverse CDF of cos theta (the GPU's construction)."""
    mu = np.asarray(mu, np.float64)
    e = np.exp(-2 * kappa)
    w = np.clip(1 + np.log(e + x1 * (1 - e)) / kappa, -1, 1)
    sw = np.sqrt(np.maximum(0, 1 - w * w))
    sgn = np.where(mu[..., 2] >= 0, 1.0, -1.0)
    a = -1 / (sgn + mu[..., 2])
    b = mu[..., 0] * mu[..., 1] * a
    t1 = np.stack([1 + sgn * mu[..., 0] ** 2 * a, sgn * b, -sgn * mu[..., 0]], -1)
    t2 = np.stack([b, sgn + mu[..., 1] ** 2 * a, -mu[..., 1]], -1)
    ph = 2 * np.pi * x2
    d = mu * w[..., None] + t1 * (sw * np.cos(ph))[..., None] + t2 * (sw * np.sin(ph))[..., None]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)
