"""The GPU's image converges to the double-precision oracle's converged image.

Reference: tests/golden/cornell64_fp64.npz, the reference's Cornell box at
64 x 64 from the oracle built with double arithmetic, mean of 8192 samples
per pixel (frames 0 .. 8191), 2^14 photon paths per frame, constant gather
radius 0.02 (so every frame draws from one estimator), with the per-pixel
variance of one sample. The GPU renders frames 2^24 .., whose seeds are
independent of the reference's.

relMSE(I, R) = mean over pixels and rgb of (I - R)^2 / (R^2 + 1e-2). For two
independent unbiased estimates of the same image, E[relMSE] after N GPU
samples = C (1/N + 1/N_ref), C = mean(var / (R^2 + 1e-2)). A bias between
the GPU and the FP64 estimator adds a term that does not fall with N.

Bars: the GPU's variance constant (from two independent halves) within
[0.8, 1.25] of the FP64 one; relMSE within [0.6, 1.6] of C (1/N + 1/N_ref)
at N = 64, 256, 1024; relMSE(1024) / relMSE(64) <= 0.12 (0.070 expected).
"""
import os

import numpy as np
import pytest

import oracle as O
import wsrt

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell64_fp64.npz")
F0 = 1 << 24


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref**2 + 1e-2)))


def load_ref():
    z = np.load(GOLDEN)
    return z["mean"].astype(np.float64), z["var"].astype(np.float64), int(z["frames"])


def test_fixture_matches_the_config_and_camera():
    z = np.load(GOLDEN)
    assert int(z["width"]) == int(z["height"]) == 64
    assert int(z["photons_per_pass"]) == 1 << 14 and float(z["gather_radius_exponent"]) == 0.0
    # the GPU test's camera (the product's default camera) is the oracle's
    # (equal values; the inverses may differ in the sign of zeros)
    cam = wsrt.CameraController().build_uniform(1.0, 3)
    ocam = O.camera_uniform(aspect=1.0, frame=3)
    np.testing.assert_array_equal(np.frombuffer(bytes(cam), np.float32), np.frombuffer(bytes(ocam), np.float32))


def test_f32_oracle_converges_to_the_fp64_image():
    """The f32 restatement, on independent frames, is within the statistical
    band of the double-precision image (CPU)."""
    R, var, nref = load_ref()
    s = O.cornell_box()
    cfg = O.default_config(photons_per_pass=1 << 14, gather_radius_exponent=0.0)
    acc = np.zeros((64 * 64, 4), np.float32)
    n = 48
    for k in range(n):
        O.Frame(s, cfg, O.camera_uniform(aspect=1.0, frame=F0 + k), 64, 64).render(acc)
    C = float(np.mean(var / (R**2 + 1e-2)))
    got = relmse(acc[:, :3] / n, R)
    want = C * (1.0 / n + 1.0 / nref)
    print(f"f32 oracle, {n} spp: relMSE {got:.4e}, expected {want:.4e}")
    assert 0.6 * want <= got <= 1.6 * want


@pytest.mark.gpu
def test_gpu_image_converges_to_the_fp64_image(device):
    R, var, nref = load_ref()
    cfg = wsrt.default_config(photons_per_pass=1 << 14, gather_radius_exponent=0.0, frames_in_flight=64)
    g = wsrt.create_cornell_box(device)
    r = wsrt.Renderer(device, g, 64, 64, cfg)
    r.set_camera(wsrt.CameraController().build_uniform(1.0, 0))
    r.frame_count = F0  # accumulation continues (frame > 0) from zero
    snaps = {}
    done = 0
    for n in (64, 256, 512, 1024):
        r.render_frames(n - done)
        done = n
        acc = r.read_accum().astype(np.float64)
        assert np.all(acc[:, 3] == n)
        snaps[n] = acc[:, :3]
    C = float(np.mean(var / (R**2 + 1e-2)))
    a = snaps[512] / 512
    b = (snaps[1024] - snaps[512]) / 512
    c_gpu = 512 * float(np.mean((a - b) ** 2 / (R**2 + 1e-2))) / 2
    print(f"variance constant: GPU {c_gpu:.4f}, FP64 oracle {C:.4f}", flush=True)
    assert 0.8 <= c_gpu / C <= 1.25
    got = {}
    for n in (64, 256, 1024):
        got[n] = relmse(snaps[n] / n, R)
        want = C * (1.0 / n + 1.0 / nref)
        print(f"  {n} spp: relMSE {got[n]:.4e}, expected {want:.4e} ({got[n] / want:.3f})", flush=True)
        assert 0.6 * want <= got[n] <= 1.6 * want
    assert got[1024] / got[64] <= 0.12
