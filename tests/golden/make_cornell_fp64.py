"""Writes cornell64_fp64.npz: the converged image of the reference's Cornell
box (cornell_box.rs) at 64 x 64 from the double-precision build of the
oracle, the reference image that the GPU's image must converge to.

Config: the reference's defaults (max depth 12, Russian roulette after
bounce 3, 20 nm spectral radius, gather cap 64), 2^14 photon paths per frame
and a constant gather radius of 0.02 (gather_radius_exponent 0), so that
every frame draws from the same estimator and frames of other indices are
independent samples of the same image. Frames 0 .. N-1 (N = 8192 samples per
pixel). Stored: the mean linear sRGB per pixel (rows as the storage
texture, row 0 the image bottom), the per-pixel variance of one sample's
rgb, and the frame count.

Run: python tests/golden/make_cornell_fp64.py [frames]   (about 7 minutes on
8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

W = H = 64
CONFIG = dict(photons_per_pass=1 << 14, gather_radius_exponent=0.0)


def main(frames: int = 8192) -> None:
    s = O.cornell_box(fp64=True)
    cfg = O.default_config(**CONFIG)
    acc = np.zeros((W * H, 4))
    sq = np.zeros((W * H, 3))
    t = time.time()
    for k in range(frames):
        cam = O.camera_uniform(aspect=W / H, frame=k)
        one = np.zeros((W * H, 4))
        O.Frame(s, cfg, cam, W, H).render(one)  # frame k > 0 adds to `one` (zero)
        acc += one
        sq += one[:, :3] ** 2
        if k % 1024 == 1023:
            print(f"{k + 1} frames, {time.time() - t:.0f} s", flush=True)
    mean = acc[:, :3] / frames
    var = sq / frames - mean**2
    np.savez_compressed(os.path.join(HERE, "cornell64_fp64.npz"), mean=mean.astype(np.float32),
                        var=var.astype(np.float32), frames=frames, width=W, height=H,
                        photons_per_pass=CONFIG["photons_per_pass"],
                        gather_radius_exponent=CONFIG["gather_radius_exponent"])


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8192)
