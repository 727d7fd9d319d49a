"""Benchmark of the render loop on one MI355X.

Workload (the caustic photon-mapping configuration): glass icosphere
(subdivision 4, 5120 triangles, radius 0.5, Cauchy A = 1.5, B = 0.0042 um^2)
on a 2-triangle diffuse plane under a point light 2 units above the sphere;
512 x 512 pixels, 2^20 photon paths per frame, 2^22 grid cells, cell size =
gather radius = 0.01, max depth 12. One step is one frame of the render
loop: photon pass, grid build, one camera path per pixel (1 spp) and
accumulation; --steps 64 renders the 64 spp image.

Prints one JSON line: Msamples/s (samples = camera paths) with rays/s and
photons/s beside it, the roofline of the dominant kernel measured with HIP
events on the renderer's stream during the timed region, and the CPU oracle
timed on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wgpu-spectral-rt_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
NODE_BYTES, TRI_BYTES, RAY_BYTES = 80, 48, 48  # BVH8 node, leaf triangle, ray in + hit out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seconds: float = 12.0) -> dict:
    """The oracle (plain-C restatement of the reference's shaders, OpenMP)
    rendering frames of the same workload until `seconds` have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import scenes

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scene = scenes.build_c2_caustic(O.Scene()).build()
    cfg = O.Config.from_buffer_copy(bytes(scenes.c2_config()))
    v = scenes.C2_VIEW
    acc = None
    frames, t0 = 0, time.perf_counter()
    while True:
        cam = O.CameraUniform.from_buffer_copy(bytes(scenes.camera(v, frames)))
        f = O.Frame(scene, cfg, cam, v.width, v.height)
        acc, _ = f.render(acc)
        frames += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(v.width * v.height * frames / dt / 1e6, 4), "unit": "Msamples/s",
            "cores": threads, "kind": "port",
            "sample": f"{frames} full frames (512x512, 2^20 photons each) in {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the render loop runs "
              "on a single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2

    import scenes
    import wsrt

    dev = wsrt.Device(0)
    scene = scenes.build_c2_caustic(wsrt.SceneBuilder()).build(dev)
    info = scene.info
    v = scenes.C2_VIEW
    cfg = scenes.c2_config()
    r = wsrt.Renderer(dev, scene, v.width, v.height, cfg)
    r.set_camera(scenes.camera(v, 0))
    r.render_frames(args.warmup)
    r.sync()
    # timed region: exactly `steps` frames submitted as captured batch graphs
    r.reset_stats()
    r.sync()
    t0 = time.perf_counter()
    r.render_frames(args.steps)
    r.sync()
    dt = time.perf_counter() - t0
    st = r.stats()
    # the same frames again with HIP events around every kernel on the
    # renderer's stream (launched one by one), for per-kernel device times
    r.reset_profile()
    r.set_profiling(True)
    r.render_frames(args.steps)
    r.sync()
    r.set_profiling(False)
    prof = r.profile()
    # traversal work per ray from the in-kernel counters (pinned against the
    # oracle's traversal of the same BVH8), measured on one extra frame
    r.reset_stats()
    r.set_traversal_counting(True)
    r.render_frames(1)
    r.sync()
    cnt = r.stats()
    r.set_traversal_counting(False)

    samples = v.width * v.height * args.steps
    msps = samples / dt / 1e6
    rays = st["extend_rays"] + st["shadow_rays"] + st["photon_rays"]
    kinds = {"photon_trace": 0, "extend": 1, "shadow": 2}
    per_ray = {}
    for k, i in kinds.items():
        n = max(cnt["trav_rays"][i], 1)
        per_ray[k] = {"nodes": cnt["trav_nodes"][i] / n, "tris": cnt["trav_tris"][i] / n}
    ray_counts = {"photon_trace": st["photon_rays"], "extend": st["extend_rays"],
                  "shadow": st["shadow_rays"]}
    kern = {}
    for name, p in prof.items():
        if p["launches"] == 0:
            continue
        e = {"ms": round(p["total_ms"], 3), "launches": p["launches"],
             "avg_us": round(1000 * p["total_ms"] / p["launches"], 2)}
        if name in per_ray:
            b = RAY_BYTES + per_ray[name]["nodes"] * NODE_BYTES + per_ray[name]["tris"] * TRI_BYTES
            e["bytes_per_ray"] = round(b, 1)
            e["GBps"] = round(ray_counts[name] * b / (p["total_ms"] * 1e-3) / 1e9, 1)
        if name == "grid_build":  # 60 N + 12 C bytes per frame, launches cover frames_in_flight frames
            n, c = cfg.photons_per_pass, cfg.hash_size
            e["GBps"] = round((60 * n + 12 * c) * args.steps / (p["total_ms"] * 1e-3) / 1e9, 1)
        kern[name] = e
    dom = max((k for k in kern if k in per_ray), key=lambda k: kern[k]["ms"])
    d = kern[dom]
    achieved = d["GBps"]
    result = {
        "metric": "Msamples/s + time-to-relMSE 0.01, guided vs plain tracer, 1 MI355X",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "caustic photon mapping: glass icosphere(4) 5120 tris + 2-tri plane + point "
                        "light, 512x512, 1 spp per step, 2^20 photons/pass, 2^22 cells, cell=radius=0.01, "
                        "max depth 12",
            "triangles": info.n_triangles, "bvh8_nodes": info.n_nodes,
            "width": v.width, "height": v.height, "photons_per_pass": cfg.photons_per_pass,
            "hash_size": cfg.hash_size, "max_depth": cfg.max_depth,
        },
        "rays_per_s": round(rays / dt, 1),
        "photons_per_s": round(st["photon_paths"] / dt, 1),
        "photons_stored_per_frame": round(st["photons_stored"] / args.steps, 1),
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes_per_ray": d["bytes_per_ray"],
            "avg_launch_us": d["avg_us"],
        },
        "kernels": kern,
        "traversal_per_ray": per_ray,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
