"""Benchmark of the render loop on one MI355X.

Workload (the caustic photon-mapping configuration): glass icosphere
(subdivision 4, 5120 triangles, radius 0.5, Cauchy A = 1.5, B = 0.0042 um^2)
on a 2-triangle diffuse plane under a point light 2 units above the sphere;
512 x 512 pixels, 2^20 photon paths per frame, 2^22 grid cells, cell size =
gather radius = 0.01, max depth 12. One step is the workload's image at 64
spp: 64 frames of the render loop (each a photon pass, grid build, one
camera path per pixel and accumulation). The timed steps run at the
frames in flight the autotuner picks, every batch one replay of its
captured HIP graph, exactly as the library renders; a batch that does not
fill frames_in_flight runs as a smaller batch with its own graph.

Prints one JSON line: Msamples/s (samples = camera paths) with rays/s and
photons/s beside it, the roofline of the dominant kernel (its device time
from HIP events on the renderer's stream over one batch at the same
setting), and the CPU oracle timed on a bounded sample of the same workload
on the host cores.
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
STEP_FRAMES = 64  # one step: the caustic workload's 512 x 512 image at 64 spp, one frame per spp
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
    scene = scenes.build_caustic(O.Scene()).build()
    cfg = O.Config.from_buffer_copy(bytes(scenes.caustic_config()))
    v = scenes.CAUSTIC_VIEW
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


# the kernels a summary names, by their short symbol (the closest-hit kernel
# is the lane-refill one where the workload's configuration uses it)
KERNEL_SYMBOLS = {"extend": ("k_extend", "void k_extend<false>", "void k_extend_refill<false>"), "shadow": ("k_shadow",),
                  "photon_trace": ("k_photon_bounce", "void k_photon_bounce<false>")}


def _symbol(kernel: str, *per_kernel: dict) -> str:
    """The short symbol of `kernel` present in every one of the per-kernel
    dictionaries."""
    for sym in KERNEL_SYMBOLS[kernel]:
        if all(sym in d for d in per_kernel):
            return sym
    raise KeyError(f"no symbol of {kernel} in the profile")
# profiled workloads: file tag -> description (--profile-workload picks the
# one a summary is written for)
PROFILE_WORKLOADS = {
    "caustic512": "caustic photon mapping, 512x512 (glass icosphere 5120 tris, 2^20 photons/frame, 2^22 cells)",
    "room1024": "occluded-light room, 2,097,182 tris, 1024x1024, plain tracer, 128 frames in flight",
    "guidedroom1024": "occluded-light room, 2,097,182 tris, 1024x1024, the guided-vs-plain leg (both tracers, "
                      "both estimators), batches launched kernel by kernel",
}
PROFILE_TAG = "caustic512"
EAGER = False  # the guided leg's renderers launch batches eagerly (--eager)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_{}_{}.txt")


def _short(name: str) -> str:
    for p in ("wsrt::(anonymous namespace)::", "wsrt::"):
        name = name.replace(p, "")
    return name.split("(")[0]


def _per_dispatch(csv_path: str, counter: str) -> dict:
    """Mean of a counter per dispatch, by kernel, from a rocprofv3
    --pmc counter_collection.csv."""
    import csv
    import collections

    tot, disp = collections.defaultdict(float), collections.defaultdict(set)
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            k = _short(row["Kernel_Name"])
            tot[k] += float(row["Counter_Value"])
            disp[k].add(row["Dispatch_Id"])
    return {k: tot[k] / len(disp[k]) for k in tot}


def summarize_traffic(pmc_dir: str, kernel: str, command: str) -> int:
    """Writes the HBM traffic of one kernel per launch from four rocprofv3
    --pmc runs in pmc_dir: calib_fetch / calib_write (FETCH_SIZE, WRITE_SIZE
    of the plain copy of --copy-calibration 256) and fetch / write (the same
    counters over the bench run). Each counter is scaled so that the copy
    reads its known size."""
    copy_kib = 256 * 1024
    cf = _per_dispatch(os.path.join(pmc_dir, "calib_fetch_counter_collection.csv"), "FETCH_SIZE")
    cw = _per_dispatch(os.path.join(pmc_dir, "calib_write_counter_collection.csv"), "WRITE_SIZE")
    copy_name = max(cf, key=lambda k: cf[k])
    f_scale, w_scale = copy_kib / cf[copy_name], copy_kib / cw[copy_name]
    fd = _per_dispatch(os.path.join(pmc_dir, "fetch_counter_collection.csv"), "FETCH_SIZE")
    wd = _per_dispatch(os.path.join(pmc_dir, "write_counter_collection.csv"), "WRITE_SIZE")
    sym = _symbol(kernel, fd, wd)
    fk, wk = fd[sym], wd[sym]
    fb, wb = fk * f_scale * 1024, wk * w_scale * 1024
    with open(TRAFFIC_FILE.format(kernel, PROFILE_TAG), "w") as f:
        f.write(f"# HBM traffic per launch of {sym} ({PROFILE_WORKLOADS[PROFILE_TAG]})\n")
        f.write(f"# command: {command}\n")
        f.write(f"# calibration: {copy_name} of {copy_kib} KiB reads FETCH_SIZE {cf[copy_name]:.1f} KiB, "
                f"WRITE_SIZE {cw[copy_name]:.1f} KiB per dispatch\n")
        f.write(f"fetch_scale = {f_scale:.5f}\nwrite_scale = {w_scale:.5f}\n")
        f.write(f"fetch_kib_per_dispatch_raw = {fk:.1f}\nwrite_kib_per_dispatch_raw = {wk:.1f}\n")
        f.write(f"fetch_bytes_per_launch = {fb:.0f}\nwrite_bytes_per_launch = {wb:.0f}\n")
        f.write(f"traffic_bytes_per_launch = {fb + wb:.0f}\n")
    print(open(TRAFFIC_FILE.format(kernel, PROFILE_TAG)).read())
    return 0


def hbm_traffic(kernel: str, tag: str = "caustic512"):
    """HBM bytes per launch of `kernel` from its committed counter summary,
    or None when there is none."""
    path = TRAFFIC_FILE.format(kernel, tag)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = float(v)
    return vals.get("traffic_bytes_per_launch")


ISSUE_FILE = os.path.join(ROOT, "profiles", "issue_{}_{}.txt")
ISSUE_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_SMEM", "SQ_INSTS_VMEM", "SQ_INSTS_LDS")


def summarize_issue(pmc_dir: str, kernel: str, command: str) -> int:
    """Writes the instructions issued per launch of one kernel (wave-level
    counts, summed over the chip) from one rocprofv3 --pmc run of the SQ
    instruction counters in pmc_dir (issue_counter_collection.csv)."""
    path = os.path.join(pmc_dir, "issue_counter_collection.csv")
    sym = _symbol(kernel, _per_dispatch(path, "SQ_INSTS_VALU"))
    with open(ISSUE_FILE.format(kernel, PROFILE_TAG), "w") as f:
        f.write(f"# instructions issued per launch of {sym} ({PROFILE_WORKLOADS[PROFILE_TAG]}), "
                "wave-level counts summed over the chip\n")
        f.write(f"# command: {command}\n")
        for c in ISSUE_COUNTERS:
            per = _per_dispatch(path, c)
            if sym in per:
                f.write(f"{c.lower()}_per_launch = {per[sym]:.1f}\n")
    print(open(ISSUE_FILE.format(kernel, PROFILE_TAG)).read())
    return 0


VALU_PEAK_WAVE_INSTS = 256 * 2.4e9  # MI355X: 256 CUs x 2.4 GHz, one wave64 VALU instruction per CU and clock


def issue_rate(kernel: str, cnt: dict, frames_in_flight: int, batch_ms: float | None = None,
               launches: int | None = None, tag: str = "caustic512"):
    """VALU and SALU instructions per BVH8 node visit of `kernel`: the
    committed per-launch instruction counts (rocprofv3 SQ counters at this
    workload and frames_in_flight) times the launches of one batch, over the
    node visits of one batch from the in-kernel counters. A wave instruction
    serves 64 lanes, so 64 x VALU per lane-visit is the issue cost of one
    useful node visit; lockstep idling inflates it."""
    path = ISSUE_FILE.format(kernel, tag)
    idx = {"extend": 1, "shadow": 2, "photon_trace": 0}.get(kernel)
    if not os.path.exists(path) or idx is None:
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = float(v)
    if launches is None:
        launches = {"extend": 12, "shadow": 12, "photon_trace": 5}[kernel]
    nodes = cnt["kernel_nodes"][idx]
    if not nodes or "sq_insts_valu_per_launch" not in vals:
        return None
    out = {"source": os.path.relpath(path, ROOT), "launches_per_batch": launches,
           "node_visits_per_batch": nodes, "frames_in_flight": frames_in_flight}
    for c in ("valu", "salu", "vmem", "lds"):
        k = f"sq_insts_{c}_per_launch"
        if k in vals:
            out[f"{c}_wave_insts_per_lane_node_visit"] = round(vals[k] * launches * 64 / nodes, 1)
    if idx == 1 and cnt.get("extend_wave_nodes"):
        out["lockstep_ratio"] = round(cnt["extend_wave_nodes"] / nodes, 3)
    if batch_ms and "sq_insts_valu_per_launch" in vals:
        # the kernel's VALU issue rate against the chip's (unpacked fp32
        # arithmetic: 64 lanes per CU and clock)
        rate = vals["sq_insts_valu_per_launch"] * launches / (batch_ms * 1e-3)
        out["valu_wave_insts_per_s"] = round(rate, 1)
        out["valu_issue_peak"] = VALU_PEAK_WAVE_INSTS
        out["valu_issue_frac"] = round(rate / VALU_PEAK_WAVE_INSTS, 4)
    return out


def summarize_kernels(trace_dir: str, first_batch: int, n_batches: int, frames_per_batch: int, command: str) -> int:
    """Writes the rocprofv3 kernel statistics of a bench run (kernel_stats.csv)
    and the per-kernel times of batches [first_batch, first_batch + n_batches)
    of the renderer (kernel_trace.csv; a batch runs from one k_clear_queues
    to the next), with their span from the first kernel's start to the last
    kernel's end."""
    import csv
    import collections

    stats = os.path.join(trace_dir, "bench_kernel_stats.csv")
    trace = os.path.join(trace_dir, "bench_kernel_trace.csv")
    out = os.path.join(ROOT, "profiles", f"kernels_{PROFILE_TAG}.txt")
    rows = []
    with open(trace) as f:
        for row in csv.DictReader(f):
            rows.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"]), _short(row["Kernel_Name"])))
    rows.sort()
    if first_batch < 0:  # the statistics only (a run of many renderers)
        rows = []
    starts = [i for i, r in enumerate(rows) if r[2] == "k_clear_queues"]
    per = collections.defaultdict(list)
    if rows:
        lo = starts[first_batch]
        hi = starts[first_batch + n_batches] if first_batch + n_batches < len(starts) else len(rows)
        sel = rows[lo:hi]
        for a, b, k in sel:
            per[k].append((b - a) / 1000.0)
        span = (max(b for _, b, _ in sel) - sel[0][0]) / 1000.0
    with open(out, "w") as f:
        f.write(f"# rocprofv3 kernel statistics, {PROFILE_WORKLOADS[PROFILE_TAG]}\n")
        f.write(f"# command: {command}\n")
        with open(stats) as g:
            rd = list(csv.DictReader(g))
        f.write(f"{'kernel':34s} {'calls':>7s} {'total_us':>11s} {'avg_us':>9s} {'pct':>6s}\n")
        for row in rd:
            f.write(f"{_short(row['Name'])[:34]:34s} {int(row['Calls']):7d} {float(row['TotalDurationNs']) / 1e3:11.1f} "
                    f"{float(row['AverageNs']) / 1e3:9.2f} {float(row['Percentage']):6.2f}\n")
        if rows:
            f.write(f"\n# batches {first_batch}..{first_batch + n_batches - 1} of the run (the timed region), "
                    "per kernel:\n")
            for k, ts in sorted(per.items(), key=lambda kv: -sum(kv[1])):
                f.write(f"#   {k[:28]:28s} launches {len(ts):4d}  total {sum(ts):10.1f} us  "
                        f"avg {sum(ts) / len(ts):9.2f} us\n")
            busy = sum(sum(ts) for ts in per.values())
            frames = n_batches * frames_per_batch
            f.write(f"#   span of the batches {span:.1f} us ({span / frames:.2f} us per frame under the tracer)\n")
            f.write(f"#   kernel time of the batches {busy:.1f} us ({busy / frames:.2f} us per frame, "
                    f"{frames} frames)\n")
    print(open(out).read())
    return 0


def copy_calibration(mib: int) -> int:
    """A plain device-to-device copy of `mib` MiB, launched 4 times: the
    hardware counters of one such dispatch read `mib` MiB fetched and `mib`
    MiB written when they count bytes one to one."""
    import torch

    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "launches": 4}))
    return 0


def time_to_relmse(frames: int = 64, warmup: int = 512, boxes: int | None = None, width: int = 1024,
                   ref_frames: int = 2048, frames_in_flight: int = 64, rate_frames_in_flight=(64, 128),
                   guided_runs: int = 3) -> dict:
    """Guided vs plain tracer on the occluded-light room (2,097,182 triangles
    at the default box count), 1024 x 1024, depth 8.

    relMSE = mean over pixels and rgb of (I - R)^2 / (R^2 + 1e-2). Both
    estimators are unbiased for the same image, so relMSE after N samples
    per pixel is C / N with C = N * E[(A - B)^2 / (R^2 + 1e-2)] / 2 for two
    independent N-spp images A, B of the same tracer. R, the reference of
    the denominator, is a plain image of `ref_frames` samples per pixel with
    seeds of its own, independent of A and B: a reference that contained A
    or B would grow with their outliers and hide them. C is measured with
    `frames_in_flight` frames per wavefront.

    Time-to-relMSE 0.01 is (C / 0.01) * (seconds per spp) plus, for the
    guided tracer, the `warmup` frames of online training it runs before
    being measured (after them the network stays fixed:
    guide_train_frames = warmup; training on kept C unchanged and cost 7 %
    more per sample). Seconds per spp is
    each tracer's steady-state rate at its faster setting among
    `rate_frames_in_flight`: before any timed window the plain tracer
    renders one untimed batch (graph capture and first touches of its
    buffers are one-time costs), and the guided tracer its training frames.
    Every timed window is whole batches. The guided tracer's C is the mean
    over `guided_runs` independently trained networks (training is
    stochastic). The guided tracer plays no Russian roulette (the plain one
    keeps the reference's); a third tracer, the plain one without roulette,
    is measured as a control."""
    import numpy as np
    import scenes
    import wsrt

    dev = wsrt.Device(0)
    nb = scenes.OCCLUDED_BOXES if boxes is None else boxes
    scene = scenes.build_occluded_room(wsrt.SceneBuilder(), n_boxes=nb).build(dev)
    v = scenes.OCCLUDED_VIEW
    view = scenes.View(v.position, v.yaw, v.pitch, width, width)

    def frames_of(r, n):
        """Mean image of the next n frames and the seconds they took."""
        a0 = r.read_accum()[:, :3].astype(np.float64)
        r.sync()
        t = time.perf_counter()
        r.render_frames(n)
        r.sync()
        dt = time.perf_counter() - t
        a1 = r.read_accum()[:, :3].astype(np.float64)
        return (a1 - a0) / n, dt

    def renderer(cfg, first):
        if EAGER:
            cfg.frame_graphs = 0
        r = wsrt.Renderer(dev, scene, width, width, cfg)
        r.set_camera(scenes.camera(view, first))
        r.frame_count = first
        return r

    def warm(r, name, S):
        """Untimed steady state for the plain tracer; timed training for the
        guided one, then one untimed batch that captures the graph of its
        frames with the network fixed (a one-time cost like the plain
        tracer's capture)."""
        _, t = frames_of(r, warmup if name == "guided" else S)
        if name == "guided":
            frames_of(r, S)
        return t if name == "guided" else 0.0

    out = {"scene": f"occluded-light room, {scene.info.n_triangles} triangles, {width}x{width}, depth 8",
           "relmse": "mean (I - R)^2 / (R^2 + 1e-2)", "target": 0.01}
    S = frames_in_flight
    out["reference_spp"] = ref_frames
    out["frames_in_flight_for_variance"] = S
    out["measured_spp"] = frames

    def pair(quirks):
        """Plain and guided tracer with the given emission weighting."""
        rp = renderer(scenes.occluded_config(frames_in_flight=128, quirks=quirks), 1 << 22)
        R, _ = frames_of(rp, ref_frames)
        rp.close()
        res = {}

        def plain(**kw):
            return scenes.occluded_config(quirks=quirks, **kw)

        def guided(**kw):
            return scenes.guided_config(guide_train_frames=warmup, quirks=quirks, **kw)

        def plain_no_roulette(**kw):
            # a control: the plain tracer without Russian roulette, as the
            # guided tracer runs (rr_after past the last bounce)
            return scenes.occluded_config(quirks=quirks, rr_after=255, **kw)

        for name, make in (("plain", plain), ("guided", guided), ("plain_no_roulette", plain_no_roulette)):
            cfg = make(frames_in_flight=S)
            r = renderer(cfg, 0)
            tw = {S: warm(r, name, S)}
            a, ta = frames_of(r, frames)
            b, tb = frames_of(r, frames)
            r.close()
            if name == "guided" and quirks == 0:  # the network's rates with training on, one batch after a warm one
                rm = renderer(scenes.guided_config(frames_in_flight=S), 0)
                rm.render_frames(S)
                out["mlp"] = mlp_rates(rm, S, cfg.guide_train_samples)
                rm.close()
            c = frames * float(np.mean((a - b) ** 2 / (R**2 + 1e-2))) / 2
            c_runs = [c]
            if name == "guided":
                # online training is stochastic (the records each network
                # version produces train the next), so C is the mean over
                # independently trained networks, each from frames seeded
                # apart from the others and from R
                for k in range(1, guided_runs):
                    r = renderer(cfg, k * 100000)
                    warm(r, name, S)
                    a, _ = frames_of(r, frames)
                    b, _ = frames_of(r, frames)
                    r.close()
                    c_runs.append(frames * float(np.mean((a - b) ** 2 / (R**2 + 1e-2))) / 2)
                c = float(np.mean(c_runs))
            rate = {S: (ta + tb) / (2 * frames)}
            for S2 in rate_frames_in_flight:
                if S2 == S:
                    continue
                r = renderer(make(frames_in_flight=S2), 0)
                tw[S2] = warm(r, name, S2)
                _, t2 = frames_of(r, S2)
                rate[S2] = t2 / S2
                r.close()
            best = min(rate, key=rate.get)
            res[name] = {"C": round(c, 3), "C_by_run": [round(x, 3) for x in c_runs],
                         "ms_per_spp": round(1000 * rate[best], 3), "frames_in_flight": best,
                         "ms_per_spp_by_frames_in_flight": {str(k): round(1000 * x, 3)
                                                            for k, x in sorted(rate.items())},
                         "relmse_at_measured_spp": round(c / frames, 4),
                         "time_to_relmse_s": round(tw[best] + c / 0.01 * rate[best], 2),
                         "training_frames_before_measuring": warmup if name == "guided" else 0}
        res["speedup_guided_vs_plain"] = round(res["plain"]["time_to_relmse_s"] / res["guided"]["time_to_relmse_s"],
                                               3)
        res["speedup_guided_vs_plain_no_roulette"] = round(
            res["plain_no_roulette"]["time_to_relmse_s"] / res["guided"]["time_to_relmse_s"], 3)
        return res

    # the reference's estimator (emission reached by a Lambert bounce counted
    # whole beside next-event estimation), then both tracers with the
    # balance heuristic (WSRT_FIX_EMISSION_MIS; the guided tracer weighs
    # light samples against its mixture pdf)
    out.update(pair(0))
    out["emission_mis"] = pair(wsrt.FIX_EMISSION_MIS)
    return out


# Algorithmic work of the guide network per sample: 168 encoded inputs, four
# hidden layers of 64, 40 outputs; forward 2 (168*64 + 3*64^2 + 64*40) FLOP,
# training (forward, backward, weight gradients) three times that.
MLP_FWD_FLOP = 2 * (168 * 64 + 3 * 64 * 64 + 64 * 40)
MLP_TRAIN_FLOP = 3 * MLP_FWD_FLOP
BF16_PEAK_TFLOPS = 2500.0


def mlp_rates(r, frames: int, train_samples: int) -> dict:
    """BF16 MFMA rate of the guide network inside the guided tracer, over
    `frames` frames with per-kernel HIP events: inference in k_guide_scatter
    (one evaluation per Lambert vertex queued to it, counted on the device)
    and the training step (preparation, forward, backward, weight gradients,
    Adam) on `train_samples` records per frame. The scatter kernel also
    encodes inputs, samples the mixture and queues the next rays, so its rate
    understates the MFMA work alone."""
    s0 = r.stats()["guide_samples"]
    r.sync()
    r.reset_profile()
    r.set_profiling(True)
    r.render_frames(frames)
    r.sync()
    r.set_profiling(False)
    prof = r.profile()
    n_inf = r.stats()["guide_samples"] - s0
    n_train = frames * train_samples
    out = {"unit": "TFLOP/s", "peak": BF16_PEAK_TFLOPS, "frames": frames,
           "flop_per_sample": {"inference": MLP_FWD_FLOP, "training": MLP_TRAIN_FLOP}}
    for key, kname, n, flop in (("inference", "guide_scatter", n_inf, MLP_FWD_FLOP),
                                ("training", "guide_train", n_train, MLP_TRAIN_FLOP)):
        ms = prof.get(kname, {}).get("total_ms", 0.0)
        tf = n * flop / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        out[key] = {"samples": int(n), "ms": round(ms, 3), "achieved": round(tf, 2),
                    "frac": round(tf / BF16_PEAK_TFLOPS, 4)}
    return out


def guide_train_steps(n: int, samples: int = 16384) -> int:
    """n Adam steps of the standalone guide network on `samples` seeded
    records (scalars in [0, 1), unit directions, weights in [0, 1)), through
    wsrt_guide_train_step: the same four kernels as one training step of the
    renderer. Prints the mean time per step (host buffers uploaded each
    step, so this is an upper bound; a kernel trace gives the kernels'
    own times)."""
    import numpy as np
    import wsrt

    dev = wsrt.Device(0)
    g = wsrt.Guide(dev, samples, 7)
    rng = np.random.default_rng(0)
    x = rng.random((samples, 7), dtype=np.float32)
    d = rng.normal(size=(samples, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(samples, dtype=np.float32)
    for _ in range(5):
        g.train_step(x, d, w)
    t = time.perf_counter()
    for _ in range(n):
        g.train_step(x, d, w)
    dt = (time.perf_counter() - t) / n
    print(json.dumps({"guide_train_steps": n, "samples_per_step": samples,
                      "ms_per_step_with_upload": round(dt * 1e3, 4)}))
    return 0


def room_throughput(batches: int = 2, warmup: int = 2) -> dict:
    """The plain tracer on the occluded-light room (2,097,182 triangles,
    1024 x 1024, depth 8, its shipped configuration: 128 frames in flight,
    lane refill, path-ordered queues): `batches` batches timed as graph
    replays after `warmup` untimed ones, then one batch with per-kernel HIP
    events and one with the traversal counters."""
    import scenes
    import wsrt

    dev = wsrt.Device(0)
    scene = scenes.build_occluded_room(wsrt.SceneBuilder()).build(dev)
    v = scenes.OCCLUDED_VIEW
    W = 1024
    view = scenes.View(v.position, v.yaw, v.pitch, W, W)
    cfg = scenes.occluded_config(frames_in_flight=128)
    S = cfg.frames_in_flight
    r = wsrt.Renderer(dev, scene, W, W, cfg)
    r.set_camera(scenes.camera(view, 0))
    r.render_frames(S * max(warmup, 1))
    r.sync()
    r.reset_stats()
    t0 = time.perf_counter()
    r.render_frames(S * batches)
    r.sync()
    dt = time.perf_counter() - t0
    st = r.stats()
    r.reset_stats()
    r.reset_profile()
    r.set_profiling(True)
    r.render_frames(S)
    r.sync()
    r.set_profiling(False)
    pst, prof = r.stats(), r.profile()
    r.reset_stats()
    r.set_traversal_counting(True)
    r.render_frames(S)
    r.sync()
    cnt = r.stats()
    r.close()
    frames = S * batches
    kern = {}
    for name, idx in (("extend", 1), ("shadow", 2)):
        p = prof[name]
        counted = max(cnt["kernel_counted_rays"][idx], 1)
        nodes, tris = cnt["kernel_nodes"][idx] / counted, cnt["kernel_tris"][idx] / counted
        bpr = RAY_BYTES + nodes * NODE_BYTES + tris * TRI_BYTES
        rays = pst["kernel_rays"][idx]
        gbs = rays * bpr / (p["total_ms"] * 1e-3) / 1e9
        kern[name] = {"ms": round(p["total_ms"], 3), "launches": p["launches"], "rays": rays,
                      "rays_per_s": round(rays / (p["total_ms"] * 1e-3), 1),
                      "nodes_per_ray": round(nodes, 3), "tris_per_ray": round(tris, 3),
                      "max_nodes_of_a_ray": cnt["ray_max_nodes"][idx], "bytes_per_ray": round(bpr, 1),
                      "GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        traffic = hbm_traffic(name, "room1024")
        if traffic:
            kern[name]["traffic_bytes_per_launch"] = traffic
        iss = issue_rate(name, cnt, S, p["total_ms"], launches=p["launches"], tag="room1024")
        if iss:
            kern[name]["issue"] = iss
    for name in ("shade", "generate", "accumulate", "order", "tail"):
        if name in prof and prof[name]["launches"]:
            kern[name] = {"ms": round(prof[name]["total_ms"], 3), "launches": prof[name]["launches"]}
    return {"workload": "occluded-light room, 1024x1024, depth 8, plain tracer, 128 frames in flight, "
                        "lane refill 16, path-ordered queues",
            "triangles": scene.info.n_triangles, "frames": frames,
            "Msamples_per_s": round(W * W * frames / dt / 1e6, 3), "ms_per_spp": round(1000 * dt / frames, 3),
            "rays_per_s": round((st["extend_rays"] + st["shadow_rays"]) / dt, 1),
            "event_timed_batch_ms": round(sum(p["total_ms"] for p in prof.values()), 3), "kernels": kern}


def heightfield_throughput(frames: int = 16) -> dict:
    """Film-coated GGX heightfield (2,093,058 triangles) under the seeded sky,
    1920 x 1080, depth 6, `frames` samples per pixel, plain tracer."""
    import scenes
    import wsrt

    dev = wsrt.Device(0)
    t = time.perf_counter()
    scene = scenes.build_heightfield(wsrt.SceneBuilder()).build(dev)
    host_s = time.perf_counter() - t
    v = scenes.HEIGHTFIELD_VIEW
    cfg = scenes.heightfield_config()
    tune = wsrt.autotune(dev, scene, v.width, v.height, cfg, scenes.camera(v, 0))
    r = wsrt.Renderer(dev, scene, v.width, v.height, cfg)
    r.set_camera(scenes.camera(v, 0))
    r.render_frames(cfg.frames_in_flight)  # captures the batch graph
    r.sync()
    r.frame_count = 0
    t = time.perf_counter()
    r.render_frames(frames)
    r.sync()
    dt = time.perf_counter() - t
    r.close()
    return {"workload": "heightfield 1024x1024 vertices (film-coated GGX tiles, sky environment), 1920x1080, "
                        f"{frames} spp, depth 6", "triangles": scene.info.n_triangles,
            "Msamples_per_s": round(v.width * v.height * frames / dt / 1e6, 3),
            "ms_per_frame": round(1000 * dt / frames, 3), "frames_in_flight": cfg.frames_in_flight,
            "autotune": tune, "scene_build_s": round(host_s, 2), "bvh_build_ms": round(scene.info.build_ms, 2)}


def small_box_throughput(frames: int = 16) -> dict:
    """34-triangle Cornell-style box (5 walls, two seeded boxes, one GGX),
    0.25 x 0.25 area light, 256 x 256, 16 spp, depth 8, photon pass on."""
    import scenes
    import wsrt

    dev = wsrt.Device(0)
    scene = scenes.build_small_box(wsrt.SceneBuilder()).build(dev)
    v = scenes.SMALL_BOX_VIEW
    cfg = scenes.small_box_config()
    tune = wsrt.autotune(dev, scene, v.width, v.height, cfg, scenes.camera(v, 0))
    r = wsrt.Renderer(dev, scene, v.width, v.height, cfg)
    r.set_camera(scenes.camera(v, 0))
    r.render_frames(cfg.frames_in_flight)
    r.sync()
    r.frame_count = 0
    t = time.perf_counter()
    r.render_frames(frames)
    r.sync()
    dt = time.perf_counter() - t
    r.close()
    return {"workload": "Cornell-style box, 34 triangles + area light, 256x256, 16 spp, depth 8, 2^20 photons/frame",
            "triangles": scene.info.n_triangles, "Msamples_per_s": round(v.width * v.height * frames / dt / 1e6, 3),
            "ms_per_frame": round(1000 * dt / frames, 3), "frames_in_flight": cfg.frames_in_flight,
            "autotune_best": tune["best"]}


def batch_throughput(n_scenes: int = 64) -> dict:
    """64 procedurally generated scenes of uneven size (triangles log-uniform
    1e3 .. 2e6, 128^2 .. 1024^2 pixels, 8 .. 64 spp) rendered as one batch,
    without and with the photon pass."""
    import scenes
    import wsrt

    dev = wsrt.Device(0)
    jobs, tris, samples = [], 0, 0
    t = time.perf_counter()
    for k in range(n_scenes):
        b, v, spp = scenes.build_batch_scene(wsrt.SceneBuilder(), k)
        sc = b.build(dev)
        jobs.append((sc, scenes.camera(v, 0), v.width, v.height, spp))
        tris += sc.info.n_triangles
        samples += v.width * v.height * spp
    host_s = time.perf_counter() - t
    out = {"workload": f"{n_scenes} procedural scenes (displaced spheres and boxes, mixed materials), depth 8",
           "scenes": n_scenes, "triangles": tris, "samples": samples, "scene_build_s": round(host_s, 2)}
    for photons in (0, 1):
        # without the photon pass: every sample of every job in one persistent
        # launch; with it (the reference's frame: photon pass, grid, camera
        # paths): per window of frames every job's passes of 2^20 photon
        # paths per frame and their grids, then one persistent launch over
        # the window's frames, one captured graph per window
        cfg = wsrt.default_config(enable_photons=photons, max_depth=8)
        batch = wsrt.Batch(dev, jobs, cfg)
        tune = batch.autotune()
        batch.render()  # the first call captures the graph(s)
        batch.sync()
        t = time.perf_counter()
        batch.render()
        batch.sync()
        dt = time.perf_counter() - t
        batch.close()
        out["photons" if photons else "no_photons"] = {
            "Msamples_per_s": round(samples / dt / 1e6, 3), "seconds": round(dt, 3), "tuning": tune,
            "photon_paths": (2**20 * sum(j[4] for j in jobs)) if photons else 0}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed 64-spp images")
    ap.add_argument("--warmup", type=int, default=4, help="untimed 64-spp images")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided-vs-plain time-to-relMSE leg")
    ap.add_argument("--no-other", action="store_true", help="skip the heightfield and batch throughput legs")
    ap.add_argument("--guided-only", action="store_true", help="only the guided-vs-plain leg (one JSON line)")
    ap.add_argument("--guided-frames", type=int, default=64)
    ap.add_argument("--eager", action="store_true",
                    help="guided leg with every batch launched kernel by kernel instead of as one captured "
                         "HIP graph (for kernel traces of the guided tracer; kernel durations are the same)")
    ap.add_argument("--frames-in-flight", type=int, default=0, help="frames per wavefront (default: config's)")
    ap.add_argument("--extend-refill", type=int, default=-1,
                    help="with --frames-in-flight: the closest-hit kernel's lane refill threshold "
                         "(default: the configuration's)")
    ap.add_argument("--guided-boxes", type=int, default=None)
    ap.add_argument("--guided-width", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (the reference point that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run) and exit")
    ap.add_argument("--guide-train-steps", type=int, default=0, metavar="N",
                    help="only run N Adam steps of the guide network over 16384 synthetic samples "
                         "(the renderer's step size), for a kernel profile of the training kernels, "
                         "and exit")
    ap.add_argument("--summarize-traffic", nargs=3, metavar=("PMC_DIR", "KERNEL", "COMMAND"),
                    help="write the per-launch HBM traffic summary of KERNEL from rocprofv3 --pmc "
                         "output and exit")
    ap.add_argument("--summarize-issue", nargs=3, metavar=("PMC_DIR", "KERNEL", "COMMAND"),
                    help="write the per-launch instruction counts of KERNEL from a rocprofv3 --pmc run "
                         "of the SQ instruction counters and exit")
    ap.add_argument("--profile-workload", choices=sorted(PROFILE_WORKLOADS), default="caustic512",
                    help="the workload a --summarize-* summary is written for")
    ap.add_argument("--room-only", action="store_true",
                    help="only the plain tracer on the occluded-light room (one JSON line; --steps batches "
                         "of 128 frames timed after --warmup)")
    ap.add_argument("--summarize-kernels", nargs=5,
                    metavar=("TRACE_DIR", "FIRST_BATCH", "N_BATCHES", "FRAMES_PER_BATCH", "COMMAND"),
                    help="write the kernel statistics and the timed batches' per-kernel times from a "
                         "rocprofv3 --kernel-trace --stats run and exit")
    args = ap.parse_args()
    global PROFILE_TAG, EAGER
    PROFILE_TAG = args.profile_workload
    EAGER = args.eager
    if args.summarize_traffic:
        return summarize_traffic(*args.summarize_traffic)
    if args.summarize_issue:
        return summarize_issue(*args.summarize_issue)
    if args.summarize_kernels:
        d, a, n, fpb, c = args.summarize_kernels
        return summarize_kernels(d, int(a), int(n), int(fpb), c)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.guide_train_steps:
        return guide_train_steps(args.guide_train_steps)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the render loop runs "
              "on a single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2

    if args.guided_only:
        print(json.dumps(time_to_relmse(args.guided_frames, boxes=args.guided_boxes, width=args.guided_width)))
        return 0
    if args.room_only:
        print(json.dumps(room_throughput(args.steps, args.warmup)))
        return 0

    import scenes
    import wsrt

    dev = wsrt.Device(0)
    scene = scenes.build_caustic(wsrt.SceneBuilder()).build(dev)
    info = scene.info
    v = scenes.CAUSTIC_VIEW
    cfg = scenes.caustic_config()
    tune = None
    if args.frames_in_flight:
        cfg.frames_in_flight = args.frames_in_flight
        if args.extend_refill >= 0:
            cfg.extend_refill = args.extend_refill
    else:
        # frames per wavefront picked by the library's autotuner (results do
        # not depend on it); the timed frames run at exactly that setting
        tune = wsrt.autotune(dev, scene, v.width, v.height, cfg, scenes.camera(v, 0))
        log(f"autotune: {tune}")
    S = cfg.frames_in_flight
    frames = args.steps * STEP_FRAMES
    r = wsrt.Renderer(dev, scene, v.width, v.height, cfg)
    r.set_camera(scenes.camera(v, 0))
    # warm-up: at least one full batch, which captures its graph, and the
    # batch size the timed frames end with, so no capture falls in the
    # timed region
    r.render_frames(max(args.warmup * STEP_FRAMES, S))
    if frames % S:
        r.render_frames(frames % S)
    r.sync()
    r.reset_stats()
    # timed region: exactly `steps` images of STEP_FRAMES frames, each batch
    # of frames_in_flight frames one replay of its captured HIP graph, as the
    # library ships
    t0 = time.perf_counter()
    r.render_frames(frames)
    r.sync()
    dt = time.perf_counter() - t0
    st = r.stats()
    # per-kernel device time: one batch of the same frames_in_flight, each
    # kernel (group) bracketed by HIP events on the renderer's stream, the
    # stream the kernels run on (profiling launches the batch eagerly)
    r.reset_stats()
    r.reset_profile()
    r.set_profiling(True)
    r.render_frames(S)
    r.sync()
    r.set_profiling(False)
    pst = r.stats()
    prof = r.profile()
    # traversal work per ray of each launching kernel from the in-kernel
    # counters (pinned against the oracle's traversal of the same BVH8),
    # measured on one more batch
    r.reset_stats()
    r.set_traversal_counting(True)
    r.render_frames(S)
    r.sync()
    cnt = r.stats()
    r.set_traversal_counting(False)

    samples = v.width * v.height * frames
    msps = samples / dt / 1e6
    rays = st["extend_rays"] + st["shadow_rays"] + st["kernel_rays"][0]
    # the launching kernels' counter indices (wsrt_frame_stats.kernel_*)
    kinds = {"photon_trace": [0], "extend": [1], "shadow": [2], "tail": [4, 5]}
    per_ray = {}
    ray_counts, ray_bytes = {}, {}
    for k, idx in kinds.items():
        counted = sum(cnt["kernel_counted_rays"][i] for i in idx)
        nodes = sum(cnt["kernel_nodes"][i] for i in idx)
        tris = sum(cnt["kernel_tris"][i] for i in idx)
        per_ray[k] = {"nodes": nodes / max(counted, 1), "tris": tris / max(counted, 1)}
        ray_counts[k] = sum(pst["kernel_rays"][i] for i in idx)
        # bytes of the profiled batch's rays: each counter's rays at its own
        # work per ray
        ray_bytes[k] = sum(pst["kernel_rays"][i] * (RAY_BYTES + (cnt["kernel_nodes"][i] * NODE_BYTES +
                                                                 cnt["kernel_tris"][i] * TRI_BYTES) /
                                                    max(cnt["kernel_counted_rays"][i], 1)) for i in idx)
    # emission: one stored-bit word per 32 paths, and for each path that
    # reaches the slab test (the cone rejects the others with one dot
    # product on registers and LDS) the root node read and its 44-byte
    # queued state
    n_emit = pst["kernel_rays"][3]
    cand = cnt["kernel_nodes"][3] / max(cnt["kernel_counted_rays"][3], 1)
    per_ray["photon_emit"] = {"slab_tested_fraction": cand}
    emit_bytes = n_emit * (4.0 / 32 + cand * (NODE_BYTES + 44))
    stored_batch = pst["photons_stored"]
    kern = {}
    over_peak = []
    for name, p in prof.items():
        if p["launches"] == 0:
            continue
        e = {"ms": round(p["total_ms"], 3), "launches": p["launches"],
             "avg_us": round(1000 * p["total_ms"] / p["launches"], 2)}
        gbs = None
        if name in per_ray and name in ray_counts and ray_counts[name]:
            e["rays"] = ray_counts[name]
            e["bytes_per_ray"] = round(ray_bytes[name] / ray_counts[name], 1)
            gbs = ray_bytes[name] / (p["total_ms"] * 1e-3) / 1e9
        elif name == "photon_emit":
            e["paths"] = n_emit
            gbs = emit_bytes / (p["total_ms"] * 1e-3) / 1e9
        elif name == "grid_build":
            # the stored-bit scan (one bit per photon slot), then per stored
            # photon the 60-byte hash, histogram, scan and scatter work and
            # 12 bytes per touched cell (at most one per stored photon)
            e["photons_stored"] = stored_batch
            gbs = (cfg.photons_per_pass * S / 8 + 72 * stored_batch) / (p["total_ms"] * 1e-3) / 1e9
        if gbs is not None:
            e["GBps"] = round(gbs, 1)
            if gbs > HBM_PEAK_GBS:
                over_peak.append(name)
        kern[name] = e
    if over_peak:
        raise AssertionError(f"byte model of {over_peak} exceeds the HBM peak: fix the model")
    dom = max((k for k in kern if "rays" in kern[k]), key=lambda k: kern[k]["ms"])
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
                        f"light, 512x512, {STEP_FRAMES} spp per step (one frame per spp: photon pass of "
                        "2^20 paths, 2^22-cell grid, one camera path per pixel), cell=radius=0.01, "
                        "max depth 12",
            "triangles": info.n_triangles, "bvh8_nodes": info.n_nodes,
            "width": v.width, "height": v.height, "spp_per_step": STEP_FRAMES,
            "photons_per_pass": cfg.photons_per_pass,
            "hash_size": cfg.hash_size, "max_depth": cfg.max_depth,
            "frames_in_flight": S, "timed_frames": frames,
            "launch": "one captured HIP graph per batch of frames_in_flight frames",
        },
        "autotune": tune,
        "ms_per_frame": round(1000 * dt / frames, 4),
        "rays_per_s": round(rays / dt, 1),
        "photon_paths_per_s": round(st["photon_paths"] / dt, 1),
        "photon_rays_traced_per_s": round(st["kernel_rays"][0] / dt, 1),
        "photons_stored_per_frame": round(st["photons_stored"] / frames, 1),
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic(dom),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_launch": round(d["bytes_per_ray"] * ray_counts[dom] / d["launches"]),
            "algorithmic_bytes_per_ray": d["bytes_per_ray"],
            "avg_launch_us": d["avg_us"],
            "issue": issue_rate(dom, cnt, S, d["ms"]),
        },
        "event_timed_batch_ms": round(sum(p["total_ms"] for p in prof.values()), 3),
        "kernels": kern,
        "traversal_per_ray": per_ray,
    }
    if not args.no_other:
        log("heightfield ...")
        r.close()
        result["other_configs"] = {"small_box": small_box_throughput(), "heightfield": heightfield_throughput()}
        log("batch ...")
        result["other_configs"]["batch"] = batch_throughput()
        log("room ...")
        result["other_configs"]["room"] = room_throughput()
    if not args.no_guided:
        log("guided vs plain time-to-relMSE ...")
        r.close()
        result["time_to_relmse"] = time_to_relmse(args.guided_frames, boxes=args.guided_boxes,
                                                  width=args.guided_width)
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
