"""Builds libwsrt.so (HIP kernels + C ABI) for gfx950 in-tree.

Each translation unit is compiled to an object under _build/ in parallel,
then linked into wgpu-spectral-rt_amd/libwsrt.so. Every build recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libwsrt.so")
SOURCES = ["api.cpp", "bvh.hip", "sort.hip", "photon.hip", "path.hip", "query.hip", "guide.hip", "batch.hip",
           "present.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WSRT_ARCH", "gfx950")

FLAGS = [
    "-x", "hip", "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
    # every f32 operation rounds as written (the oracle is built the same way)
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    warn = "\n".join(l for l in res.stderr.splitlines()
                     if "packed-fp32-ops" not in l and "warning: feature flag" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    print(build())


def build_sanitized(out_dir: str) -> str:
    """libwsrt.so with its host code (api.cpp: the C ABI, validation, the
    scene builder, geometry and camera) built with AddressSanitizer and
    UndefinedBehaviorSanitizer; the other objects are the regular build's
    (run build() first). Load it with clang's ASan runtime preloaded. Device
    code is unchanged: -Xarch_host puts the sanitizers on the host pass only."""
    os.makedirs(out_dir, exist_ok=True)
    objs = [os.path.join(BUILD, os.path.splitext(src)[0] + ".o") for src in SOURCES if src != "api.cpp"]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError(f"regular objects missing ({missing[0]}): run build() first")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    api = os.path.join(out_dir, "api.o")
    flags = [f for f in FLAGS if f != "-O3"] + ["-O1", "-g"]
    res = subprocess.run([HIPCC, *flags, *san, "-c", os.path.join(CSRC, "api.cpp"), "-o", api],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"sanitized compile failed:\n{res.stderr}")
    lib = os.path.join(out_dir, "libwsrt.so")
    res = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *san, "-shared-libsan", "-o", lib,
                          api, *objs], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"sanitized link failed:\n{res.stderr}")
    return lib


def asan_runtime() -> str:
    """The ASan runtime of the compiler build_sanitized uses."""
    out = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                         text=True).stdout.strip()
    if not os.path.isabs(out):
        res = subprocess.run([HIPCC, "--print-runtime-dir"], capture_output=True, text=True).stdout.strip()
        out = os.path.join(res, "libclang_rt.asan-x86_64.so")
    return out
