"""The oracle and the library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

* The f32 oracle built with gcc -fsanitize=address,undefined (oracle/Makefile
  `sanitized`) runs tests/test_oracle.py, which drives every oracle entry
  point (camera, geometry, scenes and BVH, grid, photon pass, camera paths,
  present), with gcc's ASan runtime preloaded.
* libwsrt.so with api.cpp (the C ABI, argument validation, the scene
  builder's host side, geometry and camera) built with the same sanitizers
  on its host pass (build.build_sanitized) runs tests/test_abi.py, with the
  ASan runtime of the compiler that built it preloaded. Without a GPU the
  device calls return WSRT_E_UNSUPPORTED, so only host code runs.

Any sanitizer report fails the child run (undefined behaviour aborts rather
than recovering; leak checking is off, as the interpreter keeps objects to
the end).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wgpu-spectral-rt_amd")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def run_child(tests, env_extra, timeout=900):
    env = dict(os.environ)
    env.update(env_extra)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    res = subprocess.run([sys.executable, "-m", "pytest", *tests, "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          "-p", "no:xdist"], cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    for r in REPORTS:
        assert r not in out, out[-4000:]
    return out


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitized"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan):
        pytest.skip("gcc's ASan runtime is not installed")
    out = run_child(["tests/test_oracle.py"], {"WSRT_ORACLE_LIB": lib, "LD_PRELOAD": asan})
    assert " passed" in out


def test_host_library_under_asan_ubsan(tmp_path):
    sys.path.insert(0, PKG)
    import build

    if not os.path.exists(os.path.join(build.BUILD, "bvh.o")):
        pytest.skip("the regular build's objects are not present")
    lib = build.build_sanitized(str(tmp_path))
    out = run_child(["tests/test_abi.py"], {"WSRT_LIB": lib, "LD_PRELOAD": build.asan_runtime()})
    assert " passed" in out
