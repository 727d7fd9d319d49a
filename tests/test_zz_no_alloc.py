"""wsrt_photon_grid_build_noalloc (the grid build that allocates nothing,
its work memory a file-scope device array) against wsrt_photon_grid_build,
bit for bit, on two calls in a row, the second with fewer photons and
cells; and its size limits. Device buffers come from the HIP runtime through
ctypes (hipMalloc / hipMemcpy / hipFree), so the entry point sees plain
device pointers as an FFI caller would pass them."""
import ctypes as C

import numpy as np
import pytest

import scenes
import wsrt

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 4)) == 0

    def upload(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def download(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def noalloc_grid(hip, device, photons, frame, cell, hash_size):
    n = len(photons)
    bufs = [DevArray(hip, photons.nbytes), DevArray(hip, 4 * n), DevArray(hip, 4 * hash_size),
            DevArray(hip, 4 * hash_size), DevArray(hip, 4 * n)]
    ph, cid, cnt, off, slot = bufs
    ph.upload(photons)
    rc = wsrt.load().wsrt_photon_grid_build_noalloc(device.h, ph.p, n, frame, cell, hash_size, cid.p, cnt.p,
                                                   off.p, slot.p, None)
    assert rc == 0, wsrt.load().wsrt_last_error()
    assert hip.hipDeviceSynchronize() == 0
    out = (cid.download(np.zeros(n, np.uint32)), cnt.download(np.zeros(hash_size, np.uint32)),
           off.download(np.zeros(hash_size, np.uint32)), slot.download(np.zeros(n, np.uint32)))
    for b in bufs:
        b.free()
    return out


def caustic_photons(device, n):
    g = scenes.build_caustic(wsrt.SceneBuilder()).build(device)
    r = wsrt.Renderer(device, g, 16, 16, scenes.caustic_config(photons_per_pass=n, frames_in_flight=1))
    r.render()
    ph, stored = r.read_photons()
    assert stored > 100
    return ph


def test_noalloc_grid_equals_the_allocating_build(device, hip):
    print("no-alloc grid build: 2^20 photons / 2^22 cells, then 40000 / 2^16", flush=True)
    big = caustic_photons(device, 1 << 20)
    rng = np.random.default_rng(5)
    small = np.zeros(40000, wsrt.PHOTON_DTYPE)
    small["position"] = rng.uniform(-0.3, 0.3, (40000, 3))
    small["energy"] = np.where(rng.random(40000) < 0.7, rng.uniform(0.1, 2.0, 40000), 0.0)
    for ph, frame, cell, hs in ((big, 7, 0.01, 1 << 22), (small, 3, 0.02, 1 << 16)):
        cid, cnt, off, slot = noalloc_grid(hip, device, ph, frame, cell, hs)
        rc, rcount, roff, rslot = wsrt.photon_grid_build(device, ph, frame, cell, hs)
        np.testing.assert_array_equal(cid, rc)
        np.testing.assert_array_equal(cnt, rcount)
        np.testing.assert_array_equal(off, roff)
        valid = int(rcount.sum())
        np.testing.assert_array_equal(slot[:valid], rslot)
        assert np.all(slot[valid:] == 0xFFFFFFFF)
        assert valid == int(np.count_nonzero(ph["energy"]))


def test_noalloc_grid_rejects_sizes_beyond_the_scratch(device):
    lib = wsrt.load()
    dummy = C.c_void_p(1)  # never dereferenced: the sizes are checked first
    assert lib.wsrt_photon_grid_build_noalloc(device.h, dummy, (1 << 20) + 1, 0, 0.02, 1 << 22, dummy, dummy,
                                              dummy, dummy, None) == -1  # WSRT_E_INVALID
    assert b"2^20" in lib.wsrt_last_error()
    assert lib.wsrt_photon_grid_build_noalloc(device.h, dummy, 16, 0, 0.02, 1 << 23, dummy, dummy, dummy,
                                              dummy, None) == -1  # WSRT_E_INVALID
    assert lib.wsrt_photon_grid_build_noalloc(device.h, dummy, 16, 0, 0.02, 3000, dummy, dummy, dummy,
                                              dummy, None) == -1  # WSRT_E_INVALID
