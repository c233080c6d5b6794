"""The allocation-free Li (nori_gpu_li_noalloc) against nori_gpu_li through the
C ABI: two calls in a row on the same rays and seeds, the second with fewer
rays, each compared with the allocating entry point's result (same samples:
equal up to the order in which a path's two next-event contributions are
added to its film entry)."""
import ctypes
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the product library is linked against (already loaded in
    this process), for the test's own device buffers."""
    ng.load_library()
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64.so" in line)
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    cam = scene.camera
    smp = np.c_[rng.random(n) * cam.width, rng.random(n) * cam.height, rng.random(n), rng.random(n)]
    return orc.camera_rays(cam, smp.astype(np.float32)), rng.integers(0, 2 ** 63, n, dtype=np.uint64)


def test_li_noalloc_matches_li(gpu, cornell):
    hip = _hip()
    gpu.upload(cornell)
    for n, seed in ((ng.LI_NOALLOC_MAX_RAYS, 1), (5000, 2)):
        _say(f"case: li_noalloc n={n}")
        rays, seeds = _rays(cornell, n, seed)
        ref = gpu.li(rays, seeds, 8)
        d_rays, d_seeds, d_out = DevBuf(hip, rays.nbytes), DevBuf(hip, seeds.nbytes), DevBuf(hip, n * 16)
        try:
            d_rays.put(rays)
            d_seeds.put(seeds)
            d_out.put(np.full((n, 4), -1.0, np.float32))  # stale contents are overwritten
            assert hip.hipDeviceSynchronize() == 0
            gpu.li_noalloc(d_rays.p.value, d_seeds.p.value, n, 8, 3, d_out.p.value)
            out = d_out.get((n, 4), np.float32)
        finally:
            for b in (d_rays, d_seeds, d_out):
                b.free()
        assert np.all(out[:, 3] == 1.0)
        assert np.isfinite(out).all()
        assert np.allclose(out[:, :3], ref, rtol=1e-5, atol=1e-7), np.abs(out[:, :3] - ref).max()
        _say(f"  mean Li {out[:, :3].mean(0)} vs {ref.mean(0)}")


def test_li_noalloc_rejects_more_rays_than_its_scratch(gpu, cornell):
    hip = _hip()
    gpu.upload(cornell)
    small = DevBuf(hip, 256)  # never read: the size is refused first
    try:
        with pytest.raises(RuntimeError, match="65536"):
            gpu.li_noalloc(small.p.value, small.p.value, ng.LI_NOALLOC_MAX_RAYS + 1, 8, 3, small.p.value)
    finally:
        small.free()
