"""rbh_envmap_build_noalloc, the entry point that allocates nothing: its
outputs on caller device buffers equal rbh_envmap_build's bit for bit, on two
calls in a row (the second with a smaller map, so the static scratch array
still holds the first call's larger contents), in FP32 and FP64 modes."""
import ctypes

import numpy as np
import pytest

from ratbite_amd import _lib as L
from ratbite_amd.envmaps import lognormal_envmap

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime the library links."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, arr):
        assert self.rt.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.rt.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes, 2) == 0

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def _noalloc(ctx, hip, tex, fp64):
    h, w, ch = tex.shape
    tex = np.ascontiguousarray(tex, np.float32)
    dt, dc, dm, dy, di = hip.alloc(tex.nbytes), hip.alloc(w * h * 2), hip.alloc(h * 2), hip.alloc(h * 4), hip.alloc(8)
    try:
        hip.h2d(dt, tex)
        print("noalloc build %dx%dx%d fp64=%d" % (w, h, ch, fp64), flush=True)
        ctx.check(L.lib().rbh_envmap_build_noalloc(ctx.h, dt, w, h, ch, L.RBH_CDF_FP64 if fp64 else 0, dc, dm, dy, di))
        ctx.check(L.lib().rbh_synchronize(ctx.h))
        cond, marg = np.zeros((h, w), np.uint16), np.zeros(h, np.uint16)
        cdfy, integral = np.zeros(h, np.float32), np.zeros(1, np.float64)
        for a, p in ((cond, dc), (marg, dm), (cdfy, dy), (integral, di)):
            hip.d2h(a, p)
        return cond, marg, cdfy, float(integral[0])
    finally:
        for p in (dt, dc, dm, dy, di):
            hip.free(p)


@pytest.mark.parametrize("fp64", [False, True])
def test_noalloc_build_equals_the_allocating_build(ctx, fp64):
    hip = _Hip()
    maps = [lognormal_envmap(2048, 1024, sigma=2.0, seed=21, sun=True),
            lognormal_envmap(256, 128, sigma=3.0, seed=22, sun=True, channels=4)]
    for tex in maps:
        ctx.envmap_build(tex, fp64=fp64)
        ref = ctx.envmap_read()
        got = _noalloc(ctx, hip, tex, fp64)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert np.array_equal(got[2], ref[2])
        assert got[3] == ref[3], (got[3], ref[3])


def test_noalloc_rejects_maps_above_its_scratch(ctx):
    rc = L.lib().rbh_envmap_build_noalloc(ctx.h, ctypes.c_void_p(16), 8, L.RBH_ENVMAP_NOALLOC_MAX_ROWS + 4, 3, 0,
                                          ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16),
                                          ctypes.c_void_p(16))
    assert rc == L.RBH_ERR_RANGE
