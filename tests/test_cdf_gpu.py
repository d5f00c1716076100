"""GPU parity of the fused env-map CDF build + inversion against the oracle.

FP32 mode must be bit-exact (CDF floats, uint16 indices, integral); FP64 mode
is compared with the oracle's FP64 distribution.
"""
import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import RbhError
from ratbite_amd._lib import RBH_ERR_INVALID, RBH_ERR_RANGE
from ratbite_amd.envmaps import lognormal_envmap

pytestmark = pytest.mark.gpu

# one to eight chunks of the reference's block per row (the marginal up to
# four), rows with and without the linear inversion's marks (up to 16384 and
# above), non-power-of-two sizes
SIZES = [(64, 32, 3), (256, 128, 4), (1024, 512, 3), (2048, 1024, 3), (4096, 2048, 3), (1000, 500, 3),
         (8192, 16, 4), (4, 4, 3), (16384, 8, 3), (32768, 4, 3), (12, 4096, 3)]


@pytest.mark.parametrize("w,h,ch", SIZES)
def test_cdf_bit_exact(ctx, w, h, ch):
    tex = lognormal_envmap(w, h, sigma=2.0, seed=w * 7 + h, sun=w >= 64, channels=ch)
    print("cdf parity", w, h, ch, flush=True)
    ctx.envmap_build(tex)
    cond, marg, cdfy, integral = ctx.envmap_read()
    ocx, ocy, oint = O.cdf_build(tex)
    ocond, omarg = O.cdf_invert(ocx, ocy)
    assert np.array_equal(cond, ocond)
    assert np.array_equal(marg, omarg)
    assert np.array_equal(cdfy, ocy)
    expect = np.float32(oint) / (np.float32(w) * np.float32(h))
    assert np.float32(integral) == expect


def test_cdf_zero_rows_and_all_zero(ctx):
    tex = lognormal_envmap(512, 64, sigma=1.0, seed=5, sun=False)
    tex[10] = 0.0
    tex[33, :, :] = -2.0
    ctx.envmap_build(tex)
    cond, marg, cdfy, _ = ctx.envmap_read()
    ocx, ocy, _ = O.cdf_build(tex)
    ocond, omarg = O.cdf_invert(ocx, ocy)
    assert np.array_equal(cond, ocond) and np.array_equal(marg, omarg)
    z = np.zeros((64, 128, 3), np.float32)
    ctx.envmap_build(z)
    cond, marg, cdfy, integral = ctx.envmap_read()
    assert integral == 0.0
    assert np.all(cdfy == 0)
    ocond, omarg = O.cdf_invert(*O.cdf_build(z)[:2])
    assert np.array_equal(cond, ocond) and np.array_equal(marg, omarg)


def test_cdf_rebuild_is_repeatable(ctx):
    tex = lognormal_envmap(2048, 1024, sigma=2.0, seed=11, sun=True)
    ctx.envmap_build(tex)
    first = ctx.envmap_read()
    ctx.envmap_rebuild(iterations=3)
    second = ctx.envmap_read()
    for a, b in zip(first, second):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("w,h", [(2048, 1024), (512, 256), (8192, 4096), (16384, 64)])
def test_cdf_fp64_mode(ctx, w, h):
    tex = lognormal_envmap(w, h, sigma=3.0, seed=w, sun=True)
    ctx.envmap_build(tex, fp64=True)
    cond, marg, cdfy, integral = ctx.envmap_read()
    dx, dy, dint = O.cdf_build_f64(tex)
    # each CDF value is the FP64 value rounded once
    assert np.max(np.abs(cdfy.astype(np.float64) - dy)) <= 1e-7
    # the integral stays in double through the division by w*h
    assert abs(integral - dint / (w * h)) / (dint / (w * h)) < 1e-12
    # indices of an FP64-exact inversion (ties at float rounding may move one slot)
    ocond, omarg = O.cdf_invert(dx.astype(np.float32), dy.astype(np.float32))
    diff = np.abs(cond.astype(np.int64) - ocond.astype(np.int64))
    assert diff.max() <= 1 and np.mean(diff == 0) > 0.999
    # error of the FP32 build against the FP64 distribution, stated by the test
    ctx.envmap_build(tex, fp64=False)
    _, _, cdfy32, _ = ctx.envmap_read()
    err32 = np.max(np.abs(cdfy32.astype(np.float64) - dy))
    print("%dx%d: FP32 marginal CDF max abs error vs FP64: %.3g" % (w, h, err32))
    assert err32 < 1e-4


def test_cdf_argument_errors(ctx):
    with pytest.raises(RbhError) as e:
        ctx.envmap_build(np.ones((8, 6, 3), np.float32))
    assert e.value.code == RBH_ERR_INVALID
    with pytest.raises(RbhError) as e:
        ctx.envmap_build(np.ones((4, 32772, 3), np.float32))
    assert e.value.code == RBH_ERR_RANGE
