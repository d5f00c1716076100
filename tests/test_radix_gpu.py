"""The stable LSD radix sort of the LBVH build and the guide's training-set
compaction, against numpy's stable argsort and nonzero on the same inputs."""
import ctypes

import numpy as np
import pytest

from ratbite_amd import _lib as L

pytestmark = pytest.mark.gpu


def _sort(ctx, keys, vals, lo, hi):
    n = len(keys)
    ok = np.zeros(n, np.uint64)
    ov = np.zeros(n, np.uint32) if vals is not None else None
    ctx.check(L.lib().rbh_test_radix_sort(ctx.h, L.ptr(keys), None if vals is None else L.ptr(vals), n, lo, hi,
                                          L.ptr(ok), None if ov is None else L.ptr(ov)))
    return ok, ov


@pytest.mark.parametrize("n", [0, 1, 63, 4095, 4096, 4097, 100003, 1 << 20])
@pytest.mark.parametrize("lo,hi", [(0, 64), (0, 7), (3, 29), (40, 64)])
def test_radix_sort_is_a_stable_sort(ctx, n, lo, hi):
    print("radix sort", n, lo, hi, flush=True)
    rng = np.random.default_rng(n * 131 + lo * 7 + hi)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    # many equal digits: every fourth key copies a neighbour's sort field
    m = len(keys[1::4])
    keys[0:4 * m:4] = keys[1::4]
    vals = np.arange(n, dtype=np.uint32)
    field = (keys >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1)
    order = np.argsort(field, kind="stable")
    ok, ov = _sort(ctx, keys, vals, lo, hi)
    assert np.array_equal(ok, keys[order])
    assert np.array_equal(ov, vals[order])
    ok2, _ = _sort(ctx, keys, None, lo, hi)
    assert np.array_equal(ok2, keys[order])


def test_radix_sort_depth_keys(ctx):
    # the LBVH's depth sort: 7-bit keys with many equal values, ids as values
    rng = np.random.default_rng(7)
    n = 300000
    keys = rng.integers(0, 40, n).astype(np.uint64)
    vals = rng.permutation(n).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    ok, ov = _sort(ctx, keys, vals, 0, 7)
    assert np.array_equal(ok, keys[order]) and np.array_equal(ov, vals[order])


@pytest.mark.parametrize("n", [0, 1, 5000, 4096 * 3 + 17, 1 << 20])
def test_select_positive_keeps_order(ctx, n):
    print("select", n, flush=True)
    rng = np.random.default_rng(n + 5)
    w = rng.normal(size=n).astype(np.float32)
    if n >= 8:
        w[:4] = [np.nan, 0.0, -0.0, np.inf]
    out = np.zeros(max(n, 1), np.uint32)
    cnt = ctypes.c_uint32()
    ctx.check(L.lib().rbh_test_select_positive(ctx.h, L.ptr(w), n, L.ptr(out), ctypes.byref(cnt)))
    ref = np.nonzero(w > 0)[0].astype(np.uint32)
    assert cnt.value == len(ref)
    assert np.array_equal(out[: cnt.value], ref)
