"""LBVH build + traversal against brute-force CPU ray casting (the OptiX
traversal it replaces is closed, so this is the pin: identical hit/miss and
primitive, t within FP32 rounding, watertight edges)."""
import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import _lib as L
from ratbite_amd import scenes

pytestmark = pytest.mark.gpu


def _upload(ctx, verts, idx):
    mats = np.zeros(1, L.MATERIAL_DTYPE)
    mats["bxdf"] = 2
    ctx.upload_textures(np.zeros(0, L.TEXTURE_DTYPE), np.zeros(0, np.uint8))
    ctx.upload_materials(mats)
    ctx.upload_geometry(verts, idx, np.zeros(16, np.uint8), np.zeros(len(idx), np.uint32),
                        np.full(len(idx), 0xFFFFFFFF, np.uint32))
    ctx.build_accel()
    return ctx.stats()


def _trace(ctx, rays, anyhit, tiny_stack=False):
    hits = np.zeros((len(rays), 4), np.float32)
    mode = int(anyhit) | (2 if tiny_stack else 0)
    ctx.check(L.lib().rbh_test_trace(ctx.h, L.ptr(rays), L.ptr(hits), len(rays), mode))
    return hits


def _brute(verts, idx, rays, anyhit):
    hits = np.zeros((len(rays), 4), np.float32)
    O.lib().orc_trace_brute(O.p(np.ascontiguousarray(verts, np.float32)), O.p(np.ascontiguousarray(idx, np.uint32)),
                            len(idx), O.p(rays), O.p(hits), len(rays), int(anyhit))
    return hits


def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 4:7], r[:, 7] = o, d, 3.4e38
    return r


@pytest.mark.parametrize("ntri", [1, 2, 37, 5000])
def test_triangle_soup_closest_and_any(ctx, ntri):
    rng = np.random.default_rng(ntri)
    c = rng.uniform(-5, 5, (ntri, 3))
    v = (c[:, None, :] + rng.normal(scale=0.7, size=(ntri, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * ntri, dtype=np.uint32).reshape(-1, 3)
    st = _upload(ctx, v, idx)
    print("ntri", ntri, "depth", st["bvhDepth"], flush=True)
    rays = _rays(rng, 20000, -6, 6)
    g = _trace(ctx, rays, False)
    b = _brute(v, idx, rays, False)
    gp, bp = g[:, 1].view(np.uint32), b[:, 1].view(np.uint32)
    assert np.array_equal(gp == 0xFFFFFFFF, bp == 0xFFFFFFFF)
    hit = bp != 0xFFFFFFFF
    same = gp[hit] == bp[hit]
    # coincident-t ties between triangles may resolve either way
    assert same.mean() > 0.999
    np.testing.assert_allclose(g[hit, 0], b[hit, 0], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g[hit][same, 2:], b[hit][same, 2:], atol=1e-5)
    rays[:, 7] = rng.uniform(0.1, 10, len(rays))
    ga = _trace(ctx, rays, True)
    ba = _brute(v, idx, rays, False)
    assert np.array_equal(ga[:, 1].view(np.uint32) == 0xFFFFFFFF, ba[:, 1].view(np.uint32) == 0xFFFFFFFF)


def test_stack_spill_gives_identical_hits(ctx):
    # the same traversal with a 2-entry LDS stack pushes almost every deferred
    # subtree through the global spill columns; results must not change
    sc = scenes.spheres(n_spheres=30, level=4, n_lights=500, width=64, height=64, spp=1, env_size=None)
    _upload(ctx, sc.verts, sc.indices)
    rng = np.random.default_rng(12)
    rays = _rays(rng, 30000, [-40, -5, -100], [40, 20, -20])
    for anyhit in (False, True):
        if anyhit:
            rays[:, 7] = rng.uniform(1, 60, len(rays))
        a = _trace(ctx, rays, anyhit)
        b = _trace(ctx, rays, anyhit, tiny_stack=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_watertight_shared_edges(ctx):
    # a closed icosphere: rays from inside must always hit; rays aimed at
    # vertices and edge midpoints must not slip through
    v, idx = scenes.icosphere(3)
    v = v.astype(np.float32)
    _upload(ctx, v, idx)
    targets = np.concatenate([v, 0.5 * (v[idx[:, 0]] + v[idx[:, 1]]), 0.5 * (v[idx[:, 1]] + v[idx[:, 2]])])
    rng = np.random.default_rng(9)
    origins = rng.uniform(-0.3, 0.3, (len(targets), 3))
    d = targets - origins
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = origins, d, 3.4e38
    g = _trace(ctx, rays, False)
    assert np.all(g[:, 1].view(np.uint32) != 0xFFFFFFFF)


def test_spheres_scene_depth_and_hits(ctx):
    sc = scenes.spheres(n_spheres=30, level=4, n_lights=500, width=64, height=64, spp=1, env_size=None)
    st = _upload(ctx, sc.verts, sc.indices)
    assert 0 < st["bvhDepth"] <= 64
    rng = np.random.default_rng(3)
    rays = _rays(rng, 4000, [-40, -5, -100], [40, 20, -20])
    g = _trace(ctx, rays, False)
    b = _brute(sc.verts, sc.indices, rays, False)
    gp, bp = g[:, 1].view(np.uint32), b[:, 1].view(np.uint32)
    assert np.array_equal(gp == 0xFFFFFFFF, bp == 0xFFFFFFFF)
    hit = bp != 0xFFFFFFFF
    np.testing.assert_allclose(g[hit, 0], b[hit, 0], rtol=1e-5, atol=1e-5)
