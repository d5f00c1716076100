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


def _trace(ctx, rays, anyhit, tiny_stack=False, binary=False, quant=False):
    hits = np.zeros((len(rays), 4), np.float32)
    mode = int(anyhit) | (2 if tiny_stack else 0) | (4 if binary else 0) | (8 if quant else 0)
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


def test_four_wide_tree_matches_binary(ctx):
    # the renderer traverses the four-wide collapse of the binary LBVH: the
    # closest t is the minimum over the same triangle tests, so it is
    # bit-identical; only exact ties between triangles may name another one
    sc = scenes.spheres(n_spheres=30, level=4, n_lights=500, width=64, height=64, spp=1, env_size=None)
    _upload(ctx, sc.verts, sc.indices)
    rng = np.random.default_rng(21)
    rays = _rays(rng, 30000, [-40, -5, -100], [40, 20, -20])
    a = _trace(ctx, rays, False)
    b = _trace(ctx, rays, False, binary=True)
    assert np.array_equal(a[:, 0].view(np.uint32), b[:, 0].view(np.uint32))
    ap, bp = a[:, 1].view(np.uint32), b[:, 1].view(np.uint32)
    assert np.array_equal(ap == 0xFFFFFFFF, bp == 0xFFFFFFFF)
    assert (ap == bp).mean() > 0.999
    a2 = _trace(ctx, rays, False, tiny_stack=True)
    assert np.array_equal(a.view(np.uint32), a2.view(np.uint32))
    rays[:, 7] = rng.uniform(1, 60, len(rays))
    ea = _trace(ctx, rays, True)[:, 1].view(np.uint32) == 0xFFFFFFFF
    eb = _trace(ctx, rays, True, binary=True)[:, 1].view(np.uint32) == 0xFFFFFFFF
    assert np.array_equal(ea, eb)


@pytest.mark.parametrize("which", ["soup", "spheres"])
def test_quantised_nodes_match_full_precision(ctx, which):
    # the 64-byte nodes hold conservative 8-bit child boxes: they prune less
    # but never a triangle the ray hits before its current best, so the
    # closest t is bit-identical and the any-hit occlusion the same
    rng = np.random.default_rng(31)
    if which == "soup":
        c = rng.uniform(-5, 5, (3000, 3))
        v = (c[:, None, :] + rng.normal(scale=0.5, size=(3000, 3, 3))).reshape(-1, 3).astype(np.float32)
        idx = np.arange(9000, dtype=np.uint32).reshape(-1, 3)
        _upload(ctx, v, idx)
        rays = _rays(rng, 30000, -6, 6)
    else:
        sc = scenes.spheres(n_spheres=30, level=4, n_lights=500, width=64, height=64, spp=1, env_size=None)
        _upload(ctx, sc.verts, sc.indices)
        rays = _rays(rng, 30000, [-40, -5, -100], [40, 20, -20])
    a = _trace(ctx, rays, False)
    b = _trace(ctx, rays, False, quant=True)
    assert np.array_equal(a[:, 0].view(np.uint32), b[:, 0].view(np.uint32))
    assert (a[:, 1].view(np.uint32) == b[:, 1].view(np.uint32)).mean() > 0.999
    b2 = _trace(ctx, rays, False, tiny_stack=True, quant=True)
    assert np.array_equal(b.view(np.uint32), b2.view(np.uint32))
    rays[:, 7] = rng.uniform(0.5, 30, len(rays))
    ea = _trace(ctx, rays, True)[:, 1].view(np.uint32) == 0xFFFFFFFF
    eb = _trace(ctx, rays, True, quant=True)[:, 1].view(np.uint32) == 0xFFFFFFFF
    assert np.array_equal(ea, eb)


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


NODE_DTYPE = np.dtype([(k, "f4") for k in ("lminx", "lmaxx", "lminy", "lmaxy", "rminx", "rmaxx", "rminy", "rmaxy",
                                           "lminz", "lmaxz", "rminz", "rmaxz")] +
                      [("left", "i4"), ("right", "i4"), ("lcnt", "u4"), ("rcnt", "u4")])
TRI_DTYPE = np.dtype([("v0", "f4", 3), ("prim", "u4"), ("v1", "f4", 3), ("material", "u4"), ("v2", "f4", 3),
                      ("light", "u4"), ("idx", "u4", 3), ("pad", "u4")])


def _read_lbvh(ctx, n):
    import ctypes
    keys = np.zeros(n, np.uint64)
    nodes = np.zeros(max(1, n - 1), NODE_DTYPE)
    tris = np.zeros(n, TRI_DTYPE)
    nn = ctypes.c_uint32()
    ctx.check(L.lib().rbh_test_lbvh_read(ctx.h, L.ptr(keys), L.ptr(nodes), L.ptr(tris), ctypes.byref(nn)))
    return keys, nodes[:nn.value], tris


@pytest.mark.parametrize("which", ["soup", "spheres", "soup_large"])
def test_morton_keys_bit_exact_and_tree_invariants(ctx, which):
    # soup_large: 300k triangles, 74 tiles of the radix sort's 4096 keys
    if which.startswith("soup"):
        rng = np.random.default_rng(21)
        ntri = 3000 if which == "soup" else 300000
        c = rng.uniform(-5, 5, (ntri, 3))
        v = (c[:, None, :] + rng.normal(scale=0.5, size=(ntri, 3, 3))).reshape(-1, 3).astype(np.float32)
        idx = np.arange(3 * ntri, dtype=np.uint32).reshape(-1, 3)
    else:
        sc = scenes.spheres(n_spheres=12, level=3, n_lights=20, width=8, height=8, spp=1, env_size=None)
        v, idx = sc.verts.astype(np.float32), sc.indices.astype(np.uint32)
    _upload(ctx, v, idx)
    n = len(idx)
    keys, nodes, tris = _read_lbvh(ctx, n)
    # Morton codes on the triangle-centroid grid, FP32 as the kernel computes
    # them: the triangle index in the low b bits (b holds n - 1), the code in
    # min(21, (64 - b) // 3) bits per axis above it, x the most significant
    tv = v[idx]                                            # n x 3 x 3
    lo, hi = tv.min(axis=1), tv.max(axis=1)
    ce = (np.float32(0.5) * (lo + hi)).astype(np.float32)
    cmin, cmax = ce.min(axis=0), ce.max(axis=0)
    ext = (cmax - cmin).astype(np.float32)
    u = np.where(ext > 0, (ce - cmin) / np.where(ext > 0, ext, 1), np.float32(0.5)).astype(np.float32)
    b = max(1, int(n - 1).bit_length())
    ab = min(21, (64 - b) // 3)
    cells = np.float32(1 << ab)
    q = np.clip(u * cells, 0, cells - 1).astype(np.uint64)
    m = np.zeros(n, np.uint64)
    for bit in range(ab):                                  # interleave bit by bit: ... x1 y1 z1 x0 y0 z0
        for axis in range(3):
            m |= ((q[:, axis] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + 2 - axis)
    ref = np.sort((m << np.uint64(b)) | np.arange(n, dtype=np.uint64))
    assert np.array_equal(keys, ref)
    # leaf order and triangle records
    assert np.array_equal(tris["prim"], (keys & np.uint64((1 << b) - 1)).astype(np.uint32))
    assert np.array_equal(tris["v0"], v[idx[tris["prim"], 0]]) and np.array_equal(tris["v2"], v[idx[tris["prim"], 2]])
    assert np.array_equal(tris["idx"], idx[tris["prim"]])
    assert np.all(tris["material"] == 0) and np.all(tris["light"] == 0xFFFFFFFF)
    # topology: every internal node but the root and every leaf is a child exactly once
    assert len(nodes) == n - 1
    ch = np.concatenate([nodes["left"], nodes["right"]])
    internal, leaves = ch[ch >= 0], ~ch[ch < 0]
    assert np.array_equal(np.sort(internal), np.arange(1, n - 1))
    assert np.array_equal(np.sort(leaves), np.arange(n))
    # boxes: a child box equals the union of the child's own boxes; leaf boxes bound their triangle
    def child_box(i, side):
        p = "l" if side == 0 else "r"
        return np.array([nodes[i][p + "minx"], nodes[i][p + "miny"], nodes[i][p + "minz"],
                         nodes[i][p + "maxx"], nodes[i][p + "maxy"], nodes[i][p + "maxz"]])
    rng = np.random.default_rng(3)
    for i in rng.choice(n - 1, min(400, n - 1), replace=False):
        for side, c in ((0, nodes[i]["left"]), (1, nodes[i]["right"])):
            b = child_box(i, side)
            if c >= 0:
                u0, u1 = child_box(c, 0), child_box(c, 1)
                assert np.array_equal(b[:3], np.minimum(u0[:3], u1[:3])) and np.array_equal(b[3:], np.maximum(u0[3:], u1[3:]))
            else:
                t = tris[~c]
                tv3 = np.stack([t["v0"], t["v1"], t["v2"]])
                assert np.array_equal(b[:3], tv3.min(axis=0)) and np.array_equal(b[3:], tv3.max(axis=0))
