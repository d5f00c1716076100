"""Emitter flux of the glTF path (rbh_emissive_flux, host/emissive.cpp) and
its UV-footprint rasteriser (rbh_rasterise_uv_triangle):

- against the reference's own rasteriser, src/core/sw_rast.h compiled from
  where it lies into oracle/_ref/libswrast_ref.so (oracle/ref/Makefile; the
  comparison is skipped where that build is absent), texel for texel and in
  the same order, on random triangles including flat-top, flat-bottom and
  degenerate ones; and against golden texel lists made by it
  (tests/golden/sw_rast_cases.json, gen_sw_rast.py), which hold everywhere;
- against what the rule means: a texel is visited exactly when its centre
  lies in the triangle with half-open bounds (rows with bottom <= y+0.5 <
  top, columns with left edge <= x+0.5 < right edge), checked by exact
  rational arithmetic away from the edges, and the count of a large
  triangle is its area within its perimeter;
- flux closed forms: an untextured emitter carries mean(factor) x area x pi,
  a constant texture multiplies that by (r+g+b)a/3, a footprint that covers
  no texel centre averages the texels under its corners, and the wrap /
  clamp / mirror rules pick the texels the loader's rules name.
"""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from ratbite_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libswrast_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "sw_rast_cases.json")


def raster(a, b, c, cap=1 << 16):
    xy = np.zeros((cap, 2), np.int32)
    n = ctypes.c_uint32()
    pts = [np.ascontiguousarray(p, np.float32) for p in (a, b, c)]
    rc = L.lib().rbh_rasterise_uv_triangle(L.ptr(pts[0]), L.ptr(pts[1]), L.ptr(pts[2]), L.ptr(xy), cap,
                                           ctypes.byref(n))
    assert rc == 0, rc
    return xy[:n.value].copy()


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libswrast_ref.so is not built (no reference checkout)")
    lib = ctypes.CDLL(REF_SO)
    lib.ref_rasterise.restype = ctypes.c_int
    lib.ref_rasterise.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int]
    return lib


def ref_raster(lib, a, b, c, cap=1 << 16):
    xy = np.zeros((cap, 2), np.int32)
    pts = [np.ascontiguousarray(p, np.float32) for p in (a, b, c)]
    n = lib.ref_rasterise(L.ptr(pts[0]), L.ptr(pts[1]), L.ptr(pts[2]), L.ptr(xy), cap)
    assert n <= cap
    return xy[:n].copy()


def random_triangles(seed, n):
    rng = np.random.default_rng(seed)
    tris = []
    for k in range(n):
        kind = k % 6
        p = rng.uniform(-20, 60, (3, 2)).astype(np.float32)
        if kind == 1:
            p[2, 1] = p[1, 1]                       # flat bottom or top
        elif kind == 2:
            p[1, 1] = p[0, 1]
        elif kind == 3:
            p = np.round(p * 2) / 2                 # corners on texel edges and centres
        elif kind == 4:
            p = p[0] + rng.uniform(-0.6, 0.6, (3, 2)).astype(np.float32)   # tiny
        elif kind == 5:
            p[2] = p[0] + (p[1] - p[0]) * np.float32(rng.uniform())          # degenerate (collinear)
        tris.append(p.astype(np.float32))
    return tris


def test_rasteriser_matches_the_reference_rasteriser():
    lib = ref_lib()
    total = 0
    for t in random_triangles(1, 3000):
        mine = raster(*t)
        ref = ref_raster(lib, *t)
        assert np.array_equal(mine, ref), t
        total += len(ref)
    print("3000 triangles, %d texels, identical lists" % total)


def test_rasteriser_matches_golden_cases():
    with open(GOLDEN) as f:
        cases = json.load(f)
    assert len(cases) >= 50
    for c in cases:
        t = np.asarray(c["tri"], np.float32).reshape(3, 2)
        got = raster(*t)
        assert got.tolist() == c["texels"], c["tri"]


def _inside_exact(t, x, y):
    """(x+0.5, y+0.5) inside the triangle: strictly, or on an edge (None)."""
    P = (Fraction(2 * x + 1, 2), Fraction(2 * y + 1, 2))
    V = [(Fraction(float(p[0])), Fraction(float(p[1]))) for p in t]
    s = []
    for i in range(3):
        a, b = V[i], V[(i + 1) % 3]
        s.append((b[0] - a[0]) * (P[1] - a[1]) - (b[1] - a[1]) * (P[0] - a[0]))
    if all(v > 0 for v in s) or all(v < 0 for v in s):
        return True
    if any(v == 0 for v in s):
        return None
    return False


def _edge_distance(t, x, y):
    P = np.array([x + 0.5, y + 0.5])
    d = []
    for i in range(3):
        a, b = t[i].astype(np.float64), t[(i + 1) % 3].astype(np.float64)
        e = b - a
        L2 = e @ e
        if L2 == 0:
            d.append(np.linalg.norm(P - a))
            continue
        u = np.clip((P - a) @ e / L2, 0, 1)
        d.append(np.linalg.norm(P - (a + u * e)))
    return min(d)


def test_rasteriser_visits_the_centres_inside_the_triangle():
    rng = np.random.default_rng(7)
    for _ in range(150):
        t = rng.uniform(-5, 25, (3, 2)).astype(np.float32)
        got = {tuple(v) for v in raster(*t)}
        assert len(got) == len(raster(*t))          # no texel twice
        lo = np.floor(t.min(axis=0)).astype(int) - 1
        hi = np.ceil(t.max(axis=0)).astype(int) + 1
        for y in range(lo[1], hi[1]):
            for x in range(lo[0], hi[0]):
                inside = _inside_exact(t, x, y)
                if inside is None or _edge_distance(t, x, y) < 1e-3:
                    continue                         # on an edge: the half-open rule decides
                assert ((x, y) in got) == inside, (t, x, y)


def test_rasteriser_count_is_the_area():
    t = np.array([[3.2, 1.1], [250.7, 40.3], [90.1, 180.9]], np.float32)
    n = len(raster(*t))
    e1, e2 = t[1] - t[0], t[2] - t[0]
    area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
    per = sum(np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3))
    assert abs(n - area) < per


def flux(verts, uvs, factor, rgba=None, addr=(0, 0)):
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    out = np.zeros(len(verts), np.float32)
    f = np.ascontiguousarray(factor, np.float32)
    if rgba is None:
        rc = L.lib().rbh_emissive_flux(L.ptr(verts), None, len(verts), L.ptr(f), None, 0, 0, 0, 0, L.ptr(out))
    else:
        uvs = np.ascontiguousarray(uvs, np.float32).reshape(-1, 6)
        rgba = np.ascontiguousarray(rgba, np.uint8)
        rc = L.lib().rbh_emissive_flux(L.ptr(verts), L.ptr(uvs), len(verts), L.ptr(f), L.ptr(rgba), rgba.shape[1],
                                       rgba.shape[0], addr[0], addr[1], L.ptr(out))
    assert rc == 0
    return out


TRI = np.array([[0, 0, 0, 2, 0, 0, 0, 3, 0]], np.float32)        # area 3


def test_untextured_flux_is_factor_area_pi():
    f = flux(TRI, None, (1.0, 2.0, 6.0))
    assert abs(f[0] - 3.0 * 3.0 * np.pi) < 1e-5 * f[0]


def test_constant_texture_scales_by_its_emission():
    tex = np.zeros((8, 8, 4), np.uint8)
    tex[...] = (255, 51, 102, 204)
    uv = np.array([[0.05, 0.05, 0.9, 0.1, 0.2, 0.95]], np.float32)
    f = flux(TRI, uv, (1.0, 1.0, 1.0), tex)
    e = (1.0 + 0.2 + 0.4) * 0.8 / 3.0
    assert abs(f[0] - e * 3.0 * np.pi) < 1e-5 * f[0]


def test_tiny_footprint_averages_its_corner_texels():
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., 3] = 255
    tex[1, 1, :3] = 255      # the texel under all three corners
    uv = np.array([[0.30, 0.30, 0.32, 0.30, 0.30, 0.32]], np.float32)   # inside texel (1, 1), no centre covered
    f = flux(TRI, uv, (1.0, 1.0, 1.0), tex)
    assert abs(f[0] - 3.0 * np.pi) < 1e-5 * f[0]


@pytest.mark.parametrize("addr,expect", [((0, 0), 3), ((1, 1), 0), ((2, 2), 3)])
def test_address_rules_pick_the_named_texels(addr, expect):
    # a footprint in column 4 of a 4-texel-wide texture: wrap reads column 0,
    # clamp column 3, and the loader's mirror rule (which mirrors the even
    # tiles, the first one included, and leaves odd tiles as they are) column 0
    tex = np.zeros((4, 4, 4), np.uint8)
    tex[..., 3] = 255
    tex[:, 0, :3] = 255
    uv = np.array([[1.0, 0.0, 1.25, 0.0, 1.0, 1.0]], np.float32)
    got = raster(np.array([4.0, 0.0]), np.array([5.0, 0.0]), np.array([4.0, 4.0]))
    assert set(int(x) for x in got[:, 0]) == {4}
    f = flux(TRI, uv, (1.0, 1.0, 1.0), tex, addr)
    assert abs(f[0] - (3.0 * np.pi if expect == 3 else 0.0)) < 1e-4
