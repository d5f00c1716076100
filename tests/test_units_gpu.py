"""Per-function GPU parity against the CPU oracle on identical inputs.

Integer work (Sobol bits, hashes, octahedral codes) is bit-exact; FP32
shading values agree within 1e-5 relative (transcendentals differ between
the device libm and glibc by a few ulp).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import _lib as L
from ratbite_amd import scenes

pytestmark = pytest.mark.gpu


def _call(ctx, name, *args):
    ctx.check(getattr(L.lib(), name)(ctx.h, *args))


def test_sobol_bit_exact(ctx):
    rng = np.random.default_rng(1)
    n = 20000
    rec = np.zeros((n, 5), np.uint32)
    rec[:, 0] = rng.integers(0, 1 << 20, n)            # sample index
    rec[:, 1] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    rec[:, 2] = 16 * rng.integers(1, 40, n)            # dimension offset
    rec[:, 3] = rng.integers(0, 9, n)                  # dimension
    rec[:, 4] = rng.integers(1, 4, n)                  # 1D/2D/3D
    out = np.zeros((n, 3), np.float32)
    _call(ctx, "rbh_test_sobol", L.ptr(rec), L.ptr(out), n)
    assert np.array_equal(out.view(np.uint32), O.sobol(rec).view(np.uint32))


def test_pixel_hash_bit_exact(ctx):
    xy = np.stack(np.meshgrid(np.arange(0, 1920, 7), np.arange(0, 1080, 5)), -1).reshape(-1, 2).astype(np.uint32)
    out = np.zeros(len(xy), np.uint32)
    _call(ctx, "rbh_test_pixel_hash", L.ptr(np.ascontiguousarray(xy)), L.ptr(out), len(xy))
    assert np.array_equal(out, O.pixel_hash(xy))


def test_wavelengths(ctx):
    u = np.linspace(0, 1, 100001, endpoint=False).astype(np.float32)
    lam = np.zeros((len(u), 4), np.float32)
    pdf = np.zeros((len(u), 4), np.float32)
    _call(ctx, "rbh_test_wavelengths", L.ptr(u), L.ptr(lam), L.ptr(pdf), len(u))
    ol, op = O.wavelengths(u)
    np.testing.assert_allclose(lam, ol, rtol=1e-5)
    np.testing.assert_allclose(pdf, op, rtol=1e-5)
    assert lam.min() >= 359.5 and lam.max() < 830.5


def test_oct_bit_exact(ctx):
    rng = np.random.default_rng(2)
    v = rng.normal(size=(50000, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[:6] = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    enc = np.zeros(len(v), np.uint32)
    dec = np.zeros((len(v), 3), np.float32)
    _call(ctx, "rbh_test_oct", L.ptr(v), L.ptr(enc), L.ptr(dec), len(v))
    oenc = np.zeros(len(v), np.uint32)
    O.lib().orc_oct_encode_u32(O.p(v), O.p(oenc), len(v))
    assert np.array_equal(enc, oenc)
    odec = np.zeros((len(v), 3), np.float32)
    O.lib().orc_oct_decode_u32(O.p(oenc), O.p(odec), len(v))
    np.testing.assert_allclose(dec, odec, atol=2e-7)


@pytest.fixture(scope="module")
def spheres_ctx(ctx):
    from ratbite_amd.renderer import Renderer
    sc = scenes.spheres(n_spheres=6, level=3, n_lights=300, width=64, height=64, spp=2, env_size=(256, 128))
    r = Renderer(sc, ctx=ctx)
    return sc, r


def test_rgb_to_spectrum(ctx, spheres_ctx):
    sc, _ = spheres_ctx
    rng = np.random.default_rng(3)
    n = 4096
    inp = np.zeros((n, 7), np.float32)
    inp[:, :3] = rng.uniform(0, 4, (n, 3))
    inp[:, 3:] = rng.uniform(360, 830, (n, 4))
    out = np.zeros((n, 4), np.float32)
    _call(ctx, "rbh_test_rgb_to_spectrum", L.ptr(inp), L.ptr(out), n)
    b = sc.tables["basis"]
    idx = np.floor(inp[:, 3:].astype(np.float64) + 0.5).astype(int) - 360   # lroundf: halves away from zero
    ref = (b[0][idx] * inp[:, 0:1] + b[1][idx] * inp[:, 1:2] + b[2][idx] * inp[:, 2:3]).astype(np.float32)
    np.testing.assert_allclose(out, ref, rtol=1e-6)


def test_light_tree_pdf_and_sample(ctx, spheres_ctx):
    sc, r = spheres_ctx
    nodes, lp, bm = r.light_tree()
    rng = np.random.default_rng(4)
    n = 3000
    pos = rng.uniform([-50, -5, -100], [50, 20, -10], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    lights = rng.integers(0, len(sc.lights), n).astype(np.uint32)
    pdf = np.zeros(n, np.float32)
    _call(ctx, "rbh_test_light_pdf", L.ptr(lights), L.ptr(pn), L.ptr(pdf), n)
    opdf = np.zeros(n, np.float32)
    O.lib().orc_light_tree_pdf(O.p(nodes), O.p(lp), O.p(bm), 1, O.p(lights), O.p(pn), O.p(opdf), n)
    rel = np.abs(pdf - opdf) / np.maximum(np.abs(opdf), 1e-30)
    print("light pdf rel median %.2g q999 %.2g max %.2g" % (np.median(rel), np.quantile(rel, 0.999), rel.max()),
          flush=True)
    np.testing.assert_allclose(pdf, opdf, rtol=1e-5, atol=1e-30)
    inp = np.ascontiguousarray(np.concatenate([rng.random((n, 1)), pos, nrm], 1), np.float32)
    ti = np.zeros((n, 2), np.uint32)
    sp = np.zeros(n, np.float32)
    _call(ctx, "rbh_test_light_sample", L.ptr(inp), L.ptr(ti), L.ptr(sp), n)
    oti = np.zeros((n, 2), np.uint32)
    osp = np.zeros(n, np.float32)
    O.lib().orc_light_tree_sample(O.p(nodes), O.p(lp), O.p(bm), 1, O.p(inp), O.p(oti), O.p(osp), n)
    same = np.all(ti == oti, axis=1)
    # a random number within a few ulp of a node split may go either way
    assert same.mean() > 0.999
    np.testing.assert_allclose(sp[same], osp[same], rtol=1e-5, atol=1e-30)


def test_light_tree_pdf_against_fp64(ctx, spheres_ctx):
    # the GPU's FP32 pdfs against the FP64 restatement (exact erf/expm1): the
    # same error profile as the FP32 oracle (tests/test_light_tree.py)
    import light_tree64 as T
    from test_light_tree import fp32_error_summary
    sc, r = spheres_ctx
    nodes, lp, bm = r.light_tree()
    rng = np.random.default_rng(14)
    n = 20000
    pos = rng.uniform([-50, -5, -100], [50, 20, -10], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    lights = rng.integers(0, len(sc.lights), n).astype(np.uint32)
    pdf = np.zeros(n, np.float32)
    _call(ctx, "rbh_test_light_pdf", L.ptr(lights), L.ptr(pn), L.ptr(pdf), n)
    p64 = T.pdf(nodes, bm, lights, pn, env=sc.env is not None)
    med, q99, wmean = fp32_error_summary(pdf, p64)
    print("gpu light pdf vs fp64: median %.2g, q99 (pdf >= 1e-4) %.2g, pdf-weighted mean %.2g" % (med, q99, wmean),
          flush=True)
    assert med < 1e-5 and q99 < 5e-3 and wmean < 2e-4


def _bxdf_records(sc, n, rng):
    rec = np.zeros((n, 64), np.float32)
    u = rec.view(np.uint32)
    mats = rng.integers(0, len(sc.materials), n)
    u[:, 0] = mats
    u[:, 1] = rng.choice([0, 4], n)                             # REGULARIZED sometimes
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 2:5] = d
    nrm = rng.normal(size=(n, 3)); nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    t = np.cross(nrm, rng.normal(size=(n, 3))); t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    rec[:, 5:8], rec[:, 8:11], rec[:, 11:14] = t, b, nrm
    rec[:, 14:18] = rng.uniform(0, 5, (n, 4))
    ld = rng.normal(size=(n, 3)); ld /= np.linalg.norm(ld, axis=1, keepdims=True)
    rec[:, 18:21] = ld
    rec[:, 21] = rng.uniform(0.1, 10, n)
    u[:, 22] = rng.random(n) < 0.2
    u[:, 23] = rng.integers(0, 1 << 16, n)
    u[:, 24] = 16 * rng.integers(1, 10, n)
    u[:, 25] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    lam, lpdf = O.wavelengths(rng.random(n).astype(np.float32))
    rec[:, 26:30], rec[:, 30:34] = lam, lpdf
    m = sc.materials[mats]
    u[:, 34] = np.where((m["flags"] & 2) != 0, 3, 1)
    rec[:, 35:38] = m["baseColor"][:, :3]
    rec[:, 38] = m["metalness"]
    rec[:, 39] = m["ior"]
    rec[:, 40] = m["roughness"]
    rec[:, 41] = m["transmission"]
    rec[:, 42:45] = m["sheenColor"]
    rec[:, 45] = m["sheenRoughness"]
    rec[:, 46:50] = rng.uniform(0.2, 1.0, (n, 4))
    rec[:, 50] = 1.0
    rec[:, 51:55] = rng.uniform(0, 1, (n, 4))
    return rec


def test_bxdf_parity(ctx, spheres_ctx):
    sc, r = spheres_ctx
    rng = np.random.default_rng(5)
    n = 20000
    rec = _bxdf_records(sc, n, rng)
    out = np.zeros((n, 16), np.float32)
    _call(ctx, "rbh_test_bxdf", L.ptr(rec), L.ptr(out), n)
    os_ = O.OracleScene(sc)
    oout = np.zeros((n, 16), np.float32)
    assert O.lib().orc_bxdf_unit(ctypes.byref(os_.s), O.p(rec), O.p(oout), n) == 0
    flags, oflags = out[:, 12].view(np.uint32), oout[:, 12].view(np.uint32)
    # lobe choice compares a random number with FP32 weights: a handful may flip
    same = flags == oflags
    same &= np.all(np.isfinite(out[:, :12]) == np.isfinite(oout[:, :12]), axis=1)
    assert same.mean() > 0.995, same.mean()
    a, b = out[same], oout[same]
    live = (flags[same] & (1 << 4)) == 0          # not terminated
    exact_dir = np.all(a[live][:, 9:12].view(np.uint32) == b[live][:, 9:12].view(np.uint32), axis=1)
    print("bit-identical sampled directions %.4f" % exact_dir.mean(), flush=True)
    # spectral values (L, throughput) and the pdf: relative error
    stats = {}
    for name, cols in (("L", slice(0, 4)), ("T", slice(4, 8)), ("pdf", slice(8, 9))):
        x, y = a[live][:, cols], b[live][:, cols]
        fin = np.isfinite(x) & np.isfinite(y) & (np.abs(y) > 1e-12)
        rel = np.abs(x - y)[fin] / np.abs(y)[fin]
        stats[name] = (np.median(rel), np.quantile(rel, 0.999), rel.max())
        print(name, "rel median %.2g q999 %.2g max %.2g" % stats[name], flush=True)
    for name, (med, q999, _) in stats.items():
        assert q999 <= 1e-5, name
        assert med < 1e-6, name
    # unit directions: absolute error
    d = np.abs(a[live][:, 9:12] - b[live][:, 9:12])
    print("dir abs q999 %.2g max %.2g" % (np.quantile(d, 0.999), d.max()))
    assert np.quantile(d, 0.999) < 1e-5
