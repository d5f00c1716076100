"""Pins of the oracle's microfacet, sheen, Fresnel, spherical-Gaussian and
sampling functions against what the mathematics fixes: normalisation
integrals by quadrature, the visible-normal identity of Smith masking,
sample histograms against the stated densities, textbook complex Fresnel
evaluated with numpy's complex arithmetic, and exact identities.

Functions pinned (reference files): GGX D, Smith G1 (src/device/surface.h:138-174),
spherical-cap and bounded VNDF samplers and pdfs (surface.h:308-372), complex
Fresnel (surface.h:176-195), F82 (surface.h:254-264), LTC sheen (surface.h:53-88),
SG integral / product / clamped-cosine integral / vMF conversions
(src/core/math_util.h:246-587), Gaussian pixel filter (src/device/filter.h:14-27),
refraction (src/device/util.h:181-198).
"""
import numpy as np
import pytest

import oracle_lib as O


def mu(fn, rec):
    rec = np.ascontiguousarray(rec, np.float32)
    assert rec.ndim == 2 and rec.shape[1] == 8
    out = np.zeros((len(rec), 4), np.float32)
    assert O.lib().orc_math_unit(fn, O.p(rec), O.p(out), len(rec)) == 0
    return out


def _rec(n, **cols):
    r = np.zeros((n, 8), np.float32)
    for k, v in cols.items():
        r[:, int(k[1:])] = v
    return r


def _hemisphere_grid(nc, nphi, lo=0.0):
    """Midpoint grid in (cos theta, phi) over cos in [lo, 1]; returns theta, phi, weight (d omega)."""
    c = lo + (np.arange(nc) + 0.5) * (1.0 - lo) / nc
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    C, PH = np.meshgrid(c, ph, indexing="ij")
    w = (1.0 - lo) / nc * 2 * np.pi / nphi
    return np.arccos(C).ravel(), PH.ravel(), w


def _dirs(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], 1)


@pytest.mark.parametrize("ax,ay", [(0.1, 0.1), (0.5, 0.5), (1.0, 1.0), (0.3, 0.8)])
def test_ggx_D_projected_area_is_one(ax, ay):
    # isotropic lobes do not depend on phi: one phi column carries the 2 pi
    th, ph, w = _hemisphere_grid(20000, 1 if ax == ay else 64)
    m = _dirs(th, ph)
    D = mu(0, np.concatenate([m, np.tile([ax, ay], (len(m), 1)), np.zeros((len(m), 3))], 1))[:, 0].astype(np.float64)
    total = np.sum(D * m[:, 2]) * w
    assert abs(total - 1.0) < 2e-3, total


@pytest.mark.parametrize("ax,ay", [(0.2, 0.2), (0.7, 0.7), (0.25, 0.6)])
@pytest.mark.parametrize("theta_o", [0.2, 0.9, 1.35])
def test_smith_G1_visible_normal_identity(ax, ay, theta_o):
    # integral of D(m) max(0, wo.m) over m equals cos(theta_o) / G1(wo)
    wo = np.array([np.sin(theta_o) * np.cos(0.7), np.sin(theta_o) * np.sin(0.7), np.cos(theta_o)])
    th, ph, w = _hemisphere_grid(6000, 360)
    m = _dirs(th, ph)
    D = mu(0, np.concatenate([m, np.tile([ax, ay], (len(m), 1)), np.zeros((len(m), 3))], 1))[:, 0].astype(np.float64)
    lhs = np.sum(D * np.maximum(0.0, m @ wo)) * w
    G1 = mu(1, np.concatenate([wo[None], [[ax, ay, 0, 0, 0]]], 1))[0, 0]
    assert abs(lhs - wo[2] / G1) / (wo[2] / G1) < 3e-3, (lhs, wo[2] / G1)


def _cap_support(wo, alpha, m, bounded):
    """Normals the spherical-cap samplers can return: in the stretched
    configuration h = ws + wi with wi on the unit sphere, wi.z >= -b, where
    b = ws.z for the visible-normal cap (Dupuy and Benyoub 2023) and
    b = k ws.z for the bounded cap (Eto and Tokuyoshi 2023)."""
    ws = np.array([wo[0] * alpha, wo[1] * alpha, wo[2]])
    ws /= np.linalg.norm(ws)
    h = np.stack([m[:, 0] / alpha, m[:, 1] / alpha, m[:, 2]], 1)
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    wi = 2 * (h @ ws)[:, None] * h - ws
    b = ws[2]
    if bounded:
        s = 1 + np.hypot(wo[0], wo[1])
        k = (1 - alpha * alpha) * s * s / (s * s + alpha * alpha * wo[2] * wo[2])
        b = k * ws[2]
    return (wi[:, 2] >= -b) & ((h @ ws) > 0)


@pytest.mark.parametrize("fn", [3, 5])
@pytest.mark.parametrize("alpha,theta_o", [(0.3, 0.4), (0.6, 1.2), (0.9, 0.1)])
def test_vndf_pdf_integrates_to_one(fn, alpha, theta_o):
    wo = np.array([np.sin(theta_o) * np.cos(1.1), np.sin(theta_o) * np.sin(1.1), np.cos(theta_o)])
    th, ph, w = _hemisphere_grid(3000, 360)
    rec = np.zeros((len(th), 8), np.float32)
    rec[:, 0:3] = wo
    rec[:, 3] = rec[:, 4] = alpha
    rec[:, 5], rec[:, 6] = th, ph
    pdf = mu(fn, rec)[:, 0].astype(np.float64)
    sup = _cap_support(wo, alpha, _dirs(th, ph), fn == 5)
    total = np.sum(pdf * sup) * w
    assert abs(total - 1.0) < 3e-3, total
    if fn == 5:
        # the bounded cap never covers more than the visible-normal cap
        assert np.all(sup <= _cap_support(wo, alpha, _dirs(th, ph), False))


@pytest.mark.parametrize("fn_s,fn_p", [(2, 3), (4, 5)])
def test_vndf_samples_follow_pdf(fn_s, fn_p):
    rng = np.random.default_rng(11)
    alpha, theta_o = 0.45, 0.8
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    n = 400000
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0:3] = wo
    rec[:, 3] = rec[:, 4] = alpha
    rec[:, 5:7] = rng.random((n, 2))
    m = mu(fn_s, rec)[:, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-5)
    assert _cap_support(wo, alpha, m * (1 - 1e-6), fn_s == 4).mean() > 0.999
    nb_c, nb_p = 16, 16
    c = np.clip(m[:, 2], 0, 1 - 1e-9)
    p = np.mod(np.arctan2(m[:, 1], m[:, 0]), 2 * np.pi)
    hist, _, _ = np.histogram2d(c, p, bins=[nb_c, nb_p], range=[[0, 1], [0, 2 * np.pi]])
    # expected mass per bin from the pdf by quadrature (8x8 points per bin)
    th, ph, w = _hemisphere_grid(nb_c * 8, nb_p * 8)
    q = np.zeros((len(th), 8), np.float32)
    q[:, 0:3] = wo
    q[:, 3] = q[:, 4] = alpha
    q[:, 5], q[:, 6] = th, ph
    pdf = mu(fn_p, q)[:, 0].astype(np.float64) * _cap_support(wo, alpha, _dirs(th, ph), fn_s == 4)
    mass = (pdf * w).reshape(nb_c, 8, nb_p, 8).sum(axis=(1, 3))
    expect = mass * n
    big = expect > 200
    z = (hist[big] - expect[big]) / np.sqrt(expect[big])
    # bins cut by the support edge carry quadrature error; the bulk must agree
    assert np.median(z ** 2) < 1.5, np.median(z ** 2)
    assert np.mean(np.abs(hist[big] - expect[big])) / np.mean(expect[big]) < 0.02
    assert hist[~big].sum() < 0.02 * n


def test_complex_fresnel_matches_textbook():
    rng = np.random.default_rng(3)
    n = 5000
    ct = rng.uniform(0.0, 1.0, n)
    eta = rng.uniform(0.1, 4.0, n)
    k = rng.uniform(0.0, 6.0, n)
    k[:500] = 0.0                      # dielectric limit
    out = mu(6, _rec(n, c0=ct, c1=eta, c2=k))[:, 0].astype(np.float64)
    e = (eta + 1j * k).astype(np.complex128)
    ct64 = ct.astype(np.float32).astype(np.float64)
    s2 = 1 - ct64 ** 2
    cost = np.sqrt(1 - s2 / e ** 2)
    rs = (ct64 - e * cost) / (ct64 + e * cost)
    rp = (e * ct64 - cost) / (e * ct64 + cost)
    ref = 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)
    # dielectric limit with eta < 1 beyond the critical angle reflects totally
    np.testing.assert_allclose(out, ref, rtol=2e-4, atol=2e-6)
    # normal incidence closed form
    n0 = mu(6, _rec(3, c0=1.0, c1=[1.5, 0.2, 2.9], c2=[0.0, 3.4, 3.1]))[:, 0]
    e0 = np.array([1.5, 0.2, 2.9]); k0 = np.array([0.0, 3.4, 3.1])
    np.testing.assert_allclose(n0, ((e0 - 1) ** 2 + k0 ** 2) / ((e0 + 1) ** 2 + k0 ** 2), rtol=1e-5)


def test_f82_with_unit_specular_colour_is_schlick():
    rng = np.random.default_rng(4)
    r = _rec(2000, c0=rng.random(2000), c1=rng.random(2000), c2=rng.random(2000), c3=rng.random(2000))
    out = mu(7, r)
    assert np.all(out[:, 3] == 0.0)
    f0 = r[:, 0].astype(np.float64)
    np.testing.assert_allclose(out[:, 0], f0 + (1 - f0) * (1 - r[:, 3].astype(np.float64)) ** 5, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("A,B,theta_o", [(1.0, 0.0, 0.3), (0.4, 0.2, 0.9), (0.15, -0.3, 1.3)])
def test_ltc_pdf_normalised_and_sampled(A, B, theta_o):
    wo = np.array([np.sin(theta_o) * np.cos(0.5), np.sin(theta_o) * np.sin(0.5), np.cos(theta_o)])
    th, ph, w = _hemisphere_grid(3000, 360)
    rec = np.zeros((len(th), 8), np.float32)
    rec[:, 0:3] = wo
    rec[:, 3], rec[:, 4] = A, B
    rec[:, 5], rec[:, 6] = th, ph
    out = mu(9, rec)
    pdf = out[:, 1].astype(np.float64)
    assert abs(np.sum(pdf) * w - 1.0) < 5e-3, np.sum(pdf) * w
    # f = pdf / |cos theta_i|
    ok = np.abs(np.cos(th)) > 1e-3
    np.testing.assert_allclose(out[ok, 0] * np.abs(np.cos(th[ok])), pdf[ok], rtol=1e-4, atol=1e-7)
    # samples follow the pdf: mean of cos(theta_i) against quadrature
    rng = np.random.default_rng(5)
    n = 200000
    s = np.zeros((n, 8), np.float32)
    s[:, 0:3] = wo
    s[:, 3], s[:, 4] = A, B
    s[:, 5:7] = rng.random((n, 2))
    wi = mu(8, s)[:, :3].astype(np.float64)
    assert np.all(wi[:, 2] >= -1e-6)
    mean_q = np.sum(pdf * np.cos(th)) * w
    assert abs(wi[:, 2].mean() - mean_q) < 4 * wi[:, 2].std() / np.sqrt(n) + 2e-3


def test_sg_integral_and_expm1():
    s = np.array([1e-4, 0.01, 0.3, 1.0, 5.0, 50.0, 1000.0], np.float32)
    x = np.array([-5.0, -0.7, -1e-3, -1e-7, 1e-5, 0.2, 0.9], np.float32)
    out = mu(10, _rec(len(s), c0=s, c1=x))
    s64 = s.astype(np.float64)
    np.testing.assert_allclose(out[:, 0], 2 * np.pi * (1 - np.exp(-2 * s64)) / s64, rtol=2e-6)
    np.testing.assert_allclose(out[:, 1], np.expm1(x.astype(np.float64)), rtol=2e-6)
    np.testing.assert_allclose(out[:, 2], np.expm1(x.astype(np.float64)) / x.astype(np.float64), rtol=2e-6)
    # the integral by quadrature for one sharpness
    th, ph, w = _hemisphere_grid(200000, 1, lo=-1.0)
    assert abs(np.sum(np.exp(3.0 * (np.cos(th) - 1))) * w - out[3, 0] * 0 - 2 * np.pi * (1 - np.exp(-6.0)) / 3.0) < 1e-4


@pytest.mark.parametrize("s", [0.05, 0.8, 4.0, 30.0])
def test_clamped_cosine_product_against_quadrature(s):
    # integral over the sphere of exp(s (v.a - 1)) max(0, v.n) / pi, cos = a.n
    cosines = np.linspace(-0.95, 0.95, 9)
    out = mu(11, _rec(len(cosines), c0=cosines, c1=s))
    th, ph, w = _hemisphere_grid(4000, 256, lo=-1.0)
    v = _dirs(th, ph)
    sg_lower = np.exp(s * (-1 - 1.0))
    for c, got in zip(cosines, out[:, 0]):
        a = np.array([np.sqrt(1 - c * c), 0.0, c])             # n = +z
        ref = np.sum(np.exp(s * (v @ a - 1)) * np.maximum(0.0, v[:, 2])) * w / np.pi
        # a fitted approximation (Tokuyoshi 2024): a few percent of the SG's
        # total clamped-cosine mass at worst
        scale = 2 * np.pi * (1 - np.exp(-2 * s)) / s / np.pi
        assert abs(got - ref) <= 0.03 * scale + 1e-6, (c, got, ref)
    # the upper and lower bounds are the cosine integrals at cos = +1 and -1
    up, lo = out[0, 1], out[0, 2]
    a = np.array([0.0, 0.0, 1.0])
    ref_up = np.sum(np.exp(s * (v @ a - 1)) * np.maximum(0.0, v[:, 2])) * w / (2 * np.pi)
    ref_lo = np.sum(np.exp(s * (-v @ a - 1)) * np.maximum(0.0, v[:, 2])) * w / (2 * np.pi)
    assert abs(up - ref_up) < 1e-3 * max(ref_up, 1e-3) + 1e-6
    assert abs(lo - ref_lo) < 1e-3 * max(ref_lo, 1e-6) + 1e-6 + 0 * sg_lower


def test_sg_product_is_an_sg():
    rng = np.random.default_rng(6)
    n = 200
    a1 = rng.normal(size=(n, 3)); a1 /= np.linalg.norm(a1, axis=1, keepdims=True)
    th2, ph2 = np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * np.pi, n)
    a2 = _dirs(th2, ph2)
    s1, s2 = rng.uniform(0.1, 20, n), rng.uniform(0.1, 20, n)
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0:3], rec[:, 3], rec[:, 4], rec[:, 5], rec[:, 6] = a1, s1, th2, ph2, s2
    out = mu(12, rec).astype(np.float64)
    sharp, logA, d1, d2 = out.T
    # rebuild the product lobe's axis from its projections on a1 and a2 and
    # compare G1(v) G2(v) with A G(v) at random v
    for i in range(n):
        v = rng.normal(size=(5, 3)); v /= np.linalg.norm(v, axis=1, keepdims=True)
        axis = s1[i] * a1[i] + s2[i] * a2[i]
        axis /= np.linalg.norm(axis)
        assert abs(axis @ a1[i] - d1[i]) < 2e-4 and abs(axis @ a2[i] - d2[i]) < 2e-4
        lhs = s1[i] * (v @ a1[i] - 1) + s2[i] * (v @ a2[i] - 1)
        rhs = logA[i] + sharp[i] * (v @ axis - 1)
        np.testing.assert_allclose(rhs, lhs, rtol=1e-4, atol=2e-4)


def test_vmf_sharpness_axis_length_are_inverse():
    # the cubic solve (Peters 2016) inverts Banerjee's approximation exactly up
    # to FP32 conditioning, which degrades as the axis length approaches 1
    s = np.array([0.05, 0.5, 2.0, 10.0, 100.0, 1000.0], np.float32)
    l = mu(13, _rec(len(s), c1=s))[:, 1]
    back = mu(13, _rec(len(s), c0=l))[:, 0]
    np.testing.assert_allclose(back[:5], s[:5], rtol=2e-3)
    np.testing.assert_allclose(back[5], s[5], rtol=1e-2)
    # Banerjee's approximation of the exact mean cosine coth(s) - 1/s of a
    # vMF lobe is good to a few percent
    s64 = s.astype(np.float64)
    np.testing.assert_allclose(l, 1 / np.tanh(s64) - 1 / s64, rtol=2e-2, atol=1e-5)


def test_gaussian_filter_radial_distribution():
    rng = np.random.default_rng(7)
    n = 200000
    out = mu(14, _rec(n, c0=rng.random(n), c1=rng.random(n)))
    r = np.hypot(out[:, 0], out[:, 1]).astype(np.float64)
    sigma, rad = 0.25, 1.5
    assert r.max() <= rad + 1e-5
    rs = np.sort(r)
    cdf = (1 - np.exp(-rs ** 2 / (2 * sigma ** 2))) / (1 - np.exp(-rad ** 2 / (2 * sigma ** 2)))
    ks = np.max(np.abs(cdf - (np.arange(1, n + 1) / n)))
    assert ks < 1.63 / np.sqrt(n) * 1.5, ks
    # isotropic: mean angle vector vanishes
    ang = np.arctan2(out[:, 1], out[:, 0])
    assert abs(np.cos(ang).mean()) < 0.01 and abs(np.sin(ang).mean()) < 0.01


def test_refraction_obeys_snell():
    rng = np.random.default_rng(8)
    n = 5000
    v = rng.normal(size=(n, 3)); v /= np.linalg.norm(v, axis=1, keepdims=True)
    eta = rng.uniform(0.5, 2.0, n)
    t = mu(15, _rec(n, c0=v[:, 0], c1=v[:, 1], c2=v[:, 2], c3=eta)).astype(np.float64)[:, :3]
    sin_i = np.linalg.norm(v[:, :2], axis=1)
    tir = (sin_i / eta) >= 1.0
    assert np.all(np.all(t[tir] == 0.0, axis=1))
    ok = ~tir
    np.testing.assert_allclose(np.linalg.norm(t[ok], axis=1), 1.0, atol=2e-5)
    np.testing.assert_allclose(np.linalg.norm(t[ok, :2], axis=1), sin_i[ok] / eta[ok], atol=3e-5)
    # transmitted to the other side, tangential direction reversed with the incident vector convention
    assert np.all(np.sign(t[ok, 2]) == -np.sign(v[ok, 2]))


def test_power_heuristic_and_its_overflow_guard():
    # f^2 / (f^2 + g^2) (src/device/mis.h:14-21); an infinite f^2 (an edge-on
    # light sample has an infinite solid-angle pdf) gives 1, not inf/inf
    f = np.array([1.0, 2.0, 0.5, 1e20, np.inf, 3.0, 0.0], np.float32)
    g = np.array([1.0, 1.0, 2.0, 1.0, 5.0, np.inf, 1.0], np.float32)
    w = mu(17, _rec(len(f), c0=f, c1=g))[:, 0]
    ref = np.array([0.5, 0.8, 1.0 / 17.0, 1.0, 1.0, 0.0, 0.0])
    np.testing.assert_allclose(w, ref, rtol=1e-6)
    assert np.all(np.isfinite(w))


def test_pow5_is_the_correctly_rounded_fifth_power():
    # the (1 - mu)^5 of the Schlick/F82 Fresnel terms: a double product rounded
    # once, equal to the exactly rounded x^5 (rational arithmetic) for every
    # tested x, and Schlick's f0 + (1 - f0) (1 - mu)^5 built on it
    from fractions import Fraction
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random(4000), np.linspace(0, 1, 1001), [1e-8, 0.5, 1.0 - 2 ** -24]]).astype(np.float32)
    f0 = rng.random(len(x)).astype(np.float32)
    out = mu(18, _rec(len(x), c0=x, c1=f0))
    exact = np.array([float(Fraction(float(v)) ** 5) for v in x]).astype(np.float32)
    assert np.array_equal(out[:, 0], exact)
    mu_ = (np.float32(1) - x).astype(np.float32)       # schlick(f0, mu) with mu = x uses (1 - x)^5
    p5 = np.array([float(Fraction(float(v)) ** 5) for v in mu_]).astype(np.float32)
    ref = (f0 + (np.float32(1) - f0) * p5).astype(np.float32)
    assert np.array_equal(out[:, 1], ref)
