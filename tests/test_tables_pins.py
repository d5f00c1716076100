"""Pins of the spectral tables and LUTs that both the product and the oracle
read (ratbite_amd/data/tables.npz, made by ratbite_amd/data/gen_tables.py),
and of the oracle's display transform (orc_resolve), against what defines
them independently of this project's code:

- the CIE 1931 2-degree colour-matching functions at tabulated wavelengths
  and the integral of y-bar (the tables hold a multi-lobe Gaussian fit);
- Planck's law for the 4000 K blackbody emitter, normalised like
  src/core/spectral.cpp:1418-1424 (sum of S * y-bar = the CIE y-bar integral);
- the IEC 61966-2-1 XYZ -> linear sRGB matrix (src/core/color.h:60-99 derives
  it from the primaries and the D65 white point);
- the RGB -> spectrum basis: reflectances that render as the sRGB primaries
  under a white (D65-balanced) illuminant and sum to one at every wavelength;
- the conductor albedo LUT (GGX, F = 1, height-correlated Smith) against a
  deterministic quadrature over half vectors;
- AgX (src/kernels/image.cu:27-60): its two matrices are inverses, it keeps
  neutral colours neutral (a transposed matrix would tint them by ~10%),
  maps black to 0 and is monotone in exposure.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import scenes

LAM = np.arange(360, 831, dtype=np.float64)

# CIE 1931 2-degree standard observer (CIE 15:2004 table, 1 nm data)
CIE1931 = {
    400: (0.01431, 0.000396, 0.06785),
    450: (0.33620, 0.03800, 1.77211),
    500: (0.00490, 0.32300, 0.27200),
    550: (0.43345, 0.99500, 0.00875),
    555: (0.51205, 1.00000, 0.00575),
    600: (1.06220, 0.63100, 0.00080),
    650: (0.28350, 0.10700, 0.00000),
    700: (0.011359, 0.004102, 0.00000),
}
CIE_Y_INTEGRAL = 106.856895      # sum of the 1 nm y-bar over 360..830 nm


@pytest.fixture(scope="module")
def tables():
    return scenes.load_tables()


def _cmf(t):
    return np.asarray(t["xyz"], np.float64).reshape(3, 471)


def test_cmfs_against_cie1931(tables):
    xyz = _cmf(tables)
    worst = 0.0
    for lam, ref in CIE1931.items():
        got = xyz[:, lam - 360]
        worst = max(worst, float(np.abs(got - ref).max()))
        assert np.allclose(got, ref, atol=0.015), (lam, got, ref)
    print("largest CMF deviation from CIE 1931: %.4f" % worst)
    # each CMF integrates to the y-bar integral within 0.2% (equal-energy white is x = y = 1/3)
    assert np.allclose(xyz.sum(axis=1), CIE_Y_INTEGRAL, rtol=2e-3)


def test_blackbody_is_planck_normalised(tables):
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lam_m = LAM * 1e-9
    b = 1.0 / (lam_m ** 5 * np.expm1(h * c / (lam_m * k * 4000.0)))
    bb = np.asarray(tables["blackbody4000"], np.float64)
    ybar = _cmf(tables)[1]
    # normalisation: sum(S * ybar) = the CIE y-bar integral
    assert abs(np.sum(bb * ybar) - CIE_Y_INTEGRAL) < 1e-4 * CIE_Y_INTEGRAL
    shape = bb / b
    assert np.allclose(shape, shape[200], rtol=1e-5)
    # Wien's displacement law: the peak of a 4000 K spectrum lies at b / T = 724.4 nm
    assert abs(LAM[np.argmax(bb)] - 2.897771955e-3 / 4000.0 * 1e9) <= 1.0


def test_xyz_to_srgb_is_the_iec_matrix(tables):
    iec = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])
    m = np.asarray(tables["xyz_to_srgb"], np.float64).reshape(3, 3)
    assert np.allclose(m, iec, atol=2e-3), m
    d65 = np.array([0.95047, 1.0, 1.08883])
    assert np.allclose(m @ d65, 1.0, atol=1e-3)
    # primaries: the xy chromaticities of the sRGB red, green and blue
    inv = np.linalg.inv(m)
    for j, (x, y) in enumerate(((0.64, 0.33), (0.30, 0.60), (0.15, 0.06))):
        X, Y, Z = inv[:, j]
        assert abs(X / (X + Y + Z) - x) < 1e-3 and abs(Y / (X + Y + Z) - y) < 1e-3


def test_rgb_basis_reproduces_the_primaries(tables):
    basis = np.asarray(tables["basis"], np.float64).reshape(3, 471)
    assert np.all(basis >= -1e-9)
    assert np.allclose(basis.sum(axis=0), 1.0, atol=1e-6)
    xyz = _cmf(tables)
    m = np.asarray(tables["xyz_to_srgb"], np.float64).reshape(3, 3)
    # a smooth illuminant near 6504 K that renders the constant reflectance 1
    # as sRGB white: the Planckian tilted by a quadratic, solved here
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lam_m = LAM * 1e-9
    P = 1.0 / (lam_m ** 5 * np.expm1(h * c / (lam_m * k * 6504.0)))
    x = (LAM - 555.0) / 100.0
    cols = np.stack([m @ (xyz * P * x ** p).sum(axis=1) for p in range(3)], axis=1)
    c1, c2, kk = np.linalg.solve(np.concatenate([cols[:, 1:], -np.ones((3, 1))], 1), -cols[:, 0])
    illum = P * (1 + c1 * x + c2 * x * x) / kk
    assert np.allclose(m @ (xyz * illum).sum(axis=1), 1.0, atol=1e-9)
    rgb = m @ (xyz * illum) @ basis.T       # column j: linear sRGB of basis spectrum j
    print("basis -> sRGB\n", np.round(rgb, 4))
    # each basis spectrum renders as its own primary
    assert np.allclose(rgb, np.eye(3), atol=3e-3)
    # the spectra are ordered blue < green < red in wavelength
    centroid = (basis * LAM).sum(axis=1) / basis.sum(axis=1)
    assert centroid[2] < centroid[1] < centroid[0]


def _ggx_albedo(mu, alpha, n=4000):
    """Directional albedo of the GGX reflection lobe with F = 1 and the
    height-correlated Smith term, by midpoint quadrature over half vectors:
    E = int D(h) G2(o, i) |o.h| / mu_o dh, i the mirror of o about h."""
    def lam(c):
        t2 = max(0.0, 1.0 - c * c) / max(c * c, 1e-300)
        return (math.sqrt(1.0 + alpha * alpha * t2) - 1.0) / 2.0

    o = np.array([math.sqrt(1.0 - mu * mu), 0.0, mu])
    th = (np.arange(n) + 0.5) / n * (math.pi / 2)
    ph = (np.arange(2 * n // 4) + 0.5) / (2 * n // 4) * 2 * math.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    dh = (math.pi / 2 / n) * (2 * math.pi / len(ph)) * np.sin(TH)
    hx, hy, hz = np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)
    a2 = alpha * alpha
    D = a2 / (math.pi * (hz * hz * (a2 - 1.0) + 1.0) ** 2)
    oh = o[0] * hx + o[2] * hz
    iz = 2.0 * oh * hz - o[2]
    t2i = np.maximum(0.0, 1.0 - iz * iz) / np.maximum(iz * iz, 1e-300)
    lam_i = (np.sqrt(1.0 + a2 * t2i) - 1.0) / 2.0
    G2 = np.where((iz > 0) & (oh > 0), 1.0 / (1.0 + lam(mu) + lam_i), 0.0)
    return float(np.sum(D * G2 * np.abs(oh) / mu * dh))


def test_conductor_albedo_lut_against_quadrature(tables):
    lut = np.asarray(tables["conductor"], np.float64).reshape(32, 32)   # [roughness][cos theta_o]
    worst = 0.0
    for j in (10, 16, 24, 31):          # roughness texel centres (alpha = r^2 >= 0.1)
        r = (j + 0.5) / 32
        for i in (2, 8, 16, 31):
            mu = (i + 0.5) / 32
            e = _ggx_albedo(mu, r * r)
            worst = max(worst, abs(lut[j, i] - e))
            assert abs(lut[j, i] - e) < 0.02, (j, i, lut[j, i], e)
    print("largest |LUT - quadrature| of the conductor albedo: %.4f" % worst)


def test_conductor_quadrature_can_fail():
    """The quadrature above separates albedos: rougher lobes lose more."""
    assert _ggx_albedo(0.5, 0.05) > _ggx_albedo(0.5, 0.5) + 0.05
    assert abs(_ggx_albedo(0.9, 0.01 ** 2 + 0.02) - 1.0) < 0.02


# ------------------------------------------------------------------ AgX
IDENT = np.eye(3, dtype=np.float32)


def _resolve_rgb(rgb_linear, m3=IDENT, exposure=1.0):
    """orc_resolve of one pixel whose accumulated value is rgb_linear (the
    colour-space matrix is the identity, so the input is linear RGB)."""
    acc = np.zeros((1, len(rgb_linear), 4))
    acc[0, :, :3] = rgb_linear
    acc[0, :, 3] = 1.0
    return O.resolve(acc, m3, exposure)[0, :, :3].astype(int)


def test_agx_matrices_are_inverses():
    agx = np.array([[0.842479062253094, 0.0423282422610123, 0.0423756549057051],
                    [0.0784335999999992, 0.878468636469772, 0.0784336],
                    [0.0792237451477643, 0.0791661274605434, 0.879142973793104]])
    inv = np.array([[1.1968790051201738155, -0.052896851757456180321, -0.052971635514443794537],
                    [-0.098020881140136776078, 1.1519031299041727435, -0.098043450117124120312],
                    [-0.099029744079720471434, -0.098961176844843346553, 1.1510736726411610622]])
    assert np.allclose(agx @ inv, np.eye(3), atol=1e-6)


def test_agx_keeps_neutrals_neutral_and_black_black():
    greys = np.array([[g, g, g] for g in (1e-4, 0.01, 0.18, 0.5, 1.0, 4.0, 16.0)])
    out = _resolve_rgb(greys)
    print("greys ->", out[:, 0])
    assert np.all(out.max(axis=1) - out.min(axis=1) <= 1)
    assert np.all(np.diff(out[:, 0]) >= 0) and out[-1, 0] > out[0, 0] + 150
    assert np.all(_resolve_rgb(np.zeros((1, 3))) == 0)
    # a coloured input keeps its dominant channel
    c = _resolve_rgb(np.array([[0.6, 0.1, 0.05], [0.05, 0.5, 0.1], [0.05, 0.1, 0.7]]))
    assert list(c.argmax(axis=1)) == [0, 1, 2]


def test_agx_monotone_in_exposure():
    rgb = np.array([[0.3, 0.2, 0.1]])
    vals = [_resolve_rgb(rgb, exposure=e)[0] for e in (0.25, 0.5, 1.0, 2.0, 4.0, 8.0)]
    assert all(np.all(b >= a) for a, b in zip(vals, vals[1:]))


def test_agx_channel_without_a_logarithm_is_black():
    """A saturated XYZ colour leaves some linear-sRGB channels negative; they
    have no logarithm, and the clamp sends them to the bottom of the AgX range
    (the product's fmax / fmin clamp) instead of propagating NaN."""
    m = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]], np.float32)
    acc = np.zeros((1, 2, 4))
    acc[0, :, 3] = 1.0
    acc[0, 0, :3] = (0.0, 1.0, 0.0)
    acc[0, 1, :3] = (0.0, 0.0, 1.0)
    out = O.resolve(acc, m, 1.0)[0, :, :3].astype(int)
    assert out[0, 0] == 0 and out[0, 2] == 0 and out[0, 1] > 200
    assert out[1, 0] == 0 and out[1, 2] > 200
