"""Pins of the oracle's math against closed forms and library routines."""
import math

import numpy as np
import pytest
import torch

import oracle_lib as O


def test_erf_fit_against_exact():
    x = np.linspace(-6, 6, 120001).astype(np.float32)
    out = np.zeros_like(x)
    O.lib().orc_sg_erf(O.p(x), O.p(out), len(x))
    exact = np.array([math.erf(float(v)) for v in x])
    assert np.max(np.abs(out - exact)) < 2.5e-7
    assert np.array_equal(out, -out[::-1])     # odd


def test_sobol_table_matches_joe_kuo_sobol():
    """Unscrambled dims 0-3 reproduce torch's Sobol engine (Joe-Kuo numbers),
    which draws in Gray-code order."""
    t = O.sobol_table()

    def brev(v):
        return int("{:032b}".format(int(v))[::-1], 2)

    pts = torch.quasirandom.SobolEngine(4, scramble=False).draw(1024).numpy()
    for n in range(1024):
        i = n ^ (n >> 1)
        for d in range(4):
            r = 0
            for k in range(32):
                if (i >> k) & 1:
                    r ^= brev(t[d][k])
            assert abs(r / 2 ** 32 - pts[n, d]) < 1e-9


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_owen_scrambled_sobol_is_stratified(kind):
    """For 2^k consecutive samples of one pixel/dimension every elementary
    interval of 1/2^k holds exactly one point per coordinate, and every
    2D interval (2^a x 2^b, a+b=k) of coordinates 0,1 holds one point."""
    k = 8
    n = 1 << k
    for hsh, off in [(0x1234567, 16), (0xdeadbeef, 48), (7, 160)]:
        rec = np.zeros((n, 5), np.uint32)
        rec[:, 0] = np.arange(n)
        rec[:, 1], rec[:, 2], rec[:, 3], rec[:, 4] = hsh, off, 2, kind
        pts = O.sobol(rec)
        assert np.all((pts >= 0) & (pts < 1))
        for c in range(kind):
            cells = np.floor(pts[:, c] * n).astype(int)
            assert np.array_equal(np.sort(cells), np.arange(n))
        if kind >= 2:
            for a in range(k + 1):
                cx = np.floor(pts[:, 0] * (1 << a)).astype(int)
                cy = np.floor(pts[:, 1] * (1 << (k - a))).astype(int)
                assert len(set(zip(cx, cy))) == n


def test_wavelength_pdf_normalised_and_matches_histogram():
    # pdf integrates to ~1 over [360, 830] (spectral.h:293-299)
    lam = np.linspace(360, 830, 470001)
    _, pdf = O.wavelengths(np.zeros(1, np.float32))
    c = np.cosh(0.0072 * (lam - 538.0))
    dens = 0.0039398042 / (c * c)
    assert abs(np.trapezoid(dens, lam) - 1.0) < 2e-3
    u = ((np.arange(200000) + 0.5) / 200000).astype(np.float32)
    l, p = O.wavelengths(u)
    hist, edges = np.histogram(l[:, 0], bins=47, range=(360, 830), density=True)
    mid = 0.5 * (edges[1:] + edges[:-1])
    cm = np.cosh(0.0072 * (mid - 538.0))
    np.testing.assert_allclose(hist, 0.0039398042 / (cm * cm), rtol=2e-2)
    # pdf returned is the density at the sampled wavelength
    cl = np.cosh(0.0072 * (l.astype(np.float64) - 538.0))
    np.testing.assert_allclose(p, 0.0039398042 / (cl * cl), rtol=1e-5)


def test_octahedral_round_trip():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(20000, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    enc = np.zeros(len(v), np.uint32)
    O.lib().orc_oct_encode_u32(O.p(v), O.p(enc), len(v))
    dec = np.zeros_like(v)
    O.lib().orc_oct_decode_u32(O.p(enc), O.p(dec), len(v))
    a = v.astype(np.float64) / np.linalg.norm(v.astype(np.float64), axis=1, keepdims=True)
    b = dec.astype(np.float64) / np.linalg.norm(dec.astype(np.float64), axis=1, keepdims=True)
    ang = 2.0 * np.arcsin(np.linalg.norm(a - b, axis=1) / 2.0)   # well conditioned for small angles
    assert ang.max() < 1.5e-4
    axes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], np.float32)
    O.lib().orc_oct_encode_u32(O.p(axes), O.p(enc), 6)
    O.lib().orc_oct_decode_u32(O.p(enc), O.p(dec), 6)
    assert np.max(np.abs(dec[:6] - axes)) < 1e-4
