"""Generates the spectral tables and BxDF look-up tables this project ships.

The renderer takes these as inputs (rbh_upload_color_tables / rbh_upload_luts),
so any caller may substitute its own. The project's tables are produced here
from published definitions, not copied:

  * CIE 1931 2-degree colour-matching functions at 1 nm, 360..830 nm, from the
    multi-lobe analytic fit of Wyman, Sloan and Shirley, "Simple Analytic
    Approximations to the CIE XYZ Color Matching Functions", JCGT 2(2), 2013.
  * RGB -> spectrum basis (three non-negative spectra summing to one) in the
    spirit of Mallett and Yuksel, "Spectral Primary Decomposition for
    Rendering with sRGB Reflectance", EGSR 2019: the smoothest basis whose
    reflectance under a 6504 K Planckian illuminant reproduces the linear
    sRGB primaries (solved as a bounded linear least-squares problem).
  * sRGB XYZ->RGB matrix built from the primaries and the D65 white point
    exactly as src/core/color.h:60-99 builds it.
  * Microfacet albedo LUTs (32 x 32 and 32^3, texel centres at (i+0.5)/32,
    axis order x = |cos theta_o|, y = roughness, z = sqrt|f0Sr|) by Monte
    Carlo integration of the reference's GGX model (height-correlated Smith
    G, alpha = roughness^2, Schlick Fresnel with f0 = f0Sr^2 and total
    internal reflection), following Turquin, "Practical multiple scattering
    compensation for microfacet models", 2019.
  * Sheen LTC table (A, B, albedo, 0) per (|cos theta_o|, alpha): a fit of the
    LTC form used by src/device/surface.h:53-88 (Zeltner et al. 2022) to the
    "Charlie" sheen BRDF of Estevez and Kulla (2017) with Neubelt-Pettineo
    visibility.

Run `python -m ratbite_amd.data.gen_tables` to regenerate tables.npz.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "tables.npz")
LAMBDA = np.arange(360, 831, dtype=np.float64)
K_CIE_Y_INTEGRAL = 106.856895  # src/core/spectral_settings.h:9


def _g(x, mu, s1, s2):
    s = np.where(x < mu, s1, s2)
    t = (x - mu) / s
    return np.exp(-0.5 * t * t)


def cie_xyz(lam=LAMBDA):
    x = 1.056 * _g(lam, 599.8, 37.9, 31.0) + 0.362 * _g(lam, 442.0, 16.0, 26.7) - 0.065 * _g(lam, 501.1, 20.4, 26.2)
    y = 0.821 * _g(lam, 568.8, 46.9, 40.5) + 0.286 * _g(lam, 530.9, 16.3, 31.1)
    z = 1.217 * _g(lam, 437.0, 11.8, 36.0) + 0.681 * _g(lam, 459.0, 26.0, 13.8)
    return np.stack([x, y, z])


def planck(lam_nm, temp):
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    l = lam_nm * 1e-9
    return (2 * h * c * c) / (l ** 5 * (np.exp(h * c / (l * k * temp)) - 1.0))


def normalize_illuminant(spd, ybar):
    """Scale so that sum(S * ybar) over 360..830 equals K_CIE_Y_INTEGRAL, the
    normalisation of src/core/spectral.cpp:1418-1424."""
    return spd * (K_CIE_Y_INTEGRAL / np.sum(spd * ybar))


def xyz_to_srgb_matrix(white_xy=(0.31271, 0.32902)):
    """src/core/color.h:60-99 for the sRGB primaries."""
    def xyY(x, y, Y=1.0):
        return np.array([x * Y / y, Y, (1 - x - y) * Y / y])
    r, g, b = xyY(0.64, 0.33), xyY(0.30, 0.60), xyY(0.15, 0.06)
    rgb = np.stack([r, g, b], axis=1)      # columns = primaries
    wp = xyY(*white_xy)
    c = np.linalg.solve(rgb, wp)
    return np.linalg.inv(rgb @ np.diag(c))  # XYZ -> linear sRGB


def white_illuminant(cmf):
    """A smooth illuminant at 1 nm whose XYZ under `cmf` maps to linear sRGB
    (1, 1, 1): the 6504 K Planckian times 1 + c1 x + c2 x^2, x = (lambda -
    555 nm) / 100 nm, with c1, c2 and the scale solved from the three
    conditions."""
    P = planck(LAMBDA, 6504.0)
    x = (LAMBDA - 555.0) / 100.0
    M = xyz_to_srgb_matrix()
    cols = np.stack([M @ (cmf * P * x ** p).sum(axis=1) for p in range(3)], axis=1)   # 3 x 3
    # k (1, 1, 1) = cols @ (1, c1, c2)  ->  cols[:, 1:] @ (c1, c2) - k 1 = -cols[:, 0]
    sys_ = np.concatenate([cols[:, 1:], -np.ones((3, 1))], axis=1)
    c1, c2, k = np.linalg.solve(sys_, -cols[:, 0])
    return P * (1.0 + c1 * x + c2 * x * x) / k


def rgb_basis(cmf, weight=1e4):
    """Smoothest non-negative partition of unity b_r + b_g + b_b = 1 whose
    reflectances render as the sRGB primaries under a D65-white illuminant
    (white_illuminant: the 6504 K Planckian tilted by a quadratic in
    wavelength until it is exactly the sRGB white under these colour-matching
    functions; the Planckian alone is 4% off white, below the daylight locus).

    Solved on a 5 nm grid over the whole 360..830 nm range the renderer
    samples (so that the 1 nm resampling keeps the conditions to ~1e-3) as
    one bounded linear least-squares
    problem (scipy's bounded-variable least squares): the rows are the first
    differences of each spectrum (smoothness) and, scaled by `weight`, the
    3 x 3 primary conditions and the partition of unity at every sample, with
    every value bounded to [0, 1]."""
    from scipy.optimize import lsq_linear
    lam5 = np.arange(360, 831, 5, dtype=np.float64)
    n = len(lam5)
    cmf5 = np.stack([np.interp(lam5, LAMBDA, c) for c in cmf])
    illum = np.interp(lam5, LAMBDA, white_illuminant(cmf))
    M = xyz_to_srgb_matrix()
    A = M @ (cmf5 * illum) * (len(LAMBDA) - 1) / (len(lam5) - 1)   # 3 x n: spectrum -> linear sRGB
    rows, rhs = [], []
    D = np.diff(np.eye(n), axis=0)          # (n-1) x n first differences
    for i in range(3):
        blk = np.zeros((n - 1, 3 * n))
        blk[:, i * n:(i + 1) * n] = D
        rows.append(blk)
        rhs.append(np.zeros(n - 1))
    for i in range(3):                      # basis i -> primary e_i
        for j in range(3):
            r = np.zeros((1, 3 * n))
            r[0, i * n:(i + 1) * n] = A[j]
            rows.append(weight * r)
            rhs.append(np.array([weight * float(i == j)]))
    S = np.hstack([np.eye(n)] * 3)          # partition of unity
    rows.append(weight * S)
    rhs.append(weight * np.ones(n))
    res = lsq_linear(np.vstack(rows), np.concatenate(rhs), bounds=(0.0, 1.0), method="bvls", tol=1e-14,
                     max_iter=10000)
    if not res.success:
        raise RuntimeError("RGB basis solve failed: %s" % res.message)
    b5 = res.x.reshape(3, n)
    # piecewise-linear resampling at 1 nm (src/core/spectral.cpp:1282-1286
    # and 1344-1380 resample the reference's basis the same way)
    return np.stack([np.interp(LAMBDA, lam5, b5[i]) for i in range(3)])


# ------------------------------------------------------------------ LUTs
def _lambda_ggx(tan2, alpha):
    return (np.sqrt(1.0 + alpha * alpha * tan2) - 1.0) / 2.0


def _tan2(cos):
    c2 = cos * cos
    return np.maximum(0.0, 1.0 - c2) / np.maximum(c2, 1e-30)


def _vndf_sample(wo, alpha, u1, u2):
    """Heitz 2018 visible-normal sampling (isotropic alpha), wo.z > 0."""
    vh = np.stack([alpha * wo[0], alpha * wo[1], wo[2]])
    vh = vh / np.linalg.norm(vh, axis=0)
    lensq = vh[0] ** 2 + vh[1] ** 2
    inv = np.where(lensq > 0, 1.0 / np.sqrt(np.maximum(lensq, 1e-30)), 0.0)
    t1 = np.where(lensq > 0, np.stack([-vh[1] * inv, vh[0] * inv, np.zeros_like(lensq)]),
                  np.stack([np.ones_like(lensq), np.zeros_like(lensq), np.zeros_like(lensq)]))
    t2 = np.cross(vh, t1, axis=0)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    p1 = r * np.cos(phi)
    p2 = r * np.sin(phi)
    s = 0.5 * (1.0 + vh[2])
    p2 = (1.0 - s) * np.sqrt(np.maximum(0.0, 1.0 - p1 * p1)) + s * p2
    nh = p1 * t1 + p2 * t2 + np.sqrt(np.maximum(0.0, 1.0 - p1 * p1 - p2 * p2)) * vh
    m = np.stack([alpha * nh[0], alpha * nh[1], np.maximum(1e-6, nh[2])])
    return m / np.linalg.norm(m, axis=0)


def _schlick(f0, c):
    return f0 + (1.0 - f0) * (1.0 - c) ** 5


def microfacet_luts(n=32, spp=2048, seed=1):
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n
    mu = c[None, :, None]          # x
    rough = c[:, None, None]       # y
    alpha = rough * rough
    u1 = rng.random((1, 1, spp))
    u2 = rng.random((1, 1, spp))
    wo = np.stack(np.broadcast_arrays(np.sqrt(1 - mu * mu), np.zeros_like(mu), mu))
    wo = np.stack([np.broadcast_to(w, (n, n, spp)) for w in wo])
    al = np.broadcast_to(alpha, (n, n, spp))
    m = _vndf_sample(wo, al, np.broadcast_to(u1, (n, n, spp)), np.broadcast_to(u2, (n, n, spp)))
    dot_om = np.sum(wo * m, axis=0)
    wr = 2 * dot_om * m - wo
    lam_o = _lambda_ggx(_tan2(wo[2]), al)
    g1 = 1.0 / (1.0 + lam_o)
    lam_r = _lambda_ggx(_tan2(wr[2]), al)
    g2r = np.where(wr[2] > 0, 1.0 / (1.0 + lam_o + lam_r), 0.0)
    cond = np.mean(g2r / g1, axis=2)                         # F = 1
    diel = {}
    for side in ("outer", "inner"):
        out_t = np.zeros((n, n, n))
        out_r = np.zeros((n, n, n))
        for k, z in enumerate(c):
            f0sr = z * z
            eta = (1 + f0sr) / (1 - f0sr) if side == "outer" else (1 - f0sr) / (1 + f0sr)
            f0 = f0sr * f0sr
            F = _schlick(f0, np.abs(dot_om))
            sin2t = np.maximum(0.0, 1.0 - dot_om * dot_om) / (eta * eta)
            tir = sin2t >= 1.0
            F = np.where(tir, 1.0, F)
            # refraction through m (wo above, m up): wt = -wo/eta + (cos/eta - sqrt(1-sin2t)) m
            cost = np.sqrt(np.maximum(0.0, 1.0 - sin2t))
            wt = -wo / eta + (dot_om / eta - cost) * m
            wt = wt / np.maximum(np.linalg.norm(wt, axis=0), 1e-30)
            lam_t = _lambda_ggx(_tan2(wt[2]), al)
            g2t = np.where((wt[2] < 0) & ~tir, 1.0 / (1.0 + lam_o + lam_t), 0.0)
            refl = F * g2r / g1
            trans = (1.0 - F) * g2t / g1
            out_r[k] = np.mean(refl, axis=2)
            out_t[k] = np.mean(refl + trans, axis=2)
        diel[side] = (np.clip(out_t, 1e-4, 1.0), np.clip(out_r, 0.0, 1.0))
    return (cond.astype(np.float32), diel["outer"][0].astype(np.float32), diel["inner"][0].astype(np.float32),
            diel["outer"][1].astype(np.float32), diel["inner"][1].astype(np.float32))


def sheen_ltc(n=32, res=48):
    """Fit (A, B) of the inverse LTC matrix [[A,0,B],[0,A,0],[0,0,1]] to the
    normalised Charlie sheen lobe; albedo by quadrature."""
    c = (np.arange(n) + 0.5) / n
    # hemisphere quadrature grid
    th = (np.arange(res) + 0.5) / res * (np.pi / 2)
    ph = (np.arange(2 * res) + 0.5) / (2 * res) * 2 * np.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    dw = (np.pi / 2 / res) * (2 * np.pi / (2 * res)) * np.sin(TH)
    wi = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)])
    out = np.zeros((n, n, 4), np.float32)
    A_grid = np.exp(np.linspace(np.log(0.01), np.log(1.5), 60))
    B_grid = np.linspace(-0.6, 0.6, 25)
    for j, alpha in enumerate(c):          # y = alpha
        inv_r = 1.0 / max(alpha, 1e-3)
        for i, muo in enumerate(c):        # x = |cos theta_o|
            wo = np.array([np.sqrt(1 - muo * muo), 0.0, muo])
            h = wi + wo[:, None, None]
            h = h / np.linalg.norm(h, axis=0)
            sin_h = np.sqrt(np.maximum(0.0, 1.0 - h[2] ** 2))
            D = (2.0 + inv_r) * sin_h ** inv_r / (2 * np.pi)
            V = 1.0 / (4.0 * (wi[2] + muo - wi[2] * muo))
            fcos = D * V * wi[2]
            albedo = np.sum(fcos * dw)
            target = fcos / max(albedo, 1e-12)
            best = (1e30, 1.0, 0.0)
            for A in A_grid:
                for B in B_grid:
                    w0 = np.stack([A * wi[0] + B * wi[2], A * wi[1], wi[2]])
                    L = np.linalg.norm(w0, axis=0)
                    pdf = np.maximum(w0[2] / L, 0.0) / np.pi * (A * A) / L ** 3
                    err = np.sum((pdf - target) ** 2 * dw)
                    if err < best[0]:
                        best = (err, A, B)
            # the Neubelt-Pettineo visibility is not energy conserving at
            # grazing angles for thin fibres; clamp the layer weight to 1
            out[j, i] = (best[1], best[2], min(albedo, 1.0), 0.0)
    return out


def generate():
    cmf = cie_xyz()
    basis = rgb_basis(cmf)
    bb = normalize_illuminant(planck(LAMBDA, 4000.0), cmf[1])
    cond, dout, din, rout, rin = microfacet_luts()
    sheen = sheen_ltc()
    np.savez_compressed(OUT, xyz=cmf.astype(np.float32), basis=basis.astype(np.float32),
                        blackbody4000=bb.astype(np.float32), xyz_to_srgb=xyz_to_srgb_matrix().astype(np.float32),
                        conductor=cond, diel_outer=dout, diel_inner=din, refl_outer=rout, refl_inner=rin, sheen=sheen)
    return OUT


if __name__ == "__main__":
    print(generate())
