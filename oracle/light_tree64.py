"""TEST INFRASTRUCTURE — FP64 evaluation of the light-tree pdf.

The same formulas as the FP32 oracle (rb_math.hpp light_tree_pdf, following
src/kernels/pathtrace.cu:974-1049 and src/core/math_util.h:246-587), evaluated
in double precision with the exact erf and expm1 instead of the FP32 fits, so
that the FP32 paths (oracle and GPU) can be given an error against it. The
packed-node decode keeps the reference's quirk (the average direction is
decoded from [0,1]-scaled components without the *2-1 remap,
light_tree_types.h:165-189).
"""
import math

import numpy as np

ENV_IMPORTANCE = 0.3
_A, _B, _C, _D, _E = (2.7360831611272558028247203765204, 17.02129778174187535455530451145,
                      4.0100826728510421403939290030394, 15.219156263147210594866010069381,
                      76.087896272360737270901154261082)
_erf = np.vectorize(math.erf, otypes=[np.float64])


def _oct_decode(ex, ey):
    x, y = ex.astype(np.float64), ey.astype(np.float64)
    z = 1.0 - np.abs(x) - np.abs(y)
    t = np.maximum(0.0, -z)
    x = x + np.where(x >= 0.0, -t, t)
    y = y + np.where(y >= 0.0, -t, t)
    v = np.stack([x, y, z], -1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _unpack(nodes, idx):
    n = nodes[idx]
    mean = np.stack([n["spatialMean"][:, k] for k in range(3)], -1).astype(np.float64)
    pad = n["packedAverageDirection"].astype(np.uint64)
    d = _oct_decode((pad & 0xFFFF).astype(np.float64) / 65535.0, (pad >> 16).astype(np.float64) / 65535.0)
    core = n["coreData"].astype(np.uint64)
    leaf = (core >> 31) != 0
    return {"mean": mean, "var": n["spatialVariance"].astype(np.float64), "flux": n["flux"].astype(np.float64),
            "dir": d, "sharp": n["sharpness"].astype(np.float64), "leaf": leaf,
            "right": np.where(leaf, 0, core).astype(np.int64),
            "count": np.where(leaf, (core >> 27) & 0xF, 0).astype(np.int64)}


def _expm1_over_x(x):
    return np.where(x == 0.0, 1.0, np.expm1(x) / np.where(x == 0.0, 1.0, x))


def _cc_upper(s):
    return np.where(s <= 1e-6, 0.5 - s / 6.0, (np.expm1(-s) + s) / np.maximum(s * s, 1e-300))


def _cc_lower(s):
    e = np.exp(-s)
    return np.where(s <= 1e-6, e * (0.5 - s / 3.0), e * (-np.expm1(-s) - s * e) / np.maximum(s * s, 1e-300))


def _cc_integral(cosine, s):
    # Tokuyoshi et al. 2024 supplement, Listing 7 (math_util.h:514-538)
    t = s * np.sqrt(0.5 * ((s + _A) * s + _B) / (((s + _C) * s + _D) * s + _E))
    tz = t * cosine
    lerp = 0.5 * (cosine * (1.0 - _erf(-tz)) + (1.0 - _erf(t))) - \
        0.5 / math.sqrt(math.pi) * np.exp(-tz * tz) * np.expm1(t * t * (cosine * cosine - 1.0)) / t
    lerp = np.clip(np.maximum(lerp, 0.5 * np.finfo(np.float32).eps), 0.0, 1.0)
    lo, up = _cc_lower(s), _cc_upper(s)
    return 2.0 * (lo + lerp * (up - lo))


def _importance(u, pos, nrm):
    ldir = u["mean"] - pos
    d2 = np.sum(ldir * ldir, -1)
    ldir = ldir / np.sqrt(d2)[:, None]
    var = np.maximum(u["var"], 2.0 ** -31 * d2)
    lsharp = d2 / var
    s1, s2 = u["sharp"], lsharp
    axis = u["dir"] * s1[:, None] + ldir * s2[:, None]
    sharp = np.linalg.norm(axis, axis=-1)
    cosine = np.clip(np.sum(u["dir"] * ldir, -1), -1.0, 1.0)
    smin = np.minimum(s1, s2)
    ratio = smin / np.maximum(np.maximum(s1, s2), np.finfo(np.float32).tiny)
    log_a = 2.0 * smin * (cosine - 1.0) / (np.sqrt(2.0 * ratio * cosine + ratio * ratio + 1.0) + ratio + 1.0)
    axis = axis / np.maximum(sharp, np.finfo(np.float32).tiny)[:, None]
    c2 = np.clip(np.sum(axis * nrm, -1), -1.0, 1.0)
    integral = 4.0 * math.pi * _expm1_over_x(-2.0 * s1)
    return u["flux"] / (var * integral) * np.exp(log_a) * _cc_integral(c2, sharp)


def pdf(nodes, masks, lights, pn, env=False):
    """FP64 pdf of picking triangle light `lights[i]` at pn[i] = (pos, normal)."""
    pn = np.asarray(pn, np.float64)
    pos, nrm = pn[:, 0:3], pn[:, 3:6]
    mask = masks[np.asarray(lights)].astype(np.uint64)
    m = len(pn)
    p = np.full(m, 1.0 - (ENV_IMPORTANCE if env else 0.0))
    cur = np.zeros(m, np.int64)
    u = _unpack(nodes, cur)
    active = ~u["leaf"]
    for depth in range(64):
        if not active.any():
            break
        a = np.nonzero(active)[0]
        ua = _unpack(nodes, cur[a])
        left, right = cur[a] + 1, ua["right"]
        ul, ur = _unpack(nodes, left), _unpack(nodes, right)
        li = _importance(ul, pos[a], nrm[a])
        ri = _importance(ur, pos[a], nrm[a])
        go_right = (mask[a] >> np.uint64(depth)) & np.uint64(1) != 0
        p[a] *= np.where(go_right, ri, li) / (li + ri)
        cur[a] = np.where(go_right, right, left)
        leaf_now = np.where(go_right, ur["leaf"], ul["leaf"])
        active[a] = ~leaf_now
    leaf = _unpack(nodes, cur)
    return p / leaf["count"]
