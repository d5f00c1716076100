"""TEST INFRASTRUCTURE — ctypes loader of the CPU oracle (oracle/_build/liboracle.so).

The oracle is the checker: the tests, the smoke check and the benchmark's
CPU-baseline leg call it. It is never used as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j8"], check=True)


def use_native_build(build_dir, jobs=16):
    """Builds the oracle with -march=native for the host this process runs
    on into build_dir and loads that copy (for the CPU-baseline timing; the
    tests use the portable build). Call before anything else loads the
    oracle; returns the seconds the build took."""
    global ORACLE_SO
    import time
    if _lib is not None:
        raise RuntimeError("the oracle is already loaded")
    t = time.perf_counter()
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j%d" % jobs, "MARCH=native", "BUILD=" + build_dir],
                   check=True)
    ORACLE_SO = os.path.join(build_dir, "liboracle.so")
    return time.perf_counter() - t


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = ctypes.CDLL(ORACLE_SO)
    return _lib


def p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def cdf_build(texels):
    h, w, ch = texels.shape
    t = np.ascontiguousarray(texels, np.float32)
    cx = np.zeros((h, w), np.float32)
    cy = np.zeros(h, np.float32)
    integ = np.zeros(1, np.float32)
    rc = lib().orc_cdf_build(p(t), w, h, ch, p(cx), p(cy), p(integ))
    assert rc == 0
    return cx, cy, float(integ[0])


def cdf_build_f64(texels):
    h, w, ch = texels.shape
    t = np.ascontiguousarray(texels, np.float32)
    cx = np.zeros((h, w), np.float64)
    cy = np.zeros(h, np.float64)
    integ = np.zeros(1, np.float64)
    rc = lib().orc_cdf_build_f64(p(t), w, h, ch, p(cx), p(cy), p(integ))
    assert rc == 0
    return cx, cy, float(integ[0])


def cdf_invert(cx, cy):
    h, w = cx.shape
    cond = np.zeros((h, w), np.uint16)
    marg = np.zeros(h, np.uint16)
    lib().orc_cdf_invert(p(np.ascontiguousarray(cx)), p(np.ascontiguousarray(cy)), w, h, p(cond), p(marg))
    return cond, marg


def cdf_block_size(w):
    return lib().orc_cdf_block_size(w)


def sobol_table():
    t = np.zeros(128, np.uint32)
    lib().orc_sobol_table(p(t))
    return t.reshape(4, 32)


def sobol(records):
    """records: (n, 5) uint32 {sample, hash, offset, dim, kind} -> (n, 3) float32"""
    r = np.ascontiguousarray(records, np.uint32)
    out = np.zeros((len(r), 3), np.float32)
    lib().orc_sobol(p(r), p(out), len(r))
    return out


def pixel_hash(xy):
    xy = np.ascontiguousarray(xy, np.uint32)
    out = np.zeros(len(xy), np.uint32)
    lib().orc_pixel_hash(p(xy), p(out), len(xy))
    return out


def wavelengths(u):
    u = np.ascontiguousarray(u, np.float32)
    lam = np.zeros((len(u), 4), np.float32)
    pdf = np.zeros((len(u), 4), np.float32)
    lib().orc_wavelengths(p(u), p(lam), p(pdf), len(u))
    return lam, pdf


# ------------------------------------------------------------ scenes
class OrcScene(ctypes.Structure):
    _fields_ = [
        ("verts", ctypes.c_void_p), ("nverts", ctypes.c_uint32),
        ("indices", ctypes.c_void_p), ("ntris", ctypes.c_uint32),
        ("attribs", ctypes.c_void_p),
        ("triMaterial", ctypes.c_void_p), ("triLight", ctypes.c_void_p),
        ("materials", ctypes.c_void_p), ("nmaterials", ctypes.c_uint32),
        ("textures", ctypes.c_void_p), ("ntextures", ctypes.c_uint32),
        ("texels", ctypes.c_void_p),
        ("spectra", ctypes.c_void_p), ("nspectra", ctypes.c_uint32),
        ("xyz", ctypes.c_void_p), ("basis", ctypes.c_void_p),
        ("lutConductor", ctypes.c_void_p), ("lutDielOuter", ctypes.c_void_p), ("lutDielInner", ctypes.c_void_p),
        ("lutReflOuter", ctypes.c_void_p), ("lutReflInner", ctypes.c_void_p), ("lutSheen", ctypes.c_void_p),
        ("lights", ctypes.c_void_p), ("nlights", ctypes.c_uint32),
        ("nodes", ctypes.c_void_p), ("nnodes", ctypes.c_uint32),
        ("lightPointers", ctypes.c_void_p), ("bitmasks", ctypes.c_void_p),
        ("envEnabled", ctypes.c_uint32), ("envWidth", ctypes.c_uint32), ("envHeight", ctypes.c_uint32),
        ("envChannels", ctypes.c_uint32),
        ("envTexels", ctypes.c_void_p), ("envCond", ctypes.c_void_p), ("envMarg", ctypes.c_void_p),
        ("envIntegral", ctypes.c_float),
        ("shapes", ctypes.c_void_p), ("nshapes", ctypes.c_uint32),
    ]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
        ("maxPathDepth", ctypes.c_uint32), ("maxReflectedPathDepth", ctypes.c_uint32),
        ("maxTransmittedPathDepth", ctypes.c_uint32),
        ("sampleOffset", ctypes.c_uint32), ("sampleCount", ctypes.c_uint32),
        ("camU", ctypes.c_float * 3), ("camV", ctypes.c_float * 3), ("camW", ctypes.c_float * 3),
        ("perspectiveScaleW", ctypes.c_float), ("perspectiveScaleH", ctypes.c_float),
        ("depthOfField", ctypes.c_uint32), ("aperture", ctypes.c_float), ("focusDistance", ctypes.c_float),
        ("flags", ctypes.c_uint32),
        ("cropX", ctypes.c_uint32), ("cropY", ctypes.c_uint32), ("cropW", ctypes.c_uint32), ("cropH", ctypes.c_uint32),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [("closestRays", ctypes.c_uint64), ("shadowRays", ctypes.c_uint64), ("pathVertices", ctypes.c_uint64),
                ("bvhSeconds", ctypes.c_double), ("renderSeconds", ctypes.c_double)]


def light_tree_build(lights, flux, shapes=None, shape_flux=None):
    """The oracle's tree over triangle lights and (optionally) shape lights."""
    n = len(lights)
    verts = np.ascontiguousarray(lights["vertices"], np.float32)
    flux = np.ascontiguousarray(flux, np.float32)
    from ratbite_amd._lib import PACKED_NODE_DTYPE, SHAPE_LIGHT_DTYPE
    m = 0 if shapes is None else len(shapes)
    tot = n + m
    nodes = np.zeros(max(1, 2 * tot), PACKED_NODE_DTYPE)
    lp = np.zeros(max(1, tot), np.uint32)
    bm = np.zeros(max(1, tot), np.uint64)
    if m:
        sh = np.ascontiguousarray(shapes, SHAPE_LIGHT_DTYPE)
        sf = np.ascontiguousarray(shape_flux, np.float32)
        rc = lib().orc_light_tree_build_shapes(p(verts) if n else None, p(flux) if n else None, n, p(sh), p(sf), m,
                                               p(nodes), p(lp), p(bm))
    else:
        rc = lib().orc_light_tree_build(p(verts), p(flux), n, p(nodes), p(lp), p(bm))
    assert rc >= 0, rc
    return nodes[:rc], lp[:tot], bm[:tot]


class OracleScene:
    """Holds the numpy buffers alive and exposes an orc_scene."""

    def __init__(self, sc, tree=None, env=None):
        t = sc.tables
        self.keep = []

        def keep(a, dt=None):
            a = np.ascontiguousarray(a if dt is None else np.asarray(a, dt))
            self.keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        s = OrcScene()
        s.verts, s.nverts = keep(sc.verts, np.float32), len(sc.verts)
        s.indices, s.ntris = keep(sc.indices, np.uint32), len(sc.indices)
        s.attribs = keep(sc.attribs, np.uint8)
        s.triMaterial = keep(sc.tri_material, np.uint32)
        s.triLight = keep(sc.tri_light, np.uint32)
        s.materials, s.nmaterials = keep(sc.materials), len(sc.materials)
        s.textures, s.ntextures = (keep(sc.textures), len(sc.textures)) if len(sc.textures) else (None, 0)
        s.texels = keep(sc.texels if sc.texels.size else np.zeros(16, np.uint8))
        s.spectra, s.nspectra = keep(sc.spectra, np.float32), len(sc.spectra)
        s.xyz = keep(t["xyz"], np.float32)
        s.basis = keep(t["basis"], np.float32)
        s.lutConductor = keep(t["conductor"], np.float32)
        s.lutDielOuter = keep(t["diel_outer"], np.float32)
        s.lutDielInner = keep(t["diel_inner"], np.float32)
        s.lutReflOuter = keep(t["refl_outer"], np.float32)
        s.lutReflInner = keep(t["refl_inner"], np.float32)
        s.lutSheen = keep(t["sheen"], np.float32)
        s.lights, s.nlights = (keep(sc.lights), len(sc.lights)) if len(sc.lights) else (None, 0)
        nshapes = getattr(sc, "nshapes", 0)
        if nshapes:
            s.shapes, s.nshapes = keep(sc.shapes), nshapes
        if tree is None and len(sc.lights) + nshapes:
            tree = light_tree_build(sc.lights, sc.flux, sc.shapes if nshapes else None,
                                    sc.shape_flux if nshapes else None)
        self.tree = tree
        if tree is not None:
            s.nodes, s.nnodes = keep(tree[0]), len(tree[0])
            s.lightPointers = keep(tree[1], np.uint32)
            s.bitmasks = keep(tree[2], np.uint64)
        if sc.env is not None:
            tex = np.ascontiguousarray(sc.env, np.float32)
            h, w, ch = tex.shape
            if env is None:
                cx, cy, integ = cdf_build(tex)
                cond, marg = cdf_invert(cx, cy)
                integral = np.float32(integ) / (np.float32(w) * np.float32(h))
                env = (cond, marg, float(integral))
            self.env = env
            s.envEnabled, s.envWidth, s.envHeight, s.envChannels = 1, w, h, ch
            s.envTexels = keep(tex)
            s.envCond = keep(env[0], np.uint16)
            s.envMarg = keep(env[1], np.uint16)
            s.envIntegral = env[2]
        self.s = s


def render(sc, params_lp, threads=0, crop=None, accum=None, oscene=None):
    """Renders with the oracle; params_lp is a ratbite_amd.LaunchParams."""
    os_ = oscene or OracleScene(sc)
    pr = OrcParams()
    for name, _ in params_lp._fields_:
        v = getattr(params_lp, name)
        if hasattr(v, "__len__"):
            getattr(pr, name)[:] = list(v)
        else:
            setattr(pr, name, v)
    if crop:
        pr.cropX, pr.cropY, pr.cropW, pr.cropH = crop
    if accum is None:
        accum = np.zeros((pr.height, pr.width, 4), np.float64)
    st = OrcStats()
    rc = lib().orc_render(ctypes.byref(os_.s), ctypes.byref(pr), p(accum), threads, ctypes.byref(st))
    assert rc == 0
    return accum, {"closestRays": st.closestRays, "shadowRays": st.shadowRays, "pathVertices": st.pathVertices,
                   "bvhSeconds": st.bvhSeconds, "renderSeconds": st.renderSeconds}


def resolve(accum, m3, exposure_scale):
    h, w = accum.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    m = np.ascontiguousarray(np.asarray(m3, np.float32).T.reshape(9))   # column-major
    lib().orc_resolve(p(np.ascontiguousarray(accum)), w, h, p(m), ctypes.c_float(exposure_scale), p(out))
    return out
