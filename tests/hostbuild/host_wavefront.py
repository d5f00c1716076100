"""TEST INFRASTRUCTURE. ctypes front end of the host build of the product's
wavefront integrator (wavefront_host.cpp): renders a scenes.Scene on the CPU
with the product's own per-path kernel bodies, optionally under
AddressSanitizer / UndefinedBehaviorSanitizer.

The light tree and the env-map tables come from the oracle (bit-identical to
the product's host builder and CDF kernel, which the CPU and GPU tests pin).
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

LIB = os.path.join(HERE, "_build", "libwavefront_host.so")
LIB_SAN = os.path.join(HERE, "_build", "libwavefront_host_san.so")


class HostScene(ctypes.Structure):
    _fields_ = [("verts", ctypes.c_void_p), ("nverts", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("ntris", ctypes.c_uint32),
                ("attribs", ctypes.c_void_p), ("attribBytes", ctypes.c_uint64),
                ("triMaterial", ctypes.c_void_p), ("triLight", ctypes.c_void_p),
                ("materials", ctypes.c_void_p), ("nmaterials", ctypes.c_uint32),
                ("textures", ctypes.c_void_p), ("ntextures", ctypes.c_uint32),
                ("texels", ctypes.c_void_p), ("texelBytes", ctypes.c_uint64),
                ("spectra", ctypes.c_void_p), ("nspectra", ctypes.c_uint32),
                ("colors", ctypes.c_void_p), ("luts", ctypes.c_void_p),
                ("lights", ctypes.c_void_p), ("nlights", ctypes.c_uint32),
                ("treeNodes", ctypes.c_void_p), ("ntreeNodes", ctypes.c_uint32),
                ("treePointers", ctypes.c_void_p), ("treeMasks", ctypes.c_void_p),
                ("envOn", ctypes.c_uint32), ("envW", ctypes.c_uint32), ("envH", ctypes.c_uint32),
                ("envC", ctypes.c_uint32), ("envTex", ctypes.c_void_p), ("envCond", ctypes.c_void_p),
                ("envMarg", ctypes.c_void_p), ("envIntegral", ctypes.c_float)]


class HostStats(ctypes.Structure):
    _fields_ = [("closestRays", ctypes.c_uint64), ("shadowRays", ctypes.c_uint64),
                ("pathVertices", ctypes.c_uint64), ("nonFiniteSamples", ctypes.c_uint64),
                ("hitNonFinite", ctypes.c_uint64), ("rayNonFinite", ctypes.c_uint64),
                ("shadowNonFinite", ctypes.c_uint64), ("queueOverflow", ctypes.c_uint64),
                ("bvhDepth", ctypes.c_uint32), ("nnodes4", ctypes.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-j2"], check=True)


def asan_runtime():
    """Path of the ASan runtime a process must preload to load the sanitized build."""
    r = subprocess.run(["make", "-s", "-C", HERE, "asan-runtime"], check=True, capture_output=True, text=True)
    return r.stdout.strip()


_libs = {}


def lib(sanitized=False):
    path = LIB_SAN if sanitized else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.host_render.restype = ctypes.c_int
        L.host_render.argtypes = [ctypes.POINTER(HostScene), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(HostStats)]
        _libs[path] = L
    return _libs[path]


class PreparedScene:
    """A scene's arrays, light tree and env-map tables, kept alive for
    host_render. tree = (packed nodes, pointers, bitmasks) and env = (cond,
    marg, integral / (W H)) override the oracle's FP32 builds."""

    def __init__(self, sc, tree=None, env=None):
        import oracle_lib as O
        t = sc.tables
        self.keep = []

        def keep(a, dt=None):
            a = np.ascontiguousarray(a if dt is None else np.asarray(a, dt))
            self.keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        s = HostScene()
        s.verts, s.nverts = keep(sc.verts, np.float32), len(sc.verts)
        s.indices, s.ntris = keep(sc.indices, np.uint32), len(sc.indices)
        attribs = np.ascontiguousarray(sc.attribs, np.uint8)
        s.attribs, s.attribBytes = keep(attribs), attribs.size
        s.triMaterial = keep(sc.tri_material, np.uint32)
        s.triLight = keep(sc.tri_light, np.uint32)
        s.materials, s.nmaterials = keep(sc.materials), len(sc.materials)
        if len(sc.textures):
            s.textures, s.ntextures = keep(sc.textures), len(sc.textures)
        texels = np.ascontiguousarray(sc.texels, np.uint8)
        s.texels, s.texelBytes = keep(texels), texels.size
        spectra = sc.spectra if len(sc.spectra) else np.ones((1, 471), np.float32)
        s.spectra, s.nspectra = keep(spectra, np.float32), len(spectra)
        colors = np.concatenate([np.asarray(t["xyz"], np.float32).reshape(3, 471),
                                 np.asarray(t["basis"], np.float32).reshape(3, 471)])
        s.colors = keep(colors)
        luts = np.concatenate([np.asarray(t[k], np.float32).ravel() for k in
                               ("conductor", "diel_outer", "diel_inner", "refl_outer", "refl_inner", "sheen")])
        s.luts = keep(luts)
        if len(sc.lights):
            if tree is None:
                tree = O.light_tree_build(sc.lights, sc.flux)
            s.lights, s.nlights = keep(sc.lights), len(sc.lights)
            s.treeNodes, s.ntreeNodes = keep(tree[0]), len(tree[0])
            s.treePointers = keep(tree[1], np.uint32)
            s.treeMasks = keep(tree[2], np.uint64)
        if sc.env is not None:
            tex = np.ascontiguousarray(sc.env, np.float32)
            h, w, ch = tex.shape
            if env is None:
                cx, cy, integ = O.cdf_build(tex)
                cond, marg = O.cdf_invert(cx, cy)
                env = (cond, marg, float(np.float32(integ) / (np.float32(w) * np.float32(h))))
            cond, marg, integral = env
            s.envOn, s.envW, s.envH, s.envC = 1, w, h, ch
            s.envTex = keep(tex)
            s.envCond = keep(cond, np.uint16)
            s.envMarg = keep(marg, np.uint16)
            s.envIntegral = integral
        self.s = s


def render(sc, lp, pass_samples=0, threads=0, sanitized=False, accum=None, prepared=None, sort_hits=True):
    """Renders launch parameters lp (a ratbite_amd LaunchParams) on the CPU.
    Returns (accum H x W x 4 float64, stats dict)."""
    ps = prepared or PreparedScene(sc)
    if accum is None:
        accum = np.zeros((lp.height, lp.width, 4), np.float64)
    accum = np.ascontiguousarray(accum, np.float64)
    st = HostStats()
    rc = lib(sanitized).host_render(ctypes.byref(ps.s), ctypes.addressof(lp), pass_samples, threads, int(sort_hits),
                                    accum.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("host_render failed with %d" % rc)
    return accum, {k: getattr(st, k) for k, _ in HostStats._fields_}
