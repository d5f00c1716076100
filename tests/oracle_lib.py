"""Loader of the CPU oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it. It is
never used as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j8"], check=True)


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
