"""Writes tests/golden/sw_rast_cases.json: texel lists of the reference's
software rasteriser (src/core/sw_rast.h, built into oracle/_ref by
oracle/ref/Makefile) for seeded triangles of every kind the rasteriser
distinguishes (general, flat bottom, flat top, corners on texel edges and
centres, tiny, collinear). Run from the repository root after building
oracle/_ref."""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from test_emissive_flux import REF_SO, random_triangles  # noqa: E402


def main():
    lib = ctypes.CDLL(REF_SO)
    lib.ref_rasterise.restype = ctypes.c_int
    lib.ref_rasterise.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int]
    cases = []
    for t in random_triangles(20261203, 96):
        t = (t * np.float32(0.25)).astype(np.float32)      # small lists
        xy = np.zeros((1 << 14, 2), np.int32)
        pts = [np.ascontiguousarray(p) for p in t]
        n = lib.ref_rasterise(*[p.ctypes.data_as(ctypes.c_void_p) for p in pts],
                              xy.ctypes.data_as(ctypes.c_void_p), 1 << 14)
        cases.append({"tri": [float(v) for v in t.reshape(-1)], "texels": xy[:n].tolist()})
    with open(os.path.join(ROOT, "tests", "golden", "sw_rast_cases.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print("wrote %d cases" % len(cases))


if __name__ == "__main__":
    main()
