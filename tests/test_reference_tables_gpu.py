"""Rendering with the reference's own BxDF LUTs as inputs on both sides.

This project generates its tables from published definitions, and they
differ from the reference's LUT binaries (most visibly in the reflective
inner table, which includes total internal reflection here). Callers who
want the reference's exact inputs upload its tables through rbh_upload_luts.
This test does that with the reference's six LUT data files (the fixture
tests/golden/reference_luts.npz, written from bin/*LUT.bin by
tests/golden/make_reference_luts.py): the GPU tracer and the oracle, given
the same tables, agree at the bar of the other image parity tests.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import scenes
from ratbite_amd.renderer import Renderer

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_luts.npz")


def test_features_scene_with_reference_luts(ctx):
    ref = np.load(FIXTURE)
    sc = scenes.features(width=96, height=96, spp=8)
    tables = dict(sc.tables)
    for k in ref.files:
        tables[k] = np.ascontiguousarray(ref[k], np.float32)
    sc.tables = tables
    r = Renderer(sc, ctx=ctx)
    r.render(spp=8, offset=0)
    acc = r.accum()
    oacc, _ = O.render(sc, sc.launch_params(offset=0, count=8), oscene=O.OracleScene(sc, tree=r.light_tree()))
    g, o = acc[..., :3] / 8, oacc[..., :3] / 8
    # the reference's inner tables hold NaN texels in their last slice; a
    # pixel that reads one is non-finite on both sides or on neither
    fin = np.isfinite(g).all(axis=2)
    assert np.array_equal(fin, np.isfinite(o).all(axis=2))
    gf, of = g[fin], o[fin]
    close = np.abs(gf - of) <= 1e-4 * np.maximum(np.abs(of), 1e-3)
    frac = close.all(axis=1).mean()
    print("features with the reference's LUTs: %d finite pixels, within 1e-4: %.4f" % (fin.sum(), frac), flush=True)
    assert fin.mean() > 0.9
    assert frac >= 0.99
    # the tables reach the image: the same scene with this project's tables differs
    own = scenes.features(width=96, height=96, spp=8)
    r2 = Renderer(own, ctx=ctx)
    r2.render(spp=8, offset=0)
    g2 = r2.accum()[..., :3] / 8
    assert np.abs(g2[fin] - gf).max() > 1e-3
