"""The interactive loop on the GPU: a GRADUAL progressive render accumulates
exactly what one render of the same samples does; after a model moves
(vertex update + LBVH refit), the camera moves (translation of the scene
into the new camera-relative frame + refit) or a material changes, the
render matches the oracle on the updated scene, and a refitted structure
traces like a rebuilt one."""
import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import interactive as I
from ratbite_amd import scenes
from ratbite_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _close(a, b):
    g, o = a[..., :3] / a[..., 3:4], b[..., :3] / b[..., 3:4]
    return (np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)).all(axis=2).mean()


def _oracle(r, sc, spp):
    acc, _ = O.render(sc, sc.launch_params(offset=0, count=spp), oscene=O.OracleScene(sc, tree=r.light_tree()))
    return acc


def test_gradual_session_equals_one_render(ctx):
    sc = scenes.cornell(width=48, height=40, spp=100)
    r = Renderer(sc, ctx=ctx)
    s = I.RenderSession(r, sample_count=100, mode=I.Mode.GRADUAL)
    n = 0
    while not s.step():
        n += 1
    prog = r.accum()
    r.render(spp=100, offset=0)
    one = r.accum()
    print("GRADUAL: %d sub-launches" % (n + 1), flush=True)
    assert n + 1 == len(I.sublaunch_schedule(I.Mode.GRADUAL, 100))
    assert np.array_equal(prog, one)


def _sphere_vertex_range(sc):
    # the last UV sphere of the Cornell box (8 spheres of 33 x 17 vertices at the end)
    nv = 33 * 17
    return len(sc.verts) - nv, nv


def test_refit_after_a_model_moves(ctx):
    sc = scenes.cornell(width=64, height=64, spp=8)
    r = Renderer(sc, ctx=ctx)
    s = I.RenderSession(r, sample_count=8)
    s.run()
    first, nv = _sphere_vertex_range(sc)
    moved = sc.verts[first:first + nv] + np.float32([0.25, 0.1, -0.2])
    s.push(I.Cmd.CHANGE_MODEL_TRANSFORM, (first, moved))
    s.run()
    refit = r.accum()
    assert np.array_equal(sc.verts[first:first + nv], moved)
    frac = _close(refit, _oracle(r, sc, 8))
    # a fresh context builds its own hierarchy over the moved geometry
    r2 = Renderer(sc)
    try:
        r2.render(spp=8, offset=0)
        rebuilt = r2.accum()
    finally:
        r2.close()
    frac2 = _close(refit, rebuilt)
    print("after the move: %.4f of pixels match the oracle, %.4f the rebuilt structure" % (frac, frac2), flush=True)
    assert frac >= 0.99 and frac2 >= 0.999


def test_camera_move_translates_the_scene(ctx):
    sc = scenes.spheres(n_spheres=10, level=2, n_lights=200, width=64, height=48, spp=4, env_size=(128, 64))
    r = Renderer(sc, ctx=ctx)
    s = I.RenderSession(r, sample_count=4)
    s.run()
    s.push(I.Cmd.CHANGE_CAMERA_POS, np.asarray(sc.camera["pos"]) + np.array([3.0, -1.0, 5.0]))
    s.run()
    frac = _close(r.accum(), _oracle(r, sc, 4))
    print("after the camera move: %.4f of pixels match the oracle" % frac, flush=True)
    assert frac >= 0.99


def test_material_change(ctx):
    sc = scenes.cornell(width=48, height=48, spp=8)
    r = Renderer(sc, ctx=ctx)
    s = I.RenderSession(r, sample_count=8)
    s.run()
    before = r.accum()
    m = sc.materials[0:1].copy()
    m["baseColor"][0, :3] = (0.9, 0.1, 0.1)
    s.push(I.Cmd.CHANGE_MODEL_MATERIAL, (0, m))
    s.run()
    after = r.accum()
    assert not np.array_equal(before, after)
    r2 = Renderer(sc)
    try:
        r2.render(spp=8, offset=0)
        assert np.array_equal(after, r2.accum())
    finally:
        r2.close()
