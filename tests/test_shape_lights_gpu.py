"""Disk and sphere lights on the GPU: image parity with the oracle (the shape
tests ahead of the BVH walk in the trace and shadow kernels, the emitter
hit with its MIS weight, NEE sampling of shapes through the light tree),
the light-tree rebuild a new set of shapes requires, and the interactive
light commands (position, size, orientation, power, add, remove) with a
camera move."""
import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import _lib as L
from ratbite_amd import interactive as I
from ratbite_amd import scenes
from ratbite_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _frac(a, b):
    g, o = a[..., :3] / a[..., 3:4], b[..., :3] / b[..., 3:4]
    close = np.abs(g - o) <= 1e-4 * np.maximum(np.abs(o), 1e-3)
    return close.all(axis=2).mean(), float(np.mean((g - o) ** 2 / (o ** 2 + 1e-2)))


def _oracle(r, sc, spp, flags=0):
    acc, _ = O.render(sc, sc.launch_params(offset=0, count=spp, flags=flags),
                      oscene=O.OracleScene(sc, tree=r.light_tree()))
    return acc


@pytest.mark.parametrize("mode", ["mixed", "only"])
def test_cornell_with_shape_lights_matches_oracle(ctx, mode):
    sc = scenes.cornell(width=96, height=96, spp=16, shapes=mode)
    print("cornell shapes=%s" % mode, flush=True)
    r = Renderer(sc, ctx=ctx)
    r.render(spp=16, offset=0)
    frac, e = _frac(r.accum(), _oracle(r, sc, 16))
    print("pixels within 1e-4: %.4f  relMSE %.3g" % (frac, e), flush=True)
    assert frac >= 0.99 and e < 1e-5


@pytest.mark.parametrize("kind", ["disk", "sphere"])
def test_shape_lit_matches_oracle(ctx, kind):
    sc = scenes.shape_lit(kind, width=64, height=64, spp=32)
    print("shape_lit %s" % kind, flush=True)
    r = Renderer(sc, ctx=ctx)
    r.render(spp=32, offset=0, flags=L.RBH_RENDER_NO_DROP)
    frac, e = _frac(r.accum(), _oracle(r, sc, 32, flags=L.RBH_RENDER_NO_DROP))
    print("pixels within 1e-4: %.4f  relMSE %.3g" % (frac, e), flush=True)
    assert frac >= 0.99 and e < 1e-5


def test_new_shapes_need_a_tree_rebuild(ctx):
    sc = scenes.cornell(width=32, height=32, spp=2, shapes="mixed")
    r = Renderer(sc, ctx=ctx)
    ctx.upload_shape_lights(sc.shapes[:1], sc.shape_flux[:1])
    with pytest.raises(L.RbhError, match="rebuild the light tree"):
        r.render(spp=2, offset=0)
    ctx.build_light_tree(sc.lights, sc.flux)
    r.render(spp=2, offset=0)
    assert len(r.ctx.get_light_tree(len(sc.lights) + 1)[1]) == len(sc.lights) + 1


def test_light_commands_and_camera_move(ctx):
    sc = scenes.cornell(width=48, height=48, spp=4, shapes="mixed")
    r = Renderer(sc, ctx=ctx)
    s = I.RenderSession(r, sample_count=4)
    s.run()
    s.push(I.Cmd.CHANGE_LIGHT_POSITION, (1, sc.shapes[1]["position"] + np.float32([0.2, -0.3, 0.1])))
    s.push(I.Cmd.CHANGE_LIGHT_SIZE, (0, 0.3))
    s.push(I.Cmd.CHANGE_LIGHT_ORIENTATION, (0, (0.2, -1.0, 0.1)))
    s.push(I.Cmd.CHANGE_LIGHT_POWER, (1, 3.5))
    s.run()
    frac, e = _frac(r.accum(), _oracle(r, sc, 4))
    print("after the light commands: %.4f within 1e-4" % frac, flush=True)
    assert frac >= 0.99
    s.push(I.Cmd.ADD_LIGHT, scenes.sphere_light(sc.shapes[1]["position"] + np.float32([0.5, 0.0, 0.0]), 0.08, 1.0,
                                                int(sc.shapes[1]["materialIndex"])))
    s.push(I.Cmd.REMOVE_LIGHT, 0)
    s.push(I.Cmd.CHANGE_CAMERA_POS, np.asarray(sc.camera["pos"]) + np.array([0.3, -0.2, -0.5]))
    s.run()
    assert sc.nshapes == 2
    frac, e = _frac(r.accum(), _oracle(r, sc, 4))
    print("after add/remove and a camera move: %.4f within 1e-4" % frac, flush=True)
    assert frac >= 0.99
