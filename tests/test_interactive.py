"""The progressive / interactive loop (ratbite_amd/interactive.py): the
reference's sub-launch sizing (rendering_interface.cpp:1859-1872) and
command processing (1534-1845), on the host with a recording stand-in for
the context; the GPU tests (test_interactive_gpu.py) run it for real."""
import numpy as np
import pytest

from ratbite_amd import interactive as I
from ratbite_amd import scenes


def test_gradual_schedule():
    sch = I.sublaunch_schedule(I.Mode.GRADUAL, 200)
    # 1, then floor(max(3, c)^1.5) capped at 64: 3^1.5 = 5.196, 5^1.5 = 11.18, 11^1.5 = 36.48, 36^1.5 = 216 -> 64
    assert [c for _, c in sch] == [1, 5, 11, 36, 64, 64, 19]
    assert [o for o, _ in sch] == [0, 1, 6, 17, 53, 117, 181]
    assert sum(c for _, c in sch) == 200


def test_immediate_schedule_and_edges():
    assert I.sublaunch_schedule(I.Mode.IMMEDIATE, 4) == [(0, 1), (1, 1), (2, 1), (3, 1)]
    assert I.sublaunch_schedule(I.Mode.GRADUAL, 0) == []
    assert I.sublaunch_schedule(I.Mode.GRADUAL, 1) == [(0, 1)]
    assert I.sublaunch_schedule(I.Mode.GRADUAL, 3) == [(0, 1), (1, 2)]
    # the formula as written: a sub-launch of 2 grows like one of 3
    assert I.next_sublaunch(I.Mode.GRADUAL, 2, 1000, 10) == 5


class FakeCtx:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def rec(*a):
            self.calls.append((name,) + tuple(a))
        return rec


class FakeRenderer:
    def __init__(self, scene):
        self.scene = scene
        self.ctx = FakeCtx()


@pytest.fixture
def session():
    sc = scenes.cornell(width=16, height=16, spp=8)
    return I.RenderSession(FakeRenderer(sc), sample_count=8)


def _names(s):
    return [c[0] for c in s.r.ctx.calls]


def test_step_runs_the_schedule(session):
    while not session.step():
        pass
    renders = [c[1] for c in session.r.ctx.calls if c[0] == "render"]
    assert [(p.sampleOffset, p.sampleCount) for p in renders] == [(0, 1), (1, 5), (6, 2)]
    assert session.processed == 8 and session.finished


def test_exposure_only_re_resolves(session):
    session.step()
    session.push(I.Cmd.CHANGE_IMAGE_EXPOSURE, 1.5)
    assert session.process_commands() is False
    assert session.offset == 1 and "resolve" in _names(session)


def test_camera_move_translates_and_restarts(session):
    session.step()
    session.step()
    pos = np.asarray(session.scene.camera["pos"], np.float64)
    v0 = session.scene.verts.copy()
    session.push(I.Cmd.CHANGE_CAMERA_POS, pos + [0.5, 0.0, -1.0])
    assert session.process_commands() is True
    tr = [c for c in session.r.ctx.calls if c[0] == "translate_scene"]
    assert len(tr) == 1 and np.allclose(tr[0][1], [0.5, 0.0, -1.0])
    assert np.allclose(session.scene.verts, v0 - np.float32([0.5, 0.0, -1.0]))
    assert session.offset == 0 and session.count == 1


def test_model_transform_refits_and_rebuilds_moved_emitters(session):
    sc = session.scene
    lt = sc.indices[sc.lights["primitiveDataIndex"]]
    first = int(lt.min())
    moved = sc.verts[first:first + 4] + np.float32([0.0, -0.1, 0.0])
    session.push(I.Cmd.CHANGE_MODEL_TRANSFORM, (first, moved))
    session.process_commands()
    names = _names(session)
    assert names.index("update_vertices") < names.index("build_light_tree") < names.index("refit_accel")
    assert np.allclose(sc.lights["vertices"].reshape(-1, 3, 3), sc.verts[lt])
    # a move of non-emitting vertices refits without touching the light tree
    session.r.ctx.calls.clear()
    session.push(I.Cmd.CHANGE_MODEL_TRANSFORM, (0, sc.verts[0:4]))
    session.process_commands()
    assert "build_light_tree" not in _names(session) and "refit_accel" in _names(session)


def test_commands_pop_last_in_first_out(session):
    session.push(I.Cmd.CHANGE_SAMPLE_COUNT, 32)
    session.push(I.Cmd.CHANGE_SAMPLE_COUNT, 16)
    session.process_commands()
    assert session.sample_count == 32          # the first pushed is applied last
    session.push(I.Cmd.CHANGE_PATH_DEPTH, (0, 5, 5))
    session.process_commands()
    assert session.depth == (1, 1, 1)
