"""Batched multi-scene launches and the launch-shape autotuner: a scene
rendered inside a batch accumulates exactly what it accumulates alone, and
no launch shape changes any result."""
import numpy as np
import pytest

from ratbite_amd import scenes
from ratbite_amd.renderer import Renderer, render_batch

pytestmark = pytest.mark.gpu


def _scenes():
    return [
        scenes.cornell(width=32, height=32, spp=8),
        scenes.spheres(n_spheres=10, level=2, n_lights=50, width=40, height=24, spp=5, env_size=(128, 64)),
        scenes.cornell(seed=0x5EED0007, width=24, height=20, spp=11, light_up=True),
        scenes.spheres(seed=0x5EED0009, n_spheres=4, level=1, n_lights=3, width=8, height=8, spp=3, env_size=None),
    ]


def test_batch_equals_individual_renders():
    rs = [Renderer(sc) for sc in _scenes()]
    try:
        alone = []
        for r in rs:
            r.render(spp=r.scene.spp, offset=0)
            alone.append(r.accum())
        done = render_batch(rs)
        print("completion ms", np.round(done, 3), flush=True)
        assert np.all(done > 0)
        for r, a in zip(rs, alone):
            b = r.accum()
            assert np.array_equal(a, b), r.scene.name
    finally:
        for r in rs:
            r.close()


def test_batch_with_small_pass_budget():
    # a pass budget below one scene's pixel count still gives the same sums
    rs = [Renderer(sc) for sc in _scenes()[:3]]
    try:
        alone = []
        for r in rs:
            r.render(spp=r.scene.spp, offset=0)
            alone.append(r.accum())
        rs[0].ctx.set_path_slots(700)
        render_batch(rs)
        rs[0].ctx.set_path_slots(0)
        for r, a in zip(rs, alone):
            assert np.array_equal(a, r.accum()), r.scene.name
    finally:
        for r in rs:
            r.close()


def test_batch_of_64_full_size_scenes():
    """The batched uneven workload at its full size (64 scenes, 1k..500k
    triangles, films up to 1024^2, 32 spp, depth 12) in one render_batch:
    every pixel of every scene receives exactly 32 finite samples, the
    completion times are ordered like the passes that finished each scene,
    and a subset of the scenes accumulates bit-identically to single renders
    in fresh contexts."""
    scs = scenes.batch(64)
    rs = [Renderer(sc) for sc in scs]
    try:
        done = np.asarray(render_batch(rs))
        print("64 scenes: %d triangles, %d pixels; completion ms p50 %.1f p99 %.1f" % (
            sum(len(sc.indices) for sc in scs), sum(sc.width * sc.height for sc in scs),
            np.percentile(done, 50), np.percentile(done, 99)), flush=True)
        assert np.all(done > 0) and np.all(np.isfinite(done))
        accs = [r.accum() for r in rs]
        for sc, a in zip(scs, accs):
            assert a.shape == (sc.height, sc.width, 4)
            assert np.all(a[..., 3] == 32), sc.name
            assert np.all(np.isfinite(a[..., :3])), sc.name
            assert np.all(a[..., 1] >= 0.0)
        # larger films finish no earlier than smaller ones of no fewer samples
        px = np.array([sc.width * sc.height for sc in scs])
        assert done[px.argmax()] >= done[px.argmin()]
        subset = [int(px.argmin()), int(px.argmax()), 7, 40]
    finally:
        for r in rs:
            r.close()
    for k in subset:
        r = Renderer(scs[k])
        try:
            r.render(spp=32, offset=0)
            assert np.array_equal(r.accum(), accs[k]), k
        finally:
            r.close()


def test_autotune_keeps_results(ctx):
    sc = scenes.cornell(width=48, height=48, spp=8)
    r = Renderer(sc, ctx=ctx)
    r.render(spp=8, offset=0)
    a = r.accum()
    t = r.autotune(trial_spp=2)
    print("tuning", t, flush=True)
    # pathSlots 0 is the automatic budget (a whole launch per pass when it fits)
    assert t["samplesPerSecond"] > 0 and t["pathSlots"] >= 0
    r.render(spp=8, offset=0)
    assert np.array_equal(a, r.accum())
    for w in (1, 3, 4):
        r.set_tuning(shade_waves=w)
        r.render(spp=8, offset=0)
        assert np.array_equal(a, r.accum()), w
    for w in (1, 7, 8):
        r.set_tuning(trace_waves=w)
        r.render(spp=8, offset=0)
        assert np.array_equal(a, r.accum()), w
        # quantised nodes find the same closest t; only a ray through a shared
        # edge (both triangles at the same t) may report the other triangle
        r.set_tuning(trace_waves=w, quant_nodes=True)
        r.render(spp=8, offset=0)
        same = (a == r.accum()).all(axis=2)
        assert same.mean() > 0.998, (w, same.mean())
    # per-lane work fetch instead of wave-synchronous ray batches
    r.set_tuning(trace_fetch=1)
    r.render(spp=8, offset=0)
    assert np.array_equal(a, r.accum())
    r.set_tuning()


def test_material_sorted_shading_keeps_results(ctx):
    # conductor hits queued from the end of the hit queue: the same sums for
    # plain, guided (the network's mixtures are indexed by queue position)
    # and batched renders
    sc = scenes.spheres(n_spheres=24, level=2, n_lights=80, width=48, height=32, spp=6, env_size=(128, 64))
    assert np.any(sc.materials["bxdf"] == scenes.BXDF_CONDUCTOR)
    r = Renderer(sc, ctx=ctx)
    try:
        r.configure_guide(seed=3)
        out = {}
        for sort in (False, True):
            r.set_tuning(sort_hits=sort)
            assert r.tuning()["sortHits"] == int(sort)
            r.render(spp=6, offset=0)
            plain = r.accum()
            r.render(spp=6, offset=0, guided=True)
            out[sort] = (plain, r.accum())
        assert np.array_equal(out[False][0], out[True][0])
        assert np.array_equal(out[False][1], out[True][1])
    finally:
        r.set_tuning()
    rs = [Renderer(s) for s in _scenes()]
    try:
        alone = []
        for q in rs:
            q.render(spp=q.scene.spp, offset=0)
            alone.append(q.accum())
        rs[0].set_tuning(sort_hits=True)
        render_batch(rs)
        for q, a in zip(rs, alone):
            assert np.array_equal(a, q.accum()), q.scene.name
    finally:
        for q in rs:
            q.close()


def test_graph_replays_match_direct_launches(ctx):
    # several passes per render replay one graph executable (updated in place);
    # eager renders and host synchronisations run between the graph renders
    sc = scenes.spheres(n_spheres=8, level=2, n_lights=60, width=40, height=24, spp=12, env_size=(128, 64))
    r = Renderer(sc, ctx=ctx)
    ctx.set_path_slots(3 * 40 * 24)            # 3 samples per pass -> 4 passes
    try:
        r.set_tuning(path_slots=3 * 40 * 24, use_graph=False)
        r.render(spp=12, offset=0)
        eager = r.accum()
        for rep in range(3):
            r.set_tuning(path_slots=3 * 40 * 24, use_graph=True)
            r.render(spp=12, offset=0)
            g = r.accum()
            assert np.array_equal(g, eager), rep
            r.set_tuning(path_slots=3 * 40 * 24, use_graph=False)
            r.render(spp=5, offset=0)           # eager work between replays
            ctx.synchronize()
    finally:
        r.set_tuning()
