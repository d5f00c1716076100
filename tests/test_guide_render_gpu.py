"""Guided rendering: one-sample MIS between the network's vMF mixture and a
closed-form BSDF sampler. It must stay unbiased for any network (checked
against the plain tracer with the same no-drop sample semantics), gather
usable training records, and, once trained, lower the error at equal samples
on a scene lit only indirectly."""
import numpy as np
import pytest

from ratbite_amd import scenes
from ratbite_amd.renderer import Renderer

pytestmark = pytest.mark.gpu


def _img(r, spp, offset=0, **kw):
    r.render(spp=spp, offset=offset, fresh=True, **kw)
    a = r.accum()
    return a[..., :3] / a[..., 3:4]


def relmse(a, ref, eps=1e-2, trim=0.0):
    """Mean over pixels of the RGB-mean relative squared error; trim drops
    that fraction of the largest per-pixel errors (at least one pixel when
    positive), the outlier-robust variant path-guiding evaluations report."""
    e = ((a - ref) ** 2 / (ref ** 2 + eps)).mean(axis=2).ravel()
    if trim > 0.0:
        k = max(1, int(np.ceil(trim * e.size)))
        e = np.sort(e)[:-k]
    return float(e.mean())


def test_guided_is_unbiased_with_untrained_network(ctx):
    # the reference's NEE and sampling branches weight the diffuse lobe
    # differently; with a small light facing the scene the BSDF strategy's MIS
    # weight for direct light is tiny, so both estimators share one limit
    sc = scenes.cornell(width=64, height=64, spp=16)
    r = Renderer(sc, ctx=ctx)
    r.configure_guide(seed=3)
    plain = _img(r, 2048, no_drop=True)
    guided = _img(r, 2048, guided=True)
    assert np.all(np.isfinite(guided))
    # whole-image means and 8x8 block means agree within Monte Carlo noise
    mp, mg = plain.mean(axis=(0, 1)), guided.mean(axis=(0, 1))
    print("image mean plain", mp, "guided", mg, flush=True)
    assert np.all(np.abs(mg - mp) / mp < 0.02)
    bp = plain.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))
    bg = guided.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))
    rel = np.abs(bg - bp) / np.maximum(bp, 1e-3)
    print("block rel diff median %.3g max %.3g" % (np.median(rel), rel.max()), flush=True)
    assert np.median(rel) < 0.03


def test_collect_train_and_gain(ctx):
    # a small emitter faces the back wall: the room is lit by the bright patch
    # it throws there, which NEE cannot target and BSDF sampling finds only by
    # chance; every surface is rough, so the variance is in that indirect light
    sc = scenes.cornell(width=64, height=64, spp=16, bounce=True)
    r = Renderer(sc, ctx=ctx)
    r.configure_guide(seed=5, lr=2e-3, capacity=1 << 21)
    ref = _img(r, 32768, no_drop=True)
    offsets = (0, 3000, 6000, 9000, 12000, 15000)

    def err(**kw):
        # six independent 128-spp renders; both the plain relMSE and the one
        # without the 0.1% largest per-pixel errors (rare caustic-like paths
        # of the reference's own BSDF sampling dominate the raw value)
        e = [(relmse(im, ref), relmse(im, ref, trim=1e-3))
             for im in (_img(r, 128, offset=o, **kw) for o in offsets)]
        return np.mean(e, axis=0)

    e_plain = err(no_drop=True)
    e_untr = err(guided=True)
    losses = []
    for it in range(32):
        r.render(spp=8, offset=100000 + 8 * it, guided=True, collect=True)
        if it == 0:
            recs = r.guide_records()
            assert len(recs) > 1000
            assert np.all(np.isfinite(recs)) and np.all(recs[:, 14] >= 0)
            d = recs[:, 11:14]
            assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-4)
            assert np.all((recs[:, 0:3] >= 0) & (recs[:, 0:3] <= 1))
        losses.append(r.train_guide(batch=1 << 16, steps=20))
    e_tr = err(guided=True)
    print("losses", np.round(losses, 3), flush=True)
    print("relMSE at 128 spp (mean of 6 renders, raw / 0.1%%-trimmed) vs 32768-spp plain: plain %.4g / %.4g, "
          "guided untrained %.4g / %.4g, guided trained %.4g / %.4g" % (*e_plain, *e_untr, *e_tr), flush=True)
    assert np.mean(losses[-8:]) < np.mean(losses[:4])
    assert e_tr[1] < e_untr[1]
    # the trained guide beats BSDF sampling at equal samples (0.098 vs 0.125
    # measured); the margin is several times the spread of six renders
    assert e_tr[1] < 0.9 * e_plain[1]


def test_fp8_guided_render_is_unbiased(ctx):
    from ratbite_amd import _lib as L
    sc = scenes.cornell(width=48, height=48, spp=16)
    r = Renderer(sc, ctx=ctx)
    r.configure_guide(seed=4)
    plain = _img(r, 1024, no_drop=True)
    g8 = _img(r, 1024, guided=True, flags=L.RBH_RENDER_GUIDE_FP8)
    assert np.all(np.isfinite(g8))
    mp, mg = plain.mean(axis=(0, 1)), g8.mean(axis=(0, 1))
    print("image mean plain", mp, "guided fp8", mg, flush=True)
    assert np.all(np.abs(mg - mp) / mp < 0.03)


def test_guided_1080p_spheres_is_unbiased():
    """The learned guide on the 1920x1080 spheres workload as configured for
    guiding (one-sample MIS at selection 0.5, 40 trained outputs, Adam at lr
    1e-3 on batches of 2^18 records from the first 32 spp): its image agrees
    with the plain no-drop image of the same number of samples within Monte
    Carlo noise, over the whole frame and over 120x120 blocks."""
    sc = scenes.spheres(spp=16)
    r = Renderer(sc)
    try:
        r.configure_guide(seed=1, lr=1e-3, capacity=1 << 22, fixed_selection=True)
        losses = []
        for it in range(8):
            r.render(spp=4, offset=4 * it, guided=True, collect=True)
            losses.append(r.train_guide(batch=1 << 18, steps=8))
        print("losses", np.round(losses, 3), flush=True)
        assert np.all(np.isfinite(losses))
        plain = _img(r, 64, offset=1 << 16, no_drop=True)
        guided = _img(r, 64, offset=1 << 16, guided=True)
        assert np.all(np.isfinite(guided))
        mp, mg = plain.mean(axis=(0, 1)), guided.mean(axis=(0, 1))
        print("image mean plain", mp, "guided", mg, flush=True)
        assert np.all(np.abs(mg - mp) / mp < 0.02)
        bp = plain.reshape(9, 120, 16, 120, 3).mean(axis=(1, 3))
        bg = guided.reshape(9, 120, 16, 120, 3).mean(axis=(1, 3))
        rel = np.abs(bg - bp) / np.maximum(bp, 1e-3)
        print("block rel diff median %.3g p90 %.3g" % (np.median(rel), np.percentile(rel, 90)), flush=True)
        assert np.median(rel) < 0.03
    finally:
        r.close()
