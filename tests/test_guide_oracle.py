"""Pins of the FP64 guiding-network oracle (oracle/guide_oracle.py): its
analytic gradient against central finite differences, the vMF mixture's
normalisation by quadrature, its sampler against its density, that the loss
over the selection probability is smallest at the true mixing fraction, and
that Adam on the loss moves the mixture towards a known target distribution."""
import numpy as np
import pytest

import guide_oracle as G


def _records(n, rng, target=None):
    rec = np.zeros((n, G.REC))
    rec[:, 0:3] = rng.random((n, 3))
    nrm = rng.normal(size=(n, 3)); nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    rec[:, 3:6], rec[:, 6:9] = nrm, wo
    rec[:, 9] = rng.random(n)
    rec[:, 10] = rng.random(n)
    d = rng.normal(size=(n, 3)) if target is None else target(n, rng)
    rec[:, 11:14] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 14] = rng.exponential(size=n)
    # BSDF-sampler density of the direction: zero for some (outside its support)
    rec[:, 15] = rng.uniform(0.0, 0.6, n) * (rng.random(n) < 0.8)
    return rec


def test_parameter_count():
    assert G.n_params() == 80 * 64 + 3 * 64 * 64 + 64 * 41 + 4 * 64 + 41 == 20329


def test_encoding_matches_direct_sin_cos():
    rng = np.random.default_rng(0)
    rec = _records(50, rng)
    f = G.encode(rec)
    for a in range(3):
        for k in range(G.N_FREQ):
            np.testing.assert_allclose(f[:, 24 * a + 2 * k], np.sin(np.pi * 2 ** k * rec[:, a]), atol=1e-9)
            np.testing.assert_allclose(f[:, 24 * a + 2 * k + 1], np.cos(np.pi * 2 ** k * rec[:, a]), atol=1e-9)
    np.testing.assert_array_equal(f[:, 72:78], rec[:, 3:9])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    ps = G.init_params(seed=2)
    # larger output-layer weights so the lobes are sharp and the mixture uneven
    ps[-1] = (ps[-1][0] * 3.0, rng.normal(size=41) * 0.5)
    rec = _records(64, rng)
    loss, grads = G.loss_grad(ps, rec)
    flat = G.flatten(ps)
    gflat = G.flatten(grads)
    idx = rng.choice(len(flat), 120, replace=False)
    idx = np.concatenate([idx, np.arange(len(flat) - 41, len(flat))])      # every output bias
    h = 1e-6
    for i in idx:
        p1 = flat.copy(); p1[i] += h
        p2 = flat.copy(); p2[i] -= h
        l1, _ = G.loss_grad(G.unflatten(p1), rec)
        l2, _ = G.loss_grad(G.unflatten(p2), rec)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - gflat[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, gflat[i])


def test_fixed_selection_mode():
    """With the selection fixed at 0.5 the loss is the batch mean of
    w * -log(q / 2 + p_bsdf / 2) with q from the mixture density, the
    gradient still matches finite differences, and the selection row (which
    the loss no longer depends on) gets exactly zero gradient."""
    rng = np.random.default_rng(11)
    ps = G.init_params(seed=12)
    ps[-1] = (ps[-1][0] * 3.0, rng.normal(size=41) * 0.5)
    rec = _records(64, rng)
    loss, grads = G.loss_grad(ps, rec, fixed_selection=True)
    fixed = G.fix_selection(ps)
    out, _ = G.forward(fixed, G.encode(rec))
    assert np.all(out[:, G.N_MIX] == 0.0) and np.allclose(G.selection(out), 0.5)
    q = G.pdf(out, rec[:, 11:14])
    direct = np.sum(rec[:, 14] * -np.log(0.5 * q + 0.5 * rec[:, 15])) / rec[:, 14].sum()
    assert abs(loss - direct) <= 1e-12 * abs(direct)
    # the learned selection's loss differs: the mode changes the objective
    assert abs(G.loss_grad(ps, rec)[0] - loss) > 1e-3
    Wg, bg = grads[-1]
    assert np.all(Wg[G.N_MIX] == 0.0) and bg[G.N_MIX] == 0.0
    flat = G.flatten(fixed)
    gflat = G.flatten(grads)
    idx = np.concatenate([rng.choice(len(flat) - 41, 60, replace=False), np.arange(len(flat) - 41, len(flat) - 1)])
    h = 1e-6
    for i in idx:
        p1 = flat.copy(); p1[i] += h
        p2 = flat.copy(); p2[i] -= h
        l1, _ = G.loss_grad(G.unflatten(p1), rec, fixed_selection=True)
        l2, _ = G.loss_grad(G.unflatten(p2), rec, fixed_selection=True)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - gflat[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, gflat[i])


def _sphere_grid(nc, nphi):
    c = -1 + (np.arange(nc) + 0.5) * 2 / nc
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    C, PH = np.meshgrid(c, ph, indexing="ij")
    s = np.sqrt(1 - C ** 2)
    d = np.stack([s * np.cos(PH), s * np.sin(PH), C], -1).reshape(-1, 3)
    return d, (2 / nc) * (2 * np.pi / nphi)


@pytest.mark.parametrize("seed", [3, 4])
def test_mixture_pdf_normalised(seed):
    rng = np.random.default_rng(seed)
    out = rng.normal(size=(1, 41)) * 2.0
    out[0, 3:40:5] = rng.uniform(-3, 5, 8)     # log kappa up to ~150
    d, w = _sphere_grid(4000, 720)
    q = G.pdf(np.repeat(out, len(d), 0), d)
    assert abs(q.sum() * w - 1.0) < 2e-3


def test_sampler_follows_pdf():
    rng = np.random.default_rng(5)
    out = rng.normal(size=(1, 41))
    out[0, 3:40:5] = rng.uniform(0, 3, 8)
    n = 200000
    s = G.sample(np.repeat(out, n, 0), rng.random((n, 3)))
    np.testing.assert_allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)
    d, w = _sphere_grid(800, 360)
    q = G.pdf(np.repeat(out, len(d), 0), d)
    for fn in (lambda x: x[:, 0], lambda x: x[:, 2] ** 2, lambda x: x[:, 0] * x[:, 1]):
        mc = fn(s).mean()
        quad = np.sum(fn(d) * q) * w
        assert abs(mc - quad) < 5 * fn(s).std() / np.sqrt(n) + 1e-3


@pytest.mark.parametrize("beta", [0.3, 0.7])
def test_selection_loss_is_smallest_at_the_mixing_fraction(beta):
    # directions drawn with probability beta from fixed lobes, otherwise from a
    # uniform "BSDF" density: the expected loss over the selection probability
    # is minimised at a = beta (its derivative, E[(q - p) / P], vanishes there)
    rng = np.random.default_rng(int(beta * 10))
    out = rng.normal(size=(1, 41))
    out[0, 3:40:5] = rng.uniform(1, 3, 8)
    n = 200000
    pick = rng.random(n) < beta
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[pick] = G.sample(np.repeat(out, pick.sum(), 0), rng.random((pick.sum(), 3)))
    rec = np.zeros((n, G.REC))
    rec[:, 11:14] = d
    rec[:, 14] = 1.0
    rec[:, 15] = 1.0 / (4.0 * np.pi)
    res = {}
    for a in (beta - 0.08, beta, beta + 0.08):
        o = np.repeat(out, n, 0)
        o[:, 40] = -np.log((G.SEL_MAX - G.SEL_MIN) / (a - G.SEL_MIN) - 1.0)
        np.testing.assert_allclose(G.selection(o[:1]), a, rtol=1e-12)
        loss, dout = G.loss_and_dout(o, rec)
        res[a] = (loss, dout[:, 40].sum())
    lo, mid, hi = (res[a] for a in sorted(res))
    assert mid[0] < lo[0] and mid[0] < hi[0]
    assert lo[1] < 0.0 < hi[1]            # d loss / d logit changes sign across beta
    assert abs(mid[1]) < 0.1 * min(abs(lo[1]), abs(hi[1]))


def test_training_moves_towards_target():
    # target: directions from one vMF lobe around +z (kappa 20), independent of x
    def target(n, rng):
        k = 20.0
        u = rng.random((n, 2))
        cz = 1 + np.log(u[:, 0] + (1 - u[:, 0]) * np.exp(-2 * k)) / k
        sz = np.sqrt(1 - cz * cz)
        return np.stack([sz * np.cos(2 * np.pi * u[:, 1]), sz * np.sin(2 * np.pi * u[:, 1]), cz], 1)

    rng = np.random.default_rng(6)
    ps = G.init_params(seed=7)
    flat = G.flatten(ps)
    opt = G.Adam(len(flat), lr=3e-3)
    first = None
    for it in range(150):
        rec = _records(256, rng, target)
        rec[:, 14] = 1.0
        rec[:, 15] = 0.0
        loss, grads = G.loss_grad(G.unflatten(flat), rec)
        first = loss if first is None else first
        flat = opt.step(flat, G.flatten(grads))
    # with no BSDF density the optimum is the target's negative entropy,
    # -log(20 / 2pi) + 1 = -0.158, plus -log of the largest selection 0.9
    assert first > 1.5 and loss < 0.3, (first, loss)
