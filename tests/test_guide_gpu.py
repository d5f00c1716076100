"""The guiding network's BF16 MFMA kernels against the FP64 oracle
(oracle/guide_oracle.py): forward outputs, loss and gradients on identical
parameters and records, with the oracle rounding layer inputs and weights to
bfloat16 where the kernels store them; and training on a known target."""
import ctypes

import numpy as np
import pytest

import guide_oracle as G
from ratbite_amd import _lib as L

pytestmark = pytest.mark.gpu


def _configure(ctx, seed=1, lr=1e-3, fixed=False):
    cfg = L.GuideConfig()
    cfg.lr, cfg.beta1, cfg.beta2, cfg.eps = lr, 0.9, 0.999, 1e-8
    cfg.sceneMin[:] = [0, 0, 0]
    cfg.sceneMax[:] = [1, 1, 1]
    cfg.seed = seed
    cfg.recordCapacity = 0
    cfg.fixedSelection = int(fixed)
    ctx.check(L.lib().rbh_guide_configure(ctx.h, ctypes.byref(cfg)))


def _records(n, rng, sharp=False):
    rec = np.zeros((n, G.REC), np.float32)
    rec[:, 0:3] = rng.random((n, 3))
    for a, b in ((3, 6), (6, 9)):
        v = rng.normal(size=(n, 3)); v /= np.linalg.norm(v, axis=1, keepdims=True)
        rec[:, a:b] = v
    rec[:, 9] = rng.random(n)
    rec[:, 10] = rng.random(n)
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 11:14] = d
    rec[:, 14] = rng.exponential(size=n)
    rec[:, 15] = rng.uniform(0.0, 0.6, n) * (rng.random(n) < 0.8)   # BSDF-sampler density
    return rec


def _params(seed, scale_out=1.0):
    ps = G.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    ps = [(W, rng.normal(size=b.shape) * 0.1) for W, b in ps]
    ps[-1] = (ps[-1][0] * scale_out, ps[-1][1])
    return np.ascontiguousarray(G.flatten(ps), np.float32)


def _set(ctx, flat):
    ctx.check(L.lib().rbh_guide_set_params(ctx.h, L.ptr(flat), len(flat)))


def test_params_round_trip(ctx):
    _configure(ctx)
    p = _params(3)
    _set(ctx, p)
    back = np.zeros_like(p)
    ctx.check(L.lib().rbh_guide_get_params(ctx.h, L.ptr(back), len(back)))
    assert np.array_equal(back, p)


@pytest.mark.parametrize("n", [1, 37, 64, 1000, 4096])
def test_forward_matches_oracle(ctx, n):
    _configure(ctx)
    p = _params(4)
    _set(ctx, p)
    rng = np.random.default_rng(n)
    rec = _records(n, rng)
    out = np.zeros((n, G.N_OUT), np.float32)
    ctx.check(L.lib().rbh_guide_infer(ctx.h, L.ptr(rec), n, L.ptr(out)))
    ps = G.unflatten(p.astype(np.float64))
    ref, _ = G.forward(ps, G.encode(rec), emulate_bf16=True)
    ref64, _ = G.forward(ps, G.encode(rec))
    scale = np.abs(ref).max()
    err = np.abs(out - ref).max() / scale
    err64 = np.abs(out - ref64).max() / scale
    print("n=%d forward max err vs bf16-emulated oracle %.2e, vs fp64 %.2e" % (n, err, err64), flush=True)
    assert err < 2e-2
    assert np.sqrt(np.mean((out - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)) < 2e-3
    assert err64 < 6e-2


def test_gradient_matches_oracle(ctx):
    _configure(ctx)
    p = _params(5, scale_out=2.0)
    _set(ctx, p)
    rng = np.random.default_rng(6)
    n = 4096
    rec = _records(n, rng)
    grad = np.zeros(L.GUIDE_PARAMS, np.float32)
    loss = ctypes.c_float()
    ctx.check(L.lib().rbh_guide_gradient(ctx.h, L.ptr(rec), n, L.ptr(grad), ctypes.byref(loss)))
    ps = G.unflatten(p.astype(np.float64))
    l_ref, g_ref = G.loss_grad(ps, rec.astype(np.float64), emulate_bf16=True)
    l64, g64 = G.loss_grad(ps, rec.astype(np.float64))
    gr = G.flatten(g_ref)
    g6 = G.flatten(g64)
    print("loss gpu %.6f oracle(bf16) %.6f fp64 %.6f" % (loss.value, l_ref, l64), flush=True)
    assert abs(loss.value - l_ref) / abs(l_ref) < 2e-3
    off = 0
    for li, (fi, fo) in enumerate(G.LAYERS):
        for name, size in (("W", fi * fo), ("b", fo)):
            a, b, c = grad[off:off + size], gr[off:off + size], g6[off:off + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            rel64 = np.linalg.norm(a - c) / np.linalg.norm(c)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            print("layer %d %s: rel L2 vs bf16 oracle %.2e, vs fp64 %.2e, cosine %.6f" % (li, name, rel, rel64, cos),
                  flush=True)
            assert rel < 2e-2 and cos > 0.9998
            # BF16 storage of activations and gradients (8-bit mantissa) moves
            # the gradient this far from the FP64 one
            assert rel64 < 0.3
            off += size


def test_fixed_selection_mode_matches_oracle(ctx):
    """Fixed selection a = 0.5: set_params zeroes the selection row, the
    selection output is exactly zero, the gradient matches the oracle's
    fixed-mode gradient (zero on the selection row) and Adam leaves the row
    at zero."""
    _configure(ctx, fixed=True)
    p = _params(5, scale_out=2.0)
    _set(ctx, p)
    back = np.zeros_like(p)
    ctx.check(L.lib().rbh_guide_get_params(ctx.h, L.ptr(back), len(back)))
    ps = G.unflatten(p.astype(np.float64))
    want = G.flatten(G.fix_selection(ps)).astype(np.float32)
    assert np.array_equal(back, want)
    rng = np.random.default_rng(6)
    n = 4096
    rec = _records(n, rng)
    out = np.zeros((n, G.N_OUT), np.float32)
    ctx.check(L.lib().rbh_guide_infer(ctx.h, L.ptr(rec), n, L.ptr(out)))
    assert np.all(out[:, G.N_MIX] == 0.0)
    grad = np.zeros(L.GUIDE_PARAMS, np.float32)
    loss = ctypes.c_float()
    ctx.check(L.lib().rbh_guide_gradient(ctx.h, L.ptr(rec), n, L.ptr(grad), ctypes.byref(loss)))
    l_ref, g_ref = G.loss_grad(ps, rec.astype(np.float64), emulate_bf16=True, fixed_selection=True)
    gr = G.flatten(g_ref)
    print("fixed selection: loss gpu %.6f oracle(bf16) %.6f" % (loss.value, l_ref), flush=True)
    assert abs(loss.value - l_ref) / abs(l_ref) < 2e-3
    fi, fo = G.LAYERS[-1]
    w0 = len(gr) - fi * fo - fo
    sel_rows = np.r_[w0 + G.N_MIX * fi: w0 + (G.N_MIX + 1) * fi, len(gr) - 1]
    assert np.all(grad[sel_rows] == 0.0)
    rel = np.linalg.norm(grad - gr) / np.linalg.norm(gr)
    cos = grad @ gr / (np.linalg.norm(grad) * np.linalg.norm(gr))
    print("fixed selection gradient: rel L2 %.2e cosine %.6f" % (rel, cos), flush=True)
    assert rel < 2e-2 and cos > 0.9998
    ctx.check(L.lib().rbh_guide_train(ctx.h, L.ptr(rec), n, 1024, 5, ctypes.byref(loss)))
    ctx.check(L.lib().rbh_guide_get_params(ctx.h, L.ptr(back), len(back)))
    assert np.all(back[sel_rows] == 0.0) and not np.array_equal(back, want)


def test_adam_training_reaches_target(ctx):
    # directions from one vMF lobe (kappa 20) around +z, uniform weights, no
    # BSDF density: the optimum loss is the target's negative entropy
    # -log(20/2pi) + 1 = -0.158 plus -log 0.9 (the largest selection) = -0.053
    _configure(ctx, seed=7, lr=3e-3)
    rng = np.random.default_rng(8)
    n = 1 << 16
    rec = _records(n, rng)
    k = 20.0
    u = rng.random((n, 2))
    cz = 1 + np.log(u[:, 0] + (1 - u[:, 0]) * np.exp(-2 * k)) / k
    sz = np.sqrt(1 - cz * cz)
    rec[:, 11:14] = np.stack([sz * np.cos(2 * np.pi * u[:, 1]), sz * np.sin(2 * np.pi * u[:, 1]), cz], 1)
    rec[:, 14] = 1.0
    rec[:, 15] = 0.0
    first = ctypes.c_float()
    ctx.check(L.lib().rbh_guide_train(ctx.h, L.ptr(rec), n, 4096, 1, ctypes.byref(first)))
    last = ctypes.c_float()
    ctx.check(L.lib().rbh_guide_train(ctx.h, L.ptr(rec), n, 4096, 300, ctypes.byref(last)))
    out = np.zeros((256, G.N_OUT), np.float32)
    ctx.check(L.lib().rbh_guide_infer(ctx.h, L.ptr(rec[:256]), 256, L.ptr(out)))
    sel = G.selection(out.astype(np.float64))
    print("loss first %.4f after 300 Adam steps %.4f, selection %.3f" % (first.value, last.value, sel.mean()),
          flush=True)
    assert first.value > 1.0 and last.value < 0.1
    # without a BSDF density every record pushes the selection to its maximum
    assert sel.min() > 0.8


def test_fp8_inference_close_to_bf16(ctx):
    _configure(ctx)
    p = _params(9)
    _set(ctx, p)
    rng = np.random.default_rng(10)
    n = 4096
    rec = _records(n, rng)
    a = np.zeros((n, G.N_OUT), np.float32)
    b = np.zeros((n, G.N_OUT), np.float32)
    ctx.check(L.lib().rbh_guide_infer(ctx.h, L.ptr(rec), n, L.ptr(a)))
    ctx.check(L.lib().rbh_guide_infer_fp8(ctx.h, L.ptr(rec), n, L.ptr(b)))
    ps = G.unflatten(p.astype(np.float64))
    ref, _ = G.forward(ps, G.encode(rec))
    e16 = np.sqrt(np.mean((a - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    e8 = np.sqrt(np.mean((b - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    print("relative RMS error vs fp64: bf16 %.3g fp8 %.3g" % (e16, e8), flush=True)
    assert e16 < 1e-2
    # E4M3 keeps 3 mantissa bits: a few percent per element, compounded over five layers
    assert e8 < 0.15
