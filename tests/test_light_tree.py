"""Light tree (src/core/light_tree.cpp, src/kernels/pathtrace.cu:974-1112).

* the product's host builder (rbh_light_tree_build_host) and the oracle's
  independent restatement produce the same topology, nodes and masks bit for
  bit;
* structural invariants of the reference (every bitmask reaches the leaf that
  holds its light, <= 2n-1 nodes, <= 8 lights per leaf, flux conserved);
* the oracle's PDF/sampling pair is pinned by invariants: the PDFs of all
  lights sum to 1 - envImportance, the PDF reported by sampling equals the
  traversal PDF of the sampled light, and sampling frequencies follow the PDF;
* the FP64 power mode (RBH_LIGHT_TREE_FP64): every node's flux is the float
  rounding of the exact sum of its lights' fluxes (math.fsum), and the FP32
  builder's flux sums stay within a stated error of the exact sums.
"""
import math
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import _lib as L
from ratbite_amd import scenes


def _lights(n, seed):
    sc = scenes.spheres(n_spheres=1, level=1, n_lights=n, width=8, height=8, env_size=None, seed=seed)
    return sc.lights, sc.flux


def product_build(lights, flux, flags=0):
    n = len(lights)
    nodes = np.zeros(max(1, 2 * n), L.PACKED_NODE_DTYPE)
    lp = np.zeros(max(1, n), np.uint32)
    bm = np.zeros(max(1, n), np.uint64)
    r = L.lib().rbh_light_tree_build_host_ex(L.ptr(np.ascontiguousarray(lights)),
                                             L.ptr(np.ascontiguousarray(flux, np.float32)), n, L.ptr(nodes), L.ptr(lp),
                                             L.ptr(bm), flags)
    assert r >= 0, r
    return nodes[:r], lp[:n], bm[:n]


def subtree_ranges(nodes):
    """[lo, hi) of lightPointers under every node (leaves hold contiguous
    ranges in depth-first order)."""
    core = nodes["coreData"].astype(np.int64)
    lo = np.zeros(len(nodes), np.int64)
    hi = np.zeros(len(nodes), np.int64)
    for i in range(len(nodes) - 1, -1, -1):
        c = int(core[i])
        if c >> 31:
            lo[i] = c & ((1 << 27) - 1)
            hi[i] = lo[i] + ((c >> 27) & 15)
        else:
            lo[i], hi[i] = lo[i + 1], hi[c]
    return lo, hi


def exact_node_flux(nodes, lp, flux):
    ids = lp.astype(np.int64) & ((1 << 30) - 1)
    f = flux.astype(np.float64)[ids]
    lo, hi = subtree_ranges(nodes)
    return np.array([math.fsum(f[a:b]) for a, b in zip(lo, hi)])


@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (9, 3), (100, 4), (2000, 5)])
def test_builder_bit_exact_with_oracle(n, seed):
    lights, flux = _lights(n, seed)
    a = product_build(lights, flux)
    b = O.light_tree_build(lights, flux)
    assert len(a[0]) == len(b[0])
    assert a[0].tobytes() == b[0].tobytes()
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def _walk(nodes, mask):
    i, depth = 0, 0
    while (int(nodes[i]["coreData"]) >> 31) == 0:
        i = int(nodes[i]["coreData"]) if (int(mask) >> depth) & 1 else i + 1
        depth += 1
    core = int(nodes[i]["coreData"])
    return i, (core >> 27) & 15, core & ((1 << 27) - 1)


def test_structural_invariants():
    lights, flux = _lights(3000, 11)
    nodes, lp, bm = product_build(lights, flux)
    n = len(lights)
    assert len(nodes) <= 2 * n - 1
    for light in range(0, n, 7):
        leaf, cnt, off = _walk(nodes, bm[light])
        assert 1 <= cnt <= 8
        assert light in set(int(x) & ((1 << 30) - 1) for x in lp[off:off + cnt])
    assert abs(float(nodes[0]["flux"]) - float(flux.astype(np.float64).sum())) / float(flux.sum()) < 1e-4


@pytest.mark.parametrize("env", [0, 1])
def test_pdf_partition_of_unity(env):
    lights, flux = _lights(500, 12)
    nodes, lp, bm = O.light_tree_build(lights, flux)
    n = len(lights)
    rng = np.random.default_rng(3)
    for _ in range(5):
        pos = rng.uniform([-60, -8, -120], [60, 25, -10]).astype(np.float32)
        nrm = rng.normal(size=3).astype(np.float32)
        nrm /= np.linalg.norm(nrm)
        pn = np.ascontiguousarray(np.tile(np.concatenate([pos, nrm]), (n, 1)), np.float32)
        ids = np.arange(n, dtype=np.uint32)
        pdf = np.zeros(n, np.float32)
        O.lib().orc_light_tree_pdf(O.p(nodes), O.p(lp), O.p(bm), env, O.p(ids), O.p(pn), O.p(pdf), n)
        total = pdf.astype(np.float64).sum()
        assert abs(total - (1.0 - 0.3 * env)) < 1e-4


def test_sample_pdf_consistency_and_frequencies():
    lights, flux = _lights(40, 13)
    nodes, lp, bm = O.light_tree_build(lights, flux)
    n = len(lights)
    rng = np.random.default_rng(4)
    pos = np.array([0.0, 5.0, -50.0], np.float32)
    nrm = np.array([0.0, 1.0, 0.0], np.float32)
    m = 200000
    inp = np.zeros((m, 7), np.float32)
    inp[:, 0] = (np.arange(m) + rng.random(m)) / m     # stratified
    inp[:, 1:4], inp[:, 4:7] = pos, nrm
    ti = np.zeros((m, 2), np.uint32)
    sp = np.zeros(m, np.float32)
    O.lib().orc_light_tree_sample(O.p(nodes), O.p(lp), O.p(bm), 0, O.p(inp), O.p(ti), O.p(sp), m)
    assert np.all(ti[:, 0] == 0)
    ids = ti[:, 1]
    pn = np.ascontiguousarray(np.tile(np.concatenate([pos, nrm]), (n, 1)), np.float32)
    pdf = np.zeros(n, np.float32)
    O.lib().orc_light_tree_pdf(O.p(nodes), O.p(lp), O.p(bm), 0, O.p(np.arange(n, dtype=np.uint32)), O.p(pn), O.p(pdf), n)
    np.testing.assert_allclose(sp, pdf[ids], rtol=1e-5)
    freq = np.bincount(ids, minlength=n) / m
    np.testing.assert_allclose(freq, pdf, atol=3e-3)


def _fp64_setup(n_lights=2000, n=20000, seed=4):
    import light_tree64 as T
    lights, flux = _lights(n_lights, 5)
    nodes, lp, bm = O.light_tree_build(lights, flux)
    rng = np.random.default_rng(seed)
    pos = rng.uniform([-60, -8, -120], [60, 25, -10], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    ids = rng.integers(0, len(lights), n).astype(np.uint32)
    return T, nodes, lp, bm, pn, ids, len(lights)


def fp32_error_summary(p32, p64):
    """Relative error of FP32 light pdfs against FP64: median, 99th
    percentile over pdfs >= 1e-4, and the mean weighted by the FP64 pdf (the
    error a sampled light sees on average)."""
    rel = np.abs(p32.astype(np.float64) - p64) / p64
    big = p64 >= 1e-4
    return float(np.median(rel)), float(np.quantile(rel[big], 0.99)), float(np.sum(p64 * rel) / np.sum(p64))


def test_fp64_light_pdf_partition_and_fp32_error():
    # the FP64 restatement (exact erf/expm1) is a distribution over all lights
    # at any shading point, and the FP32 oracle (the reference's formulation,
    # erfc = 1 - erf from an FP32 fit) stays close to it where lights are
    # likely; unlikely lights of a node much dimmer than its sibling carry
    # large relative errors (cancellation in 1 - erf), harmless to the
    # estimator, which samples and weights with the same FP32 pdf
    T, nodes, lp, bm, pn, ids, nl = _fp64_setup()
    all_ids = np.arange(nl, dtype=np.uint32)
    for k in range(3):
        total = T.pdf(nodes, bm, all_ids, np.tile(pn[k], (nl, 1))).sum()
        assert abs(total - 1.0) < 1e-9, total
    p32 = np.zeros(len(ids), np.float32)
    O.lib().orc_light_tree_pdf(O.p(nodes), O.p(lp), O.p(bm), 0, O.p(ids), O.p(pn), O.p(p32), len(ids))
    p64 = T.pdf(nodes, bm, ids, pn)
    med, q99, wmean = fp32_error_summary(p32, p64)
    print("fp32 oracle vs fp64: median %.2g, q99 (pdf >= 1e-4) %.2g, pdf-weighted mean %.2g" % (med, q99, wmean))
    assert med < 1e-5 and q99 < 5e-3 and wmean < 2e-4


@pytest.mark.parametrize("n,seed", [(2000, 5), (100000, 6)])
def test_fp64_power_mode_node_flux(n, seed):
    # the FP64 mode rounds each node's flux once from an exact-enough double
    # sum; the reference's FP32 builder sums packed (rounded) child fluxes
    lights, flux = _lights(n, seed)
    flux = np.ascontiguousarray(flux, np.float32)
    nodes64, lp64, bm64 = product_build(lights, flux, L.RBH_LIGHT_TREE_FP64)
    ex64 = exact_node_flux(nodes64, lp64, flux)
    assert np.array_equal(nodes64["flux"], ex64.astype(np.float32))
    nodes32, lp32, _ = product_build(lights, flux)
    ex32 = exact_node_flux(nodes32, lp32, flux)
    rel32 = np.abs(nodes32["flux"].astype(np.float64) - ex32) / ex32
    print("FP32 builder node flux vs exact sums (%d lights): max rel %.3g, mean %.3g" % (n, rel32.max(), rel32.mean()))
    assert rel32.max() < 1e-5
    # the FP64-mode tree is a valid tree: every light reachable, a partition of unity
    for light in range(0, n, max(1, n // 500)):
        leaf, cnt, off = _walk(nodes64, bm64[light])
        assert light in set(int(x) & ((1 << 30) - 1) for x in lp64[off:off + cnt])
    assert np.array_equal(np.sort(lp64.astype(np.int64) & ((1 << 30) - 1)), np.arange(n))


def test_fp64_power_mode_rejects_unknown_flags():
    lights, flux = _lights(10, 1)
    n = len(lights)
    nodes = np.zeros(2 * n, L.PACKED_NODE_DTYPE)
    lp, bm = np.zeros(n, np.uint32), np.zeros(n, np.uint64)
    r = L.lib().rbh_light_tree_build_host_ex(L.ptr(lights), L.ptr(np.ascontiguousarray(flux, np.float32)), n,
                                             L.ptr(nodes), L.ptr(lp), L.ptr(bm), 2)
    assert r == L.RBH_ERR_INVALID


def test_fp64_light_pdf_error_at_100k_lights():
    # C5-sized tree (100k emissive triangles, built in the FP64 power mode):
    # the FP32 pdf (the reference's formulation, equal to the GPU's within
    # 1e-5) against the FP64 restatement, at 20k random shading points
    import light_tree64 as T
    lights, flux = _lights(100000, 7)
    nodes, lp, bm = product_build(lights, flux, L.RBH_LIGHT_TREE_FP64)
    rng = np.random.default_rng(8)
    n = 20000
    pos = rng.uniform([-60, -8, -120], [60, 25, -10], (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pn = np.ascontiguousarray(np.concatenate([pos, nrm], 1), np.float32)
    ids = rng.integers(0, len(lights), n).astype(np.uint32)
    p32 = np.zeros(n, np.float32)
    O.lib().orc_light_tree_pdf(O.p(nodes), O.p(lp), O.p(bm), 0, O.p(ids), O.p(pn), O.p(p32), n)
    p64 = T.pdf(nodes, bm, ids, pn)
    ok = p64 > 0
    med, q99, wmean = fp32_error_summary(p32[ok], p64[ok])
    print("100k lights, fp32 vs fp64 pdf: median %.2g, q99 (pdf >= 1e-4) %.2g, pdf-weighted mean %.2g"
          % (med, q99, wmean))
    assert med < 1e-5 and wmean < 5e-4
