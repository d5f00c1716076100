"""The product's per-path kernel bodies (ratbite_amd/csrc/device/wavefront.hpp
and the traversal of device/traverse.hpp), compiled for the host by
tests/hostbuild and run on the CPU: against the oracle, through the furnace,
and under AddressSanitizer + UndefinedBehaviorSanitizer (out-of-range
accesses and invalid float-to-integer conversions of the product code fail
the run with their source line)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from ratbite_amd import _lib as L
from ratbite_amd import scenes

HB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostbuild")
sys.path.insert(0, HB)
import host_wavefront as H  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def built():
    H.build()


def _close_frac(a, o):
    g, r = a[..., :3] / a[..., 3:4], o[..., :3] / o[..., 3:4]
    fin = np.isfinite(g).all(axis=2) & np.isfinite(r).all(axis=2)
    close = np.abs(g - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3)
    return close.all(axis=2)[fin].mean(), fin.mean()


def test_product_kernels_on_cpu_match_oracle_cornell():
    sc = scenes.cornell(width=48, height=48, spp=4)
    lp = sc.launch_params(offset=0, count=4)
    a, st = H.render(sc, lp)
    o, ost = O.render(sc, lp)
    frac, fin = _close_frac(a, o)
    print("cornell: %.4f of pixels within 1e-4 of the oracle" % frac)
    assert fin == 1.0 and frac >= 0.999
    assert st["closestRays"] == ost["closestRays"] and st["shadowRays"] == ost["shadowRays"]
    assert np.all(a[..., 3] == 4)


def test_product_kernels_on_cpu_match_oracle_spheres():
    sc = scenes.spheres(n_spheres=10, level=3, n_lights=300, width=48, height=32, spp=4, env_size=(128, 64))
    lp = sc.launch_params(offset=0, count=4)
    a, st = H.render(sc, lp)
    o, ost = O.render(sc, lp)
    frac, fin = _close_frac(a, o)
    print("spheres: %.4f of pixels within 1e-4 of the oracle" % frac)
    assert fin > 0.99 and frac >= 0.99
    assert abs(st["closestRays"] - ost["closestRays"]) <= 0.001 * ost["closestRays"]


@pytest.mark.parametrize("flags", [0, L.RBH_RENDER_NO_DROP])
def test_product_kernels_on_cpu_match_oracle_features(flags):
    """Every surface-attribute and texture branch (quaternion frames, normal
    maps, the material textures, TEX_COORD_2, alpha cutoff and blend,
    PureDielectric, no vertex attributes): scenes.features."""
    sc = scenes.features(width=64, height=64, spp=4)
    lp = sc.launch_params(offset=0, count=4, flags=flags)
    a, st = H.render(sc, lp)
    o, ost = O.render(sc, lp)
    frac, fin = _close_frac(a, o)
    print("features: %.4f of pixels within 1e-4 of the oracle" % frac)
    assert fin == 1.0 and frac >= 0.999
    assert st["closestRays"] == ost["closestRays"] and st["shadowRays"] == ost["shadowRays"]


def test_product_kernels_furnace():
    from test_oracle_furnace import furnace_ratios

    def host(sc, lp):
        return H.render(sc, lp)[0]

    wall, ratio = furnace_ratios(host, "metal", 0.3, 128, L.RBH_RENDER_NO_DROP)
    print("product furnace metal: wall %.5f, sphere / wall %.4f" % (wall, ratio))
    assert abs(wall - 1.0) < 2e-3 and abs(ratio - 1.0) < 0.01
    # the dispersive smooth glass sphere (secondary wavelengths terminated at
    # refraction) conserves energy too
    wall, ratio = furnace_ratios(host, "glass", 0.0, 64, L.RBH_RENDER_NO_DROP)
    print("product furnace dispersive glass: wall %.5f, sphere / wall %.4f" % (wall, ratio))
    assert abs(wall - 1.0) < 2e-3 and abs(ratio - 1.0) < 0.015


SAN_SCRIPT = r'''
import sys
sys.path.insert(0, %(root)r); sys.path.insert(0, %(oracle)r); sys.path.insert(0, %(hb)r)
import numpy as np
import host_wavefront as H
from ratbite_amd import scenes, _lib as L
cases = [
    scenes.cornell(width=32, height=32, spp=2),
    scenes.cornell(width=24, height=20, spp=2, bounce=True),
    scenes.spheres(n_spheres=8, level=2, n_lights=200, width=40, height=24, spp=2, env_size=(64, 32)),
    scenes.features(width=32, height=32, spp=2),
]
for sc in cases:
    for flags in (0, L.RBH_RENDER_NO_DROP):
        a, st = H.render(sc, sc.launch_params(offset=(1 << 24) + 7, count=2, flags=flags), sanitized=True, pass_samples=1)
        assert st["queueOverflow"] == 0
print("sanitized runs done", flush=True)
'''


def test_product_kernels_under_address_and_ub_sanitizers():
    rt = H.asan_runtime()
    assert os.path.exists(rt), rt
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = SAN_SCRIPT % {"root": root, "oracle": os.path.join(root, "oracle"), "hb": HB}
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "sanitized runs done" in out, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
