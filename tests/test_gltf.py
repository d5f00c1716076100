"""glTF ingestion (ratbite_amd/gltf.py) against the glTF 2.0 specification's
meaning of the fixture written by tests/gltf_fixture.py and the mapping of
the reference's loader (src/core/scene.cpp:92-641,
rendering_interface.cpp:47-307): world positions from the node hierarchy
after the change of coordinates (x, y, z) -> (-x, z, y); indices of every
width; the tangent-frame quaternion and octahedral-normal attributes decoded
back to the transformed normals and tangents; texture coordinates (float
and normalised u16); materials, textures and samplers; emitters and their
flux (factor x strength x texture emission x model-space area x pi x
|det|); the .glb container gives the same arrays as .gltf + .bin + PNG;
and a render of the loaded scene on the CPU build of the product's kernels
agrees with the oracle's."""
import math
import os
import sys

import numpy as np
import pytest

import gltf_fixture as F
from ratbite_amd import gltf, scenes
from ratbite_amd import _lib as L

C3 = np.array([[-1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("gltf")
    a, b = str(d / "scene.gltf"), str(d / "scene.glb")
    F.write_gltf(a)
    F.write_glb(b)
    return a, b


@pytest.fixture(scope="module")
def loaded(files):
    return gltf.load(files[0])


def _trs(t=(0, 0, 0), r=(0, 0, 0, 1), s=(1, 1, 1)):
    x, y, z, w = r
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(s)
    M[:3, 3] = t
    return M


def _instances():
    """(mesh name, glTF world matrix) in the loader's depth-first order."""
    rot = (0.0, math.sin(0.3), 0.0, math.cos(0.3))
    root = _trs((0.5, -1.0, -4.0), rot, (1.5, 1.5, 1.5))
    quad = root @ _trs((0.0, 0.5, 0.0))
    boxes = root @ np.array([[1, 0, 0, -1.5], [0, 1, 0, 0.2], [0, 0, 1, 0.3], [0, 0, 0, 1]], float)
    em_a = boxes @ _trs((0.0, 2.0, 0.0), s=(0.5, 0.5, 0.5))
    em_b = _trs((2.0, 2.0, -1.0), s=(0.25, 0.5, 1.0))
    tris = _trs((-2.0, -0.5, 1.0))
    return [("quad", quad), ("box", boxes), ("emitter", em_a), ("emitter", em_b), ("tris", tris)]


def _local_positions(name):
    if name in ("quad", "emitter"):
        return F.quad_arrays()[0]
    if name == "box":
        return F.box_arrays()[0]
    return np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [2, 1, 0], [1, 2, 0]], np.float32)


def test_positions_follow_the_node_hierarchy(loaded):
    off = 0
    for name, M in _instances():
        p = _local_positions(name).astype(np.float64)
        want = (C3 @ (p @ M[:3, :3].T + M[:3, 3]).T).T
        got = loaded["verts"][off:off + len(p)]
        assert np.allclose(got, want, atol=1e-5), name
        off += len(p)
    assert off == len(loaded["verts"])


def test_indices_of_every_width(loaded):
    qi = F.quad_arrays()[4].reshape(-1, 3)
    bi = F.box_arrays()[3].reshape(-1, 3)
    want = [qi, bi + 4, qi + 28, qi + 32, np.arange(6).reshape(-1, 3) + 36]
    assert np.array_equal(loaded["indices"], np.concatenate(want))
    assert np.array_equal(loaded["tri_material"], np.array([0, 0] + [1] * 12 + [2, 2, 2, 2, 3, 3]))


def _quat_cols(q):
    x, y, z, w = q.T
    c0 = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], 1)
    c2 = np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], 1)
    return c0, c2


def _oct_decode(e):
    x, y = e[:, 0].astype(np.float64), e[:, 1].astype(np.float64)
    z = 1.0 - np.abs(x) - np.abs(y)
    t = np.maximum(-z, 0.0)
    x = x + np.where(x >= 0, -t, t)
    y = y + np.where(y >= 0, -t, t)
    v = np.stack([x, y, z], 1)
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_vertex_attributes(loaded):
    a = loaded["attribs"].view(np.float32).reshape(-1, 8)
    inst = _instances()
    # quad: tangent frames
    M = inst[0][1][:3, :3]
    n_want = C3 @ (np.linalg.inv(M).T @ np.array([0, 0, 1.0]))
    n_want /= np.linalg.norm(n_want)
    t_want = C3 @ (M @ np.array([1.0, 0, 0]))
    t_want /= np.linalg.norm(t_want)
    q = a[0:4, 0:4].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    t, n = _quat_cols(q)
    assert np.allclose(n, n_want, atol=1e-6) and np.allclose(t, t_want, atol=1e-6)
    assert np.all(q[:, 3] >= 0)                                  # right-handed (tangent w = +1)
    uv = F.quad_arrays()[3]
    assert np.array_equal(a[0:4, 4:6], uv)
    uv2 = (np.clip(uv * 0.5 + 0.25, 0, 1) * 65535).round().astype(np.uint16)
    assert np.array_equal(a[0:4, 6:8], uv2.astype(np.float32) * np.float32(1.0 / 65535.0))
    # box: octahedral normals
    M = inst[1][1][:3, :3]
    bn = F.box_arrays()[1].astype(np.float64)
    want = (C3 @ (np.linalg.inv(M).T @ bn.T)).T
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.allclose(_oct_decode(a[4:28, 0:2]), want, atol=1e-5)


def test_materials(loaded):
    m = loaded["materials"]
    assert len(m) == 4
    S = scenes
    m0 = m[0]
    assert m0["textures"] == S.TX_BASE_COLOR | S.TX_MET_ROUGH | S.TX_NORMAL | S.TX_TRANSMISSION | S.TX_SHEEN_COLOR
    assert list(m0["tex"]) == [0, 1, 2, 3, 0, -1, -1]
    assert m0["texCoordSets"] == (1 << 1) | (1 << 2) | (1 << 4)
    assert m0["factors"] == (S.FC_BASE_COLOR | S.FC_METALNESS | S.FC_CUTOFF | S.FC_TRANSMISSION |
                             S.FC_SHEEN_COLOR | S.FC_SHEEN_ROUGH)
    assert m0["flags"] == S.MF_DOUBLE_SIDED | S.MF_SHEEN
    assert np.allclose(m0["baseColor"], [1, 0.8, 0.6, 1]) and m0["metalness"] == np.float32(0.25)
    assert m0["ior"] == np.float32(1.33) and m0["alphaCutoff"] == np.float32(0.4)
    assert m0["transmission"] == np.float32(0.5) and m0["sheenRoughness"] == np.float32(0.3)
    assert m0["attributes"] == S.AT_FRAME | S.AT_TEX1 | S.AT_TEX2
    m1 = m[1]
    assert m1["textures"] == 0
    assert m1["factors"] == S.FC_BASE_COLOR | S.FC_METALNESS | S.FC_ROUGHNESS | S.FC_BLEND
    assert m1["metalness"] == 0.0 and m1["roughness"] == np.float32(0.45)
    assert m1["attributes"] == S.AT_NORMAL | S.AT_TEX1
    m2 = m[2]
    assert m2["textures"] == S.TX_EMISSION and m2["tex"][6] == 4 and m2["factors"] == S.FC_EMISSION
    assert np.allclose(m2["emissive"], [4.0, 2.0, 1.0])
    assert m2["attributes"] == S.AT_TEX1
    assert m[3]["factors"] == S.FC_EMISSION and m[3]["attributes"] == 0
    assert np.all(m["bxdf"] == S.BXDF_COMPLEX) and np.all(m["attributeStride"] == 32)


def test_textures_and_samplers(loaded):
    d = loaded["textures"]
    ims = F.images()
    assert list(d["format"]) == [0, 0, 0, 1, 0]
    assert list(d["srgb"]) == [1, 0, 0, 0, 0]                    # only the base colour is sRGB
    assert list(d["filter"]) == [1, 0, 1, 0, 1]
    assert [tuple(x[:2]) for x in d["addr"]] == [(0, 0), (1, 2), (0, 0), (1, 2), (0, 0)]
    blob = loaded["texels"]

    def texels(i, shape):
        o = int(d[i]["offset"])
        return blob[o:o + int(np.prod(shape))].reshape(shape)

    assert np.array_equal(texels(0, (16, 16, 4)), ims["base"])
    nm = np.concatenate([ims["normal"], np.full((8, 8, 1), 255, np.uint8)], -1)
    assert np.array_equal(texels(1, (8, 8, 4)), nm)
    assert np.array_equal(texels(3, (8, 8)), ims["grey"])
    assert all(int(x) % 256 == 0 for x in d["offset"])


def test_emitters_and_flux(loaded):
    lt = loaded["light_tris"]
    assert list(lt) == [14, 15, 16, 17, 18, 19]
    inst = _instances()
    em_mean = (200 + 100 + 50) / 255.0 / 3.0
    want = []
    for name, M in inst[2:4]:
        det = abs(np.linalg.det(M[:3, :3]))
        want += [(4 + 2 + 1) / 3.0 * em_mean * 2.0 * math.pi * det] * 2
    want += [2.0 * 0.5 * math.pi] * 2
    assert np.allclose(loaded["flux"], want, rtol=2e-6)
    uv = F.quad_arrays()[3]
    qi = F.quad_arrays()[4].reshape(-1, 3)
    assert np.array_equal(loaded["light_uvs"][:2], uv[qi].reshape(-1, 6))


def test_glb_equals_gltf(files, loaded):
    g = gltf.load(files[1])
    for k in ("verts", "indices", "attribs", "tri_material", "materials", "textures", "texels", "light_tris",
              "light_uvs", "flux"):
        assert np.array_equal(g[k], loaded[k]), k


def test_bad_files_are_refused(tmp_path):
    p = tmp_path / "bad.gltf"
    p.write_text('{"asset": {"version": "2.0"}, "meshes": [], "scenes": [{"nodes": []}]}')
    with pytest.raises(gltf.GltfError):
        gltf.load(str(p))


def test_loaded_scene_renders_like_the_oracle(files, tmp_path):
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostbuild"))
    import host_wavefront as H
    import oracle_lib as O
    from ratbite_amd import hdr
    from ratbite_amd.envmaps import lognormal_envmap
    H.build()
    env = lognormal_envmap(64, 32, sigma=1.0, seed=3, sun=False)
    hp = str(tmp_path / "sky.hdr")
    hdr.write_hdr(hp, env)
    sc = gltf.scene(files[0], env=hp, width=48, height=40, spp=4, depth=6)
    assert sc.env.shape == (32, 64, 4)
    lp = sc.launch_params(offset=0, count=4)
    a, st = H.render(sc, lp)
    o, ost = O.render(sc, lp)
    g, r = a[..., :3] / a[..., 3:4], o[..., :3] / o[..., 3:4]
    assert np.isfinite(g).all() and np.isfinite(r).all()
    close = (np.abs(g - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3)).all(axis=2).mean()
    print("glTF scene: %.4f of pixels within 1e-4 of the oracle" % close)
    assert close >= 0.999
    assert st["closestRays"] == ost["closestRays"]
    assert (g[..., 1] > 0).mean() > 0.5
