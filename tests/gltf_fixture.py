"""Writes a small glTF 2.0 scene (as .gltf + .bin + PNG files, or as one .glb)
that exercises what the loader maps: u8 / u16 / u32 / missing indices,
normals with and without tangents, two texture-coordinate sets (float and
normalised u16), RGBA, RGB and grey PNG images, samplers (nearest / linear,
wrap / clamp / mirror), base-colour / metal-roughness / normal /
transmission / sheen / emission textures and factors, KHR_materials_ior,
_transmission, _sheen and _emissive_strength, alpha MASK and BLEND,
doubleSided, a node hierarchy with TRS and matrix transforms, one mesh
instanced twice, and an emissive primitive with a texture."""
import io
import json
import math
import struct

import numpy as np


def _png(arr):
    from PIL import Image
    b = io.BytesIO()
    mode = {2: "L", 3: "RGB", 4: "RGBA"}[arr.shape[2] if arr.ndim == 3 else 2]
    Image.fromarray(arr if arr.ndim == 2 or arr.shape[2] != 1 else arr[..., 0], mode).save(b, "PNG")
    return b.getvalue()


class _Buf:
    def __init__(self):
        self.data = bytearray()
        self.views = []
        self.accessors = []

    def add(self, arr, component, type_, target=None, normalized=False, minmax=False):
        arr = np.ascontiguousarray(arr)
        while len(self.data) % 4:
            self.data.append(0)
        off = len(self.data)
        self.data += arr.tobytes()
        view = {"buffer": 0, "byteOffset": off, "byteLength": arr.nbytes}
        if target:
            view["target"] = target
        self.views.append(view)
        acc = {"bufferView": len(self.views) - 1, "componentType": component, "count": len(arr), "type": type_}
        if normalized:
            acc["normalized"] = True
        if minmax:
            acc["min"] = arr.min(axis=0).tolist()
            acc["max"] = arr.max(axis=0).tolist()
        self.accessors.append(acc)
        return len(self.accessors) - 1

    def add_blob(self, raw):
        while len(self.data) % 4:
            self.data.append(0)
        off = len(self.data)
        self.data += raw
        self.views.append({"buffer": 0, "byteOffset": off, "byteLength": len(raw)})
        return len(self.views) - 1


def quad_arrays(size=1.0):
    p = np.array([[-size, -size, 0], [size, -size, 0], [size, size, 0], [-size, size, 0]], np.float32)
    n = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    t = np.tile(np.array([[1, 0, 0, 1]], np.float32), (4, 1))
    uv = np.array([[0, 1], [1, 1], [1, 0], [0, 0]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3])
    return p, n, t, uv, idx


def box_arrays():
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sgn
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            v = np.cross(n, u)
            c = n * 0.5
            pts = [c - 0.5 * u - 0.5 * v, c + 0.5 * u - 0.5 * v, c + 0.5 * u + 0.5 * v, c - 0.5 * u + 0.5 * v]
            faces.append((pts, n))
    P, N, UV, I = [], [], [], []
    for pts, n in faces:
        b = len(P)
        P += pts
        N += [n] * 4
        UV += [[0, 0], [1, 0], [1, 1], [0, 1]]
        I += [b, b + 1, b + 2, b, b + 2, b + 3]
    return (np.array(P, np.float32), np.array(N, np.float32), np.array(UV, np.float32),
            np.array(I, np.uint32))


def images():
    rng = np.random.default_rng(5)
    base = np.concatenate([rng.integers(0, 256, (16, 16, 3)), np.full((16, 16, 1), 255)], -1).astype(np.uint8)
    base[::2, :, 3] = 64                                            # alpha rows for the mask
    nm = np.zeros((8, 8, 3), np.uint8)
    nm[..., 0] = 128 + rng.integers(-40, 40, (8, 8))
    nm[..., 1] = 128 + rng.integers(-40, 40, (8, 8))
    nm[..., 2] = 230
    mr = np.concatenate([np.zeros((8, 8, 1)), rng.integers(0, 256, (8, 8, 2))], -1).astype(np.uint8)
    grey = rng.integers(0, 256, (8, 8)).astype(np.uint8)
    emis = np.zeros((4, 4, 4), np.uint8)
    emis[...] = (200, 100, 50, 255)                                 # constant emission
    return {"base": base, "normal": nm, "metrough": mr, "grey": grey, "emission": emis}


def document():
    """The JSON document, its binary buffer and the images (name -> array)."""
    buf = _Buf()
    ims = images()
    j = {"asset": {"version": "2.0"}, "scene": 0, "extensionsUsed": [
        "KHR_materials_ior", "KHR_materials_transmission", "KHR_materials_sheen", "KHR_materials_emissive_strength"]}
    # meshes
    qp, qn, qt, quv, qi = quad_arrays()
    quv2 = (np.clip(quv * 0.5 + 0.25, 0, 1) * 65535).round().astype(np.uint16)
    quad = {"POSITION": buf.add(qp, 5126, "VEC3", minmax=True), "NORMAL": buf.add(qn, 5126, "VEC3"),
            "TANGENT": buf.add(qt, 5126, "VEC4"), "TEXCOORD_0": buf.add(quv, 5126, "VEC2"),
            "TEXCOORD_1": buf.add(quv2, 5123, "VEC2", normalized=True)}
    quad_idx16 = buf.add(qi.astype(np.uint16), 5123, "SCALAR")
    quad_idx8 = buf.add(qi.astype(np.uint8), 5121, "SCALAR")
    bp, bn, buv, bi = box_arrays()
    box = {"POSITION": buf.add(bp, 5126, "VEC3", minmax=True), "NORMAL": buf.add(bn, 5126, "VEC3"),
           "TEXCOORD_0": buf.add(buv, 5126, "VEC2")}
    box_idx = buf.add(bi, 5125, "SCALAR")
    tri_p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [2, 1, 0], [1, 2, 0]], np.float32)
    tri = {"POSITION": buf.add(tri_p, 5126, "VEC3", minmax=True)}
    # images in the buffer (the .glb keeps them there; the .gltf writer moves them to files)
    img_views = {k: buf.add_blob(_png(v if v.ndim == 3 else v)) for k, v in ims.items()}
    j["images"] = [{"bufferView": img_views[k], "mimeType": "image/png", "name": k}
                   for k in ("base", "normal", "metrough", "grey", "emission")]
    j["samplers"] = [{"magFilter": 9729, "wrapS": 10497, "wrapT": 10497},
                     {"magFilter": 9728, "wrapS": 33071, "wrapT": 33648}]
    j["textures"] = [{"source": 0, "sampler": 0}, {"source": 1, "sampler": 1}, {"source": 2},
                     {"source": 3, "sampler": 1}, {"source": 4, "sampler": 0}]
    j["materials"] = [
        {   # 0: textured, alpha mask, transmission, sheen, ior, double-sided
            "pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "baseColorFactor": [1, 0.8, 0.6, 1],
                                     "metallicRoughnessTexture": {"index": 2, "texCoord": 1},
                                     "metallicFactor": 0.25},
            "normalTexture": {"index": 1, "texCoord": 1}, "alphaMode": "MASK", "alphaCutoff": 0.4,
            "doubleSided": True,
            "extensions": {"KHR_materials_ior": {"ior": 1.33},
                           "KHR_materials_transmission": {"transmissionFactor": 0.5,
                                                          "transmissionTexture": {"index": 3}},
                           "KHR_materials_sheen": {"sheenColorFactor": [0.9, 0.9, 1.0],
                                                   "sheenRoughnessFactor": 0.3,
                                                   "sheenColorTexture": {"index": 0, "texCoord": 1}}}},
        {   # 1: factors only, blend
            "pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.5, 0.9, 0.7], "metallicFactor": 0.0,
                                     "roughnessFactor": 0.45}, "alphaMode": "BLEND"},
        {   # 2: textured emitter with a strength
            "emissiveFactor": [1.0, 0.5, 0.25], "emissiveTexture": {"index": 4},
            "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0}}},
        {   # 3: untextured emitter
            "emissiveFactor": [2.0, 2.0, 2.0]},
    ]
    j["meshes"] = [
        {"name": "quad", "primitives": [{"attributes": quad, "indices": quad_idx16, "material": 0}]},
        {"name": "box", "primitives": [{"attributes": box, "indices": box_idx, "material": 1}]},
        {"name": "emitter", "primitives": [{"attributes": {"POSITION": quad["POSITION"],
                                                           "TEXCOORD_0": quad["TEXCOORD_0"]},
                                            "indices": quad_idx8, "material": 2}]},
        {"name": "tris", "primitives": [{"attributes": tri, "material": 3},
                                        {"attributes": tri, "mode": 1}]},       # a line list: skipped
    ]
    rot = [0.0, math.sin(0.3), 0.0, math.cos(0.3)]
    j["nodes"] = [
        {"name": "root", "translation": [0.5, -1.0, -4.0], "rotation": rot, "scale": [1.5, 1.5, 1.5],
         "children": [1, 2]},
        {"name": "quad", "mesh": 0, "translation": [0.0, 0.5, 0.0]},
        {"name": "boxes", "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -1.5, 0.2, 0.3, 1], "mesh": 1,
         "children": [3]},
        {"name": "emitter a", "mesh": 2, "translation": [0.0, 2.0, 0.0], "scale": [0.5, 0.5, 0.5]},
        {"name": "emitter b", "mesh": 2, "translation": [2.0, 2.0, -1.0], "scale": [0.25, 0.5, 1.0]},
        {"name": "tris", "mesh": 3, "translation": [-2.0, -0.5, 1.0]},
    ]
    j["scenes"] = [{"nodes": [0, 4, 5]}]
    return j, buf, ims


def write_gltf(path):
    """.gltf + .bin, the images as PNG files next to it."""
    import os
    j, buf, ims = document()
    d = os.path.dirname(path)
    names = ("base", "normal", "metrough", "grey", "emission")
    for i, k in enumerate(names):
        fn = "tex_%s.png" % k
        with open(os.path.join(d, fn), "wb") as f:
            f.write(_png(ims[k]))
        j["images"][i] = {"uri": fn}
    j["buffers"] = [{"byteLength": len(buf.data), "uri": os.path.basename(path) + ".bin"}]
    j["bufferViews"] = buf.views
    j["accessors"] = buf.accessors
    with open(path + ".bin", "wb") as f:
        f.write(bytes(buf.data))
    with open(path, "w") as f:
        json.dump(j, f)


def write_glb(path):
    j, buf, _ = document()
    j["buffers"] = [{"byteLength": len(buf.data)}]
    j["bufferViews"] = buf.views
    j["accessors"] = buf.accessors
    js = json.dumps(j).encode()
    js += b" " * ((-len(js)) % 4)
    bb = bytes(buf.data) + b"\0" * ((-len(buf.data)) % 4)
    total = 12 + 8 + len(js) + 8 + len(bb)
    with open(path, "wb") as f:
        f.write(b"glTF" + struct.pack("<II", 2, total))
        f.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
        f.write(struct.pack("<II", len(bb), 0x004E4942) + bb)
