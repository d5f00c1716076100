"""glTF 2.0 ingestion: meshes, materials, textures and emitters in the
renderer's input layout.

Follows the reference's loader (src/core/scene.cpp:15-641) and the material
and attribute packing of src/core/rendering_interface.cpp:47-307, with the
scene graph flattened into one world-space mesh instead of an instance
hierarchy:

  * .gltf (JSON with external or base64 "data:" buffers) and .glb; images
    from files or buffer views, decoded to 8-bit (one channel stays R8,
    every other layout becomes RGBA8), samplers mapped to nearest / linear
    and wrap / clamp / mirror; a texture used as a base colour is sampled as
    sRGB, every other texture as linear (the reference marks only base
    colour);
  * positions, normals and tangents change coordinates (x, y, z) ->
    (-x, z, y); node transforms (matrix, or translation-rotation-scale)
    compose from the root; texture coordinates in float, normalised u8 or
    normalised u16; indices u8 / u16 / u32 (non-indexed primitives get
    0..n-1); only triangle lists are kept;
  * each primitive becomes one material (ComplexSurface): base-colour,
    metal-roughness, normal, transmission (KHR_materials_transmission),
    sheen (KHR_materials_sheen, present when its colour factor is not
    black), ior (KHR_materials_ior), emission (emissiveFactor x
    KHR_materials_emissive_strength, with its texture), alpha mode
    (MASK -> cutoff, BLEND -> blend), double-sidedness; a factor counts as
    present when it differs from its glTF default, as the reference decides;
    the texture-coordinate set of each slot is recorded;
  * vertex attributes: a tangent frame quaternion (normal and tangent
    present; the sign of w carries the bitangent handedness) or an
    octahedral normal, then TEX_COORD_1 and TEX_COORD_2, in a 32-byte
    record per vertex;
  * emitters: the triangles of a primitive with a non-zero emissive factor,
    with the flux the reference computes in model space
    (rbh_emissive_flux: factor mean x rasterised emission-texture mean x
    area x pi) times |det| of the instance's 3x3 transform
    (scene.cpp:61, rendering_interface.cpp:1148); zero-flux triangles are
    not lights.
"""
import base64
import io
import json
import math
import os
import struct

import numpy as np

from . import _lib as L
from .scenes import (AT_FRAME, AT_NORMAL, AT_TEX1, AT_TEX2, BXDF_COMPLEX, FC_BASE_COLOR, FC_BLEND, FC_CUTOFF,
                     FC_EMISSION, FC_METALNESS, FC_ROUGHNESS, FC_SHEEN_COLOR, FC_SHEEN_ROUGH, FC_TRANSMISSION,
                     MF_DOUBLE_SIDED, MF_SHEEN, NO_LIGHT, TX_BASE_COLOR, TX_EMISSION, TX_MET_ROUGH, TX_NORMAL,
                     TX_SHEEN_COLOR, TX_SHEEN_ROUGH, TX_TRANSMISSION, Scene, camera_basis, emissive_records,
                     load_tables, oct_encode)

ATTRIB_STRIDE = 32          # frame quaternion / octahedral normal at 0, TEX_COORD_1 at 16, TEX_COORD_2 at 24
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_WRAP = {10497: 0, 33071: 1, 33648: 2}
_SLOTS = ("base", "normal", "metrough", "transmission", "sheencolor", "sheenrough", "emission")


class GltfError(ValueError):
    pass


# ------------------------------------------------------------------ files
def _read_container(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"glTF":
        version, length = struct.unpack_from("<II", data, 4)
        if version != 2:
            raise GltfError("glTF binary version %d is not supported" % version)
        pos, doc, binchunk = 12, None, None
        while pos + 8 <= min(length, len(data)):
            clen, ctype = struct.unpack_from("<II", data, pos)
            chunk = data[pos + 8:pos + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                binchunk = chunk
            pos += 8 + clen
        if doc is None:
            raise GltfError("glb without a JSON chunk")
        return doc, binchunk
    return json.loads(data.decode("utf-8")), None


def _load_uri(base_dir, uri):
    if uri.startswith("data:"):
        head, _, payload = uri.partition(",")
        if not head.endswith(";base64"):
            raise GltfError("only base64 data URIs are supported")
        return base64.b64decode(payload)
    from urllib.parse import unquote
    with open(os.path.join(base_dir, unquote(uri)), "rb") as f:
        return f.read()


class _Doc:
    def __init__(self, path):
        self.path = path
        self.dir = os.path.dirname(os.path.abspath(path))
        self.j, binchunk = _read_container(path)
        self.buffers = []
        for i, b in enumerate(self.j.get("buffers", [])):
            if "uri" in b:
                self.buffers.append(_load_uri(self.dir, b["uri"]))
            elif i == 0 and binchunk is not None:
                self.buffers.append(binchunk)
            else:
                raise GltfError("buffer %d has no data" % i)

    def view_bytes(self, vi):
        v = self.j["bufferViews"][vi]
        buf = self.buffers[v["buffer"]]
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v.get("byteStride", 0)

    def accessor(self, ai):
        a = self.j["accessors"][ai]
        if "sparse" in a:
            raise GltfError("sparse accessors are not supported")
        dt = np.dtype(_COMPONENT[a["componentType"]])
        nc = _NCOMP[a["type"]]
        count = a["count"]
        if "bufferView" not in a:
            return np.zeros((count, nc), dt)
        raw, stride = self.view_bytes(a["bufferView"])
        off = a.get("byteOffset", 0)
        elem = dt.itemsize * nc
        stride = stride or elem
        if count and off + (count - 1) * stride + elem > len(raw):
            raise GltfError("accessor %d reads past its buffer view" % ai)
        arr = np.ndarray((count, nc), dt, buffer=raw, offset=off, strides=(stride, dt.itemsize))
        return np.array(arr)


# ----------------------------------------------------------------- images
def _decode_image(data):
    """8-bit texels as stb_image's stbi_load(..., 1 channel stays 1, else 4)
    gives them: (h, w) uint8 or (h, w, 4) uint8."""
    from PIL import Image
    im = Image.open(io.BytesIO(data))
    mode = im.mode
    if mode in ("I;16", "I;16B", "I;16L", "I"):
        a = np.asarray(im).astype(np.uint32)
        return (a >> 8).astype(np.uint8) if a.max() > 255 or mode.startswith("I;16") else a.astype(np.uint8)
    if mode == "L":
        return np.asarray(im, np.uint8).copy()
    if mode == "1":
        return np.asarray(im.convert("L"), np.uint8).copy()
    return np.asarray(im.convert("RGBA"), np.uint8).copy()


def _quat_from_frames(t, b, n):
    """glm::quat_cast of the matrices with columns (t, b, n), rows of (k, 3)."""
    m00, m10, m20 = t[:, 0], t[:, 1], t[:, 2]
    m01, m11, m21 = b[:, 0], b[:, 1], b[:, 2]
    m02, m12, m22 = n[:, 0], n[:, 1], n[:, 2]
    four = np.stack([m00 + m11 + m22, m00 - m11 - m22, m11 - m00 - m22, m22 - m00 - m11], 1)
    big = four.argmax(axis=1)
    v = np.sqrt(four[np.arange(len(t)), big] + 1.0) * 0.5
    mult = 0.25 / v
    q = np.zeros((len(t), 4))
    sel = big == 0
    q[sel] = np.stack([(m21 - m12)[sel] * mult[sel], (m02 - m20)[sel] * mult[sel], (m10 - m01)[sel] * mult[sel],
                       v[sel]], 1)
    sel = big == 1
    q[sel] = np.stack([v[sel], (m10 + m01)[sel] * mult[sel], (m02 + m20)[sel] * mult[sel],
                       (m21 - m12)[sel] * mult[sel]], 1)
    sel = big == 2
    q[sel] = np.stack([(m10 + m01)[sel] * mult[sel], v[sel], (m21 + m12)[sel] * mult[sel],
                       (m02 - m20)[sel] * mult[sel]], 1)
    sel = big == 3
    q[sel] = np.stack([(m02 + m20)[sel] * mult[sel], (m21 + m12)[sel] * mult[sel], v[sel],
                       (m10 - m01)[sel] * mult[sel]], 1)
    return q


def _node_matrix(node):
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T     # column-major
    t = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * s[None, :]
    M[:3, 3] = t
    return M


# The loader's change of coordinates, (x, y, z) -> (-x, z, y)
C3 = np.array([[-1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])


class _Prim:
    pass


def load(path, transform=None):
    """Reads a .gltf / .glb file into flattened arrays (a dict): verts
    (world space), indices, attribs, tri_material, materials, textures,
    texels, the emitter records `lights` (world space) with their `flux`,
    and `bounds`. transform: an optional 4x4 world-from-model matrix applied
    after the change of coordinates."""
    doc = _Doc(path)
    j = doc.j
    W = np.eye(4) if transform is None else np.asarray(transform, np.float64)

    # images and textures
    images = []
    for im in j.get("images", []):
        if "uri" in im:
            if im["uri"].startswith("data:"):
                raise GltfError("base64 images are not supported")
            images.append(_decode_image(_load_uri(doc.dir, im["uri"])))
        else:
            raw, _ = doc.view_bytes(im["bufferView"])
            images.append(_decode_image(raw))
    textures = []
    for t in j.get("textures", []):
        filt, ax, ay = 1, 0, 0
        if "sampler" in t:
            sm = j["samplers"][t["sampler"]]
            filt = 0 if sm.get("magFilter") == 9728 else 1
            ax = _WRAP.get(sm.get("wrapS", 10497), 0)
            ay = _WRAP.get(sm.get("wrapT", 10497), 0)
        textures.append({"image": t.get("source", 0), "filter": filt, "addr": (ax, ay), "srgb": 0})

    def tex_ref(info):
        if not info or "index" not in info:
            return -1, 0
        return int(info["index"]), int(info.get("texCoord", 0))

    # primitives (one material each)
    prims = []
    for mi, mesh in enumerate(j.get("meshes", [])):
        plist = []
        for pi, prim in enumerate(mesh.get("primitives", [])):
            if prim.get("mode", 4) != 4:
                continue
            attrs = prim["attributes"]
            p = _Prim()
            p.pos = doc.accessor(attrs["POSITION"]).astype(np.float32) @ C3.T.astype(np.float32)
            nv = len(p.pos)
            p.nrm = doc.accessor(attrs["NORMAL"]).astype(np.float32) @ C3.T.astype(np.float32) \
                if "NORMAL" in attrs else None
            if "TANGENT" in attrs:
                tg = doc.accessor(attrs["TANGENT"]).astype(np.float32)
                p.tan = np.concatenate([tg[:, :3] @ C3.T.astype(np.float32), tg[:, 3:4]], 1)
            else:
                p.tan = None
            p.uv = []
            for s in (0, 1):
                key = "TEXCOORD_%d" % s
                if key not in attrs:
                    break
                a = doc.accessor(attrs[key])
                acc = j["accessors"][attrs[key]]
                if a.dtype == np.uint8:
                    a = a.astype(np.float32) * np.float32(1.0 / 255.0)
                elif a.dtype == np.uint16:
                    a = a.astype(np.float32) * np.float32(1.0 / 65535.0)
                elif a.dtype != np.float32:
                    raise GltfError("unsupported texture-coordinate component type %d" % acc["componentType"])
                p.uv.append(a[:, :2].astype(np.float32))
            if "indices" in prim:
                idx = doc.accessor(prim["indices"])[:, 0]
                p.idx_u16 = idx.dtype != np.uint32
                p.idx = idx.astype(np.uint32).reshape(-1, 3)
            else:
                p.idx_u16 = nv <= 65536
                p.idx = np.arange(nv - nv % 3, dtype=np.uint32).reshape(-1, 3)
            if len(p.idx) and p.idx.max() >= nv:
                raise GltfError("index out of range in mesh %d primitive %d" % (mi, pi))
            p.mat = _material(j, prim.get("material"), textures, tex_ref, p)
            p.emissive = p.mat["_emissive"]
            p.flux = None
            plist.append(p)
        prims.append(plist)

    # model-space emitter flux (before instancing)
    for plist in prims:
        for p in plist:
            if not p.emissive:
                continue
            m = p.mat
            tri = p.pos[p.idx]                                      # (n, 3, 3)
            ef = np.asarray(m["emissive"], np.float32)
            em = m["_emTex"]
            if em >= 0:
                uvset = p.uv[m["_emSet"]] if m["_emSet"] < len(p.uv) else np.zeros((len(p.pos), 2), np.float32)
                tuv = np.ascontiguousarray(uvset[p.idx].reshape(-1, 6), np.float32)
                t = textures[em]
                img = images[t["image"]]
                if img.ndim != 3:
                    # the reference's flux pass reads four bytes per texel; a
                    # one-channel emission image is widened to grey RGBA here
                    img = np.repeat(img[..., None], 4, axis=2)
                    img[..., 3] = 255
                rgba = np.ascontiguousarray(img, np.uint8)
                p.uv_em = tuv
                p.flux = _flux(tri, tuv, ef, rgba, t["addr"])
            else:
                p.uv_em = np.zeros((len(p.idx), 6), np.float32)
                p.flux = _flux(tri, None, ef, None, (0, 0))

    # materials and textures in the renderer's layout
    mats = []
    for plist in prims:
        for p in plist:
            p.mat_index = len(mats)
            mats.append(_device_material(p))
    descs = np.zeros(len(textures), L.TEXTURE_DTYPE)
    blobs, off = [], 0
    for i, t in enumerate(textures):
        img = images[t["image"]]
        one = img.ndim == 2
        descs[i]["width"], descs[i]["height"], descs[i]["depth"] = img.shape[1], img.shape[0], 1
        descs[i]["format"] = 1 if one else 0
        descs[i]["srgb"] = t["srgb"]
        descs[i]["addr"] = (t["addr"][0], t["addr"][1], 1)
        descs[i]["filter"] = t["filter"]
        descs[i]["offset"] = off
        raw = np.ascontiguousarray(img).reshape(-1)
        pad = (-len(raw)) % 256
        blobs += [raw, np.zeros(pad, np.uint8)]
        off += len(raw) + pad

    # scene graph
    out_v, out_i, out_a, out_m = [], [], [], []
    light_tris, light_uvs, light_flux = [], [], []
    count = {"verts": 0, "tris": 0}

    def visit(ni, parent):
        node = j["nodes"][ni]
        # the node's matrix conjugated by the change of coordinates
        nm = _node_matrix(node)
        local = np.eye(4)
        local[:3, :3] = C3 @ nm[:3, :3] @ C3
        local[:3, 3] = C3 @ nm[:3, 3]
        M = parent @ local
        if "mesh" in node:
            world = W @ M
            A = world[:3, :3]
            det = float(np.linalg.det(A))
            nmat = np.linalg.inv(A).T
            for p in prims[node["mesh"]]:
                pos = p.pos.astype(np.float64) @ A.T + world[:3, 3]
                rec = np.zeros((len(pos), ATTRIB_STRIDE // 4), np.float32)
                if p.nrm is not None:
                    n = p.nrm.astype(np.float64) @ nmat.T
                    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
                    if p.tan is not None:
                        t = p.tan[:, :3].astype(np.float64) @ A.T
                        t /= np.maximum(np.linalg.norm(t, axis=1, keepdims=True), 1e-30)
                        w = p.tan[:, 3].astype(np.float64) * (1.0 if det >= 0 else -1.0)
                        q = _quat_from_frames(t, np.cross(n, t), n)
                        # the sign of w carries the handedness (rendering_interface.cpp:248-268)
                        q[q[:, 3] * w < 0.0] *= -1.0
                        zero = q[:, 3] == 0.0
                        q[zero, 3] = np.where(w[zero] > 0.0, 0.0, -0.0)
                        rec[:, 0:4] = q
                    else:
                        rec[:, 0:2] = oct_encode(n)
                for s_, uv in enumerate(p.uv[:2]):
                    rec[:, 4 + 2 * s_:6 + 2 * s_] = uv
                out_v.append(pos)
                out_i.append(p.idx + count["verts"])
                out_a.append(rec)
                out_m.append(np.full(len(p.idx), p.mat_index, np.uint32))
                if p.emissive:
                    ids = np.nonzero(p.flux > 0.0)[0]
                    light_tris.append(count["tris"] + ids)
                    light_uvs.append(p.uv_em[ids])
                    light_flux.append(p.flux[ids] * np.float32(abs(det)))
                count["verts"] += len(pos)
                count["tris"] += len(p.idx)
        for c in node.get("children", []):
            visit(c, M)

    roots = j["scenes"][j.get("scene", 0)]["nodes"] if j.get("scenes") else list(range(len(j.get("nodes", []))))
    for r in roots:
        visit(r, np.eye(4))
    if not out_v:
        raise GltfError("the scene holds no triangle meshes")
    verts = np.concatenate(out_v).astype(np.float32)
    cat = (lambda xs, dt, shape: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt))
    return {"verts": verts, "indices": np.concatenate(out_i).astype(np.uint32),
            "attribs": np.concatenate(out_a).view(np.uint8).reshape(-1),
            "tri_material": np.concatenate(out_m), "materials": np.array(mats, L.MATERIAL_DTYPE),
            "textures": descs, "texels": np.concatenate(blobs) if blobs else np.zeros(0, np.uint8),
            "light_tris": cat(light_tris, np.uint32, (0,)), "light_uvs": cat(light_uvs, np.float32, (0, 6)),
            "flux": cat(light_flux, np.float32, (0,)), "bounds": (verts.min(axis=0), verts.max(axis=0))}


def _flux(tri, uvs, factor, rgba, addr):
    import ctypes
    n = len(tri)
    v = np.ascontiguousarray(tri.reshape(n, 9), np.float32)
    f = np.ascontiguousarray(factor, np.float32)
    out = np.zeros(n, np.float32)
    if rgba is None:
        rc = L.lib().rbh_emissive_flux(L.ptr(v), None, n, L.ptr(f), None, 0, 0, 0, 0, L.ptr(out))
    else:
        rc = L.lib().rbh_emissive_flux(L.ptr(v), L.ptr(uvs), n, L.ptr(f), L.ptr(rgba), rgba.shape[1], rgba.shape[0],
                                       int(addr[0]), int(addr[1]), L.ptr(out))
    if rc != 0:
        raise GltfError("rbh_emissive_flux failed (%d)" % rc)
    return out


def _material(j, mi, textures, tex_ref, prim):
    """MaterialDescriptor of scene.cpp:222-358 for one primitive."""
    d = {"base": -1, "normal": -1, "metrough": -1, "transmission": -1, "sheencolor": -1, "sheenrough": -1,
         "emission": -1, "sets": {}, "bcf": None, "met": None, "rough": None, "tr": None, "shc": None, "shr": None,
         "sheen": False, "alpha": 0, "cutoff": 0.5, "ior": 1.5, "double": False, "emissive": (0.0, 0.0, 0.0),
         "_emissive": False, "_emTex": -1, "_emSet": 0}
    if mi is None:
        return d
    m = j["materials"][mi]
    ext = m.get("extensions", {})
    d["double"] = bool(m.get("doubleSided", False))
    d["ior"] = float(ext.get("KHR_materials_ior", {}).get("ior", 1.5))
    pbr = m.get("pbrMetallicRoughness")
    if pbr is not None:
        ti, ts = tex_ref(pbr.get("baseColorTexture"))
        if ti >= 0:
            textures[ti]["srgb"] = 1
            d["base"], d["sets"]["base"] = ti, ts
        bcf = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        if any(float(x) != 1.0 for x in bcf):
            d["bcf"] = [float(x) for x in bcf]
        mode = m.get("alphaMode", "OPAQUE")
        d["alpha"] = {"MASK": 1, "BLEND": 2}.get(mode, 0)
        d["cutoff"] = float(m.get("alphaCutoff", 0.5))
        mf, rf = float(pbr.get("metallicFactor", 1.0)), float(pbr.get("roughnessFactor", 1.0))
        if mf != 1.0:
            d["met"] = mf
        if rf != 1.0:
            d["rough"] = rf
        ti, ts = tex_ref(pbr.get("metallicRoughnessTexture"))
        if ti >= 0:
            d["metrough"], d["sets"]["metrough"] = ti, ts
    tr = ext.get("KHR_materials_transmission")
    if tr is not None:
        ti, ts = tex_ref(tr.get("transmissionTexture"))
        if ti >= 0:
            d["transmission"], d["sets"]["transmission"] = ti, ts
        tf = float(tr.get("transmissionFactor", 0.0))
        if tf != 0.0:
            d["tr"] = tf
    ti, ts = tex_ref(m.get("normalTexture"))
    if ti >= 0:
        d["normal"], d["sets"]["normal"] = ti, ts
    sh = ext.get("KHR_materials_sheen")
    if sh is not None:
        scf = [float(x) for x in sh.get("sheenColorFactor", [0.0, 0.0, 0.0])]
        if not all(x == 0.0 for x in scf):
            d["sheen"] = True
            if any(x != 1.0 for x in scf):
                d["shc"] = scf
            d["shr"] = float(sh.get("sheenRoughnessFactor", 0.0))
            ti, ts = tex_ref(sh.get("sheenColorTexture"))
            if ti >= 0:
                d["sheencolor"], d["sets"]["sheencolor"] = ti, ts
            ti, ts = tex_ref(sh.get("sheenRoughnessTexture"))
            if ti >= 0:
                d["sheenrough"], d["sets"]["sheenrough"] = ti, ts
    ef = [float(x) for x in m.get("emissiveFactor", [0.0, 0.0, 0.0])]
    if not all(x == 0.0 for x in ef):
        strength = float(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0)) \
            if "KHR_materials_emissive_strength" in ext else None
        if strength is not None:
            ef = [float(np.float32(x) * np.float32(strength)) for x in ef]
        d["emissive"] = tuple(ef)
        d["_emissive"] = True
        ti, ts = tex_ref(m.get("emissiveTexture"))
        if ti >= 0:
            d["emission"], d["sets"]["emission"] = ti, ts
            d["_emTex"], d["_emSet"] = ti, ts
    return d


def _device_material(p):
    """MaterialData of rendering_interface.cpp:72-190 plus the attribute
    layout of :200-233 (fixed 32-byte records here)."""
    d = p.mat
    m = np.zeros(1, L.MATERIAL_DTYPE)[0]
    m["bxdf"] = BXDF_COMPLEX
    m["flags"] = (MF_DOUBLE_SIDED if d["double"] else 0) | (MF_SHEEN if d["sheen"] else 0)
    attrs = 0
    if p.nrm is not None and p.tan is not None:
        attrs |= AT_FRAME
    elif p.nrm is not None:
        attrs |= AT_NORMAL
    if len(p.uv) >= 1:
        attrs |= AT_TEX1
    if len(p.uv) >= 2:
        attrs |= AT_TEX2
    m["attributes"] = attrs
    m["attributeStride"] = ATTRIB_STRIDE
    m["normalOffset"], m["frameOffset"], m["texCoord1Offset"], m["texCoord2Offset"] = 0, 0, 16, 24
    m["mfRoughness"] = 0.0
    m["ior"] = d["ior"]
    m["alphaCutoff"] = d["cutoff"]
    m["tex"][:] = -1
    bits = {"base": TX_BASE_COLOR, "normal": TX_NORMAL, "metrough": TX_MET_ROUGH, "transmission": TX_TRANSMISSION,
            "sheencolor": TX_SHEEN_COLOR, "sheenrough": TX_SHEEN_ROUGH, "emission": TX_EMISSION}
    tex, sets = 0, 0
    for slot, name in enumerate(_SLOTS):
        ti = d[name]
        if ti < 0:
            continue
        if name in ("sheencolor", "sheenrough") and not d["sheen"]:
            continue
        tex |= bits[name]
        m["tex"][slot] = ti
        if d["sets"].get(name, 0) == 1:
            sets |= 1 << slot
    m["textures"] = tex
    m["texCoordSets"] = sets
    fac = 0
    if d["bcf"] is not None:
        fac |= FC_BASE_COLOR
        m["baseColor"][:] = d["bcf"]
    else:
        m["baseColor"][:] = (1.0, 1.0, 1.0, 1.0)
    if d["alpha"] == 1:
        fac |= FC_CUTOFF
    elif d["alpha"] == 2:
        fac |= FC_BLEND
    if d["met"] is not None:
        fac |= FC_METALNESS
        m["metalness"] = d["met"]
    if d["rough"] is not None:
        fac |= FC_ROUGHNESS
        m["roughness"] = d["rough"]
    if d["tr"] is not None:
        fac |= FC_TRANSMISSION
        m["transmission"] = d["tr"]
    if d["sheen"]:
        if d["shc"] is not None:
            fac |= FC_SHEEN_COLOR
            m["sheenColor"][:] = d["shc"]
        fac |= FC_SHEEN_ROUGH
        m["sheenRoughness"] = d["shr"]
    if d["_emissive"]:
        fac |= FC_EMISSION
        m["emissive"][:] = d["emissive"]
    m["factors"] = fac
    return m


def scene(path, camera=None, env=None, width=512, height=512, spp=16, depth=8, transform=None, fov_deg=45.0,
          tables=None):
    """A Scene (ratbite_amd.scenes.Scene) from a glTF file. camera: dict with
    pos / target / up (world space, after the change of coordinates), or None
    to look at the scene's bounding box from its +z side. env: a (h, w, 3|4)
    float image or the path of a Radiance .hdr file (width and height
    multiples of 4, as the reference requires), or None."""
    g = load(path, transform=transform)
    lo, hi = g["bounds"]
    if camera is None:
        c = 0.5 * (lo + hi)
        r = float(np.linalg.norm(hi - lo)) * 0.5 or 1.0
        d = r / math.tan(math.radians(fov_deg) * 0.5) * 1.1
        camera = {"pos": c + np.array([0.0, 0.0, d]), "target": c, "up": (0.0, 1.0, 0.0)}
    cam = camera_basis(camera["pos"], camera["target"], camera.get("up", (0.0, 1.0, 0.0)))
    cam["fov"] = math.radians(fov_deg)
    if isinstance(env, str):
        from .hdr import read_hdr
        env = read_hdr(env)
    if env is not None:
        h, w = env.shape[:2]
        if h % 4 or w % 4:
            raise GltfError("environment map dimensions must be multiples of 4")
    verts = (g["verts"].astype(np.float64) - cam["pos"]).astype(np.float32)
    tri_light = np.full(len(g["indices"]), NO_LIGHT, np.uint32)
    tri_light[g["light_tris"]] = np.arange(len(g["light_tris"]), dtype=np.uint32)
    lights = emissive_records(verts, g["indices"], g["light_tris"], g["tri_material"],
                              np.ascontiguousarray(g["light_uvs"], np.float32))
    t = tables or load_tables()
    return Scene(verts=verts, indices=g["indices"], attribs=g["attribs"], tri_material=g["tri_material"],
                 tri_light=tri_light, materials=g["materials"], spectra=np.ones((1, 471), np.float32),
                 textures=g["textures"], texels=g["texels"], lights=lights, flux=g["flux"], env=env, camera=cam,
                 width=width, height=height, depth=depth, spp=spp, tables=t, name=os.path.basename(path))
