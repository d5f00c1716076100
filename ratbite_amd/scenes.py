"""Seeded synthetic workloads for the benchmark configurations.

A Scene holds flattened, camera-relative world-space geometry plus every
table the renderer needs, as numpy arrays in the byte layouts of
include/ratbite_hip.h. The same arrays feed the HIP library and the CPU
oracle.

  cornell(...)   256x256 Cornell box: 5 wall quads, 8 UV spheres (32x16),
                 one emissive quad with a 4000 K blackbody spectrum, glossy-
                 diffuse OpenPBR materials, no env map, depth 8, 16 spp.
  spheres(...)   1920x1080 field of displaced level-5 icospheres, textured
                 emissive triangles feeding the light tree, a log-normal
                 env map with a sun, a 40/25/20/15 mix of glossy-diffuse /
                 conductor (Cauchy eta, k) / translucent / sheen materials,
                 depth 16, 64 spp.
"""
import math
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import EMISSIVE_TRI_DTYPE, MATERIAL_DTYPE, SHAPE_LIGHT_DTYPE, TEXTURE_DTYPE, LaunchParams
from .envmaps import lognormal_envmap

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = os.path.join(HERE, "data", "tables.npz")

# MaterialData bitfields (src/core/material.h:19-57)
AT_NORMAL, AT_FRAME, AT_COLOR, AT_TEX1, AT_TEX2 = 1, 2, 4, 8, 16
TX_BASE_COLOR, TX_NORMAL, TX_MET_ROUGH, TX_TRANSMISSION, TX_SHEEN_COLOR, TX_SHEEN_ROUGH, TX_EMISSION = 1, 2, 4, 8, 16, 32, 64
FC_BASE_COLOR, FC_METALNESS, FC_ROUGHNESS, FC_TRANSMISSION, FC_CUTOFF, FC_BLEND = 1, 2, 4, 8, 16, 32
FC_SHEEN_COLOR, FC_SHEEN_ROUGH, FC_EMISSION = 64, 128, 256
MF_DOUBLE_SIDED, MF_SHEEN, MF_SPECTRAL_EMISSION = 1, 2, 4
BXDF_CONDUCTOR, BXDF_DIELECTRIC, BXDF_COMPLEX = 0, 1, 2
NO_LIGHT = 0xFFFFFFFF
ATTRIB_STRIDE = 16   # oct normal (2 f32) + uv0 (2 f32)


def load_tables():
    t = np.load(TABLES)
    return {k: t[k] for k in t.files}


@dataclass
class Scene:
    verts: np.ndarray
    indices: np.ndarray
    attribs: np.ndarray
    tri_material: np.ndarray
    tri_light: np.ndarray
    materials: np.ndarray
    spectra: np.ndarray
    textures: np.ndarray
    texels: np.ndarray
    lights: np.ndarray
    flux: np.ndarray
    env: np.ndarray = None
    camera: dict = field(default_factory=dict)
    width: int = 256
    height: int = 256
    depth: int = 8
    spp: int = 16
    tables: dict = field(default_factory=dict)
    name: str = ""
    shapes: np.ndarray = None       # disk and sphere lights (SHAPE_LIGHT_DTYPE), camera-relative
    shape_flux: np.ndarray = None

    @property
    def nshapes(self):
        return 0 if self.shapes is None else len(self.shapes)

    def launch_params(self, offset=0, count=None, width=None, height=None, depth=None, flags=0):
        p = LaunchParams()
        p.width = width or self.width
        p.height = height or self.height
        d = depth or self.depth
        p.maxPathDepth = d
        p.maxReflectedPathDepth = d
        p.maxTransmittedPathDepth = d
        p.sampleOffset = offset
        p.sampleCount = self.spp if count is None else count
        for k, name in (("u", "camU"), ("v", "camV"), ("w", "camW")):
            getattr(p, name)[:] = [float(x) for x in self.camera[k]]
        fov = self.camera.get("fov", math.radians(70.0))
        p.perspectiveScaleW = float(np.float32(math.tan(fov * 0.5) * (p.width / p.height)))
        p.perspectiveScaleH = float(np.float32(math.tan(fov * 0.5)))
        p.depthOfField = 0
        p.aperture = 0.0
        p.focusDistance = 1.0
        p.flags = flags
        return p

    @property
    def ntris(self):
        return len(self.indices)


# ----------------------------------------------------------------- helpers
def oct_encode(n):
    """src/core/math_util.h:147-162 on arrays (attribute normals)."""
    n = np.asarray(n, np.float32)
    s = np.abs(n).sum(axis=1, keepdims=True)
    p = n[:, :2] / s
    neg = n[:, 2] <= 0
    sx = np.where(p[:, 0] >= 0, 1.0, -1.0)
    sy = np.where(p[:, 1] >= 0, 1.0, -1.0)
    q = p.copy()
    q[neg, 0] = (1.0 - np.abs(p[neg, 1])) * sx[neg]
    q[neg, 1] = (1.0 - np.abs(p[neg, 0])) * sy[neg]
    return q.astype(np.float32)


def pack_attribs(normals, uvs):
    n = len(normals)
    a = np.zeros((n, 4), np.float32)
    a[:, 0:2] = oct_encode(normals)
    a[:, 2:4] = uvs
    return a.view(np.uint8).reshape(-1)


def camera_basis(pos, target, up):
    """src/core/camera.h:33-38"""
    v = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    v /= np.linalg.norm(v)
    u = np.cross(v, np.asarray(up, np.float64))
    u /= np.linalg.norm(u)
    w = np.cross(u, v)
    return {"pos": np.asarray(pos, np.float64), "u": u.astype(np.float32), "v": v.astype(np.float32),
            "w": w.astype(np.float32)}


def material(bxdf=BXDF_COMPLEX, base=(1, 1, 1), metal=0.0, rough=0.5, ior=1.5, transmission=None,
             sheen=None, emissive=None, emission_tex=-1, spectral_emission=None, eta_k=None):
    m = np.zeros(1, MATERIAL_DTYPE)[0]
    m["bxdf"] = bxdf
    m["attributes"] = AT_NORMAL | AT_TEX1
    m["attributeStride"] = ATTRIB_STRIDE
    m["normalOffset"] = 0
    m["texCoord1Offset"] = 8
    m["ior"] = ior
    m["tex"][:] = -1
    m["baseColor"][:] = (base[0], base[1], base[2], 1.0)
    m["metalness"] = metal
    m["roughness"] = rough
    m["mfRoughness"] = rough
    m["factors"] = FC_BASE_COLOR | FC_METALNESS | FC_ROUGHNESS
    if transmission is not None:
        m["transmission"] = transmission
        m["factors"] |= FC_TRANSMISSION
    if sheen is not None:
        m["flags"] |= MF_SHEEN
        m["sheenColor"][:] = sheen[0]
        m["sheenRoughness"] = sheen[1]
        m["factors"] |= FC_SHEEN_COLOR | FC_SHEEN_ROUGH
    if emissive is not None:
        m["emissive"][:] = emissive
        m["factors"] |= FC_EMISSION
    if emission_tex >= 0:
        m["textures"] |= TX_EMISSION
        m["tex"][6] = emission_tex
    if spectral_emission is not None:
        m["flags"] |= MF_SPECTRAL_EMISSION
        m["emissionSpectrum"] = spectral_emission[0]
        m["emissive"][:] = (spectral_emission[1], 0, 0)
    if eta_k is not None:
        m["iorSpectrum"], m["acSpectrum"] = eta_k
    return m


def emissive_records(verts, indices, tri_ids, tri_material, uvs_per_tri):
    rec = np.zeros(len(tri_ids), EMISSIVE_TRI_DTYPE)
    v = verts[indices[tri_ids]]                  # (n, 3, 3)
    rec["vertices"] = v.reshape(-1, 9)
    rec["uvs"] = uvs_per_tri.reshape(-1, 6)
    rec["materialIndex"] = tri_material[tri_ids]
    rec["primitiveDataIndex"] = tri_ids
    return rec


def triangle_flux(rec, materials, tex_avg):
    """Emitted flux of each light as src/core/scene.cpp:519-607 computes it:
    mean emissive factor x mean textured emission over the UV footprint x
    area x pi (spectral emitters: luminance-normalised spectrum x scale)."""
    v = rec["vertices"].reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    m = materials[rec["materialIndex"]]
    f = m["emissive"].mean(axis=1).astype(np.float64)
    spectral = (m["flags"] & MF_SPECTRAL_EMISSION) != 0
    f = np.where(spectral, m["emissive"][:, 0], f)
    has_tex = (m["textures"] & TX_EMISSION) != 0
    tscale = np.where(has_tex, tex_avg[np.maximum(m["tex"][:, 6], 0)] if len(tex_avg) else 1.0, 1.0)
    return (f * tscale * area * np.pi).astype(np.float32)


def disk_light(pos, normal, radius, power, material):
    """A disk light record (world space) with the frame scene.h:386-392 builds
    for a disk: tangent = normalize(cross(z, n)) (x instead of z when n is
    within 0.9999 of +-z), bitangent = n x tangent."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    a = np.array([0.0, 0.0, 1.0]) if abs(n[2]) < 0.9999 else np.array([1.0, 0.0, 0.0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    r = np.zeros((), SHAPE_LIGHT_DTYPE)
    r["position"], r["radius"], r["powerScale"] = pos, radius, power
    r["frame"] = _quat_exact(t, b, n)
    r["materialIndex"], r["type"] = material, 0
    return r


def sphere_light(pos, radius, power, material):
    r = np.zeros((), SHAPE_LIGHT_DTYPE)
    r["position"], r["radius"], r["powerScale"] = pos, radius, power
    r["frame"] = (0.0, 0.0, 0.0, 1.0)
    r["materialIndex"], r["type"] = material, 1
    return r


def shape_flux(shapes, materials):
    """Flux of each shape light, as triangle_flux weighs an emitter: mean
    emissive factor (spectral emitters: the spectrum's scale) x area x pi x
    the light's power scale; a sphere's area is its whole surface."""
    m = materials[shapes["materialIndex"]]
    f = m["emissive"].mean(axis=1).astype(np.float64)
    f = np.where((m["flags"] & MF_SPECTRAL_EMISSION) != 0, m["emissive"][:, 0], f)
    r = shapes["radius"].astype(np.float64)
    area = np.where(shapes["type"] == 0, np.pi * r * r, 4.0 * np.pi * r * r)
    return (f * area * np.pi * shapes["powerScale"]).astype(np.float32)


def _quat_exact(t, b, n):
    """quat_from_frame of one frame without the w >= 1e-6 floor that vertex
    frames need."""
    m = np.stack([t, b, n], 1)
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        r = math.sqrt(1.0 + tr) * 2.0
        q = [(m[2, 1] - m[1, 2]) / r, (m[0, 2] - m[2, 0]) / r, (m[1, 0] - m[0, 1]) / r, 0.25 * r]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        r = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2.0
        q = [0.25 * r, (m[0, 1] + m[1, 0]) / r, (m[0, 2] + m[2, 0]) / r, (m[2, 1] - m[1, 2]) / r]
    elif m[1, 1] > m[2, 2]:
        r = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2.0
        q = [(m[0, 1] + m[1, 0]) / r, 0.25 * r, (m[1, 2] + m[2, 1]) / r, (m[0, 2] - m[2, 0]) / r]
    else:
        r = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2.0
        q = [(m[0, 2] + m[2, 0]) / r, (m[1, 2] + m[2, 1]) / r, 0.25 * r, (m[1, 0] - m[0, 1]) / r]
    q = np.array(q)
    return q / np.linalg.norm(q)


def value_noise_texture(rng, size=256, cells=8):
    g = rng.random((cells + 1, cells + 1, 3))
    g[cells] = g[0]
    g[:, cells] = g[:, 0]
    t = (np.arange(size) + 0.5) / size * cells
    i = np.floor(t).astype(int)
    f = t - i
    f = f * f * (3 - 2 * f)
    a = g[i][:, i]
    b = g[i + 1][:, i]
    c = g[i][:, i + 1]
    d = g[i + 1][:, i + 1]
    fy = f[:, None, None]
    fx = f[None, :, None]
    val = (a * (1 - fy) + b * fy) * (1 - fx) + (c * (1 - fy) + d * fy) * fx
    rgba = np.concatenate([val, np.ones((size, size, 1))], axis=2)
    return (np.clip(rgba, 0, 1) * 255 + 0.5).astype(np.uint8)


def uv_footprint_average(tex):
    """Mean (r+g+b)*a/3 of the texels whose centres fall inside the UV
    triangle (0,0) (1,0) (0,1), the rasterised footprint of
    src/core/scene.cpp:571-586 for that triangle."""
    h, w = tex.shape[:2]
    ys, xs = np.mgrid[0:h, 0:w]
    inside = ((xs + 0.5) / w + (ys + 0.5) / h) <= 1.0
    t = tex.astype(np.float64) / 255.0
    e = (t[..., 0] + t[..., 1] + t[..., 2]) * t[..., 3] / 3.0
    return float(e[inside].mean())


def _transform(verts, cam_pos):
    return (np.asarray(verts, np.float64) - cam_pos).astype(np.float32)


# ------------------------------------------------------------------ shapes
def quad(p0, p1, p2, p3):
    v = np.array([p0, p1, p2, p3], np.float64)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return v, idx, np.repeat(n[None], 4, 0), uv


def uv_sphere(center, radius, seg=32, rings=16):
    verts, norms, uvs = [], [], []
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(seg + 1):
            ph = 2 * math.pi * s / seg
            n = np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
            verts.append(np.asarray(center) + radius * n)
            norms.append(n)
            uvs.append((s / seg, r / rings))
    idx = []
    for r in range(rings):
        for s in range(seg):
            a = r * (seg + 1) + s
            b = a + seg + 1
            if r != 0:
                idx.append((a, a + 1, b))
            if r != rings - 1:
                idx.append((a + 1, b + 1, b))
    return np.array(verts), np.array(idx, np.uint32), np.array(norms), np.array(uvs, np.float32)


def icosphere(level):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = np.array(f, np.int64)
    for _ in range(level):
        cache = {}
        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(nf, np.int64)
    return np.array(verts), faces.astype(np.uint32)


def vertex_normals(verts, idx):
    v = verts[idx]
    fn = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n = np.zeros_like(verts)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-20)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _finish(parts, mats, spectra, tex_list, cam, env, width, height, depth, spp, name, tables, light_uv=None,
            shapes=None):
    verts = np.concatenate([p[0] for p in parts])
    norms = np.concatenate([p[2] for p in parts])
    uvs = np.concatenate([p[3] for p in parts])
    idx, tm, off = [], [], 0
    for p in parts:
        idx.append(p[1] + off)
        tm.append(np.full(len(p[1]), p[4], np.uint32) if np.isscalar(p[4]) else p[4])
        off += len(p[0])
    idx = np.concatenate(idx).astype(np.uint32)
    tri_material = np.concatenate(tm).astype(np.uint32)
    tri_light = np.full(len(idx), NO_LIGHT, np.uint32)
    materials = np.array(mats, MATERIAL_DTYPE)
    em = ((materials["factors"][tri_material] & FC_EMISSION) != 0)
    light_ids = np.nonzero(em)[0].astype(np.uint32)
    tri_light[light_ids] = np.arange(len(light_ids), dtype=np.uint32)
    verts_c = _transform(verts, cam["pos"])
    if light_uv is None:
        luv = uvs[idx[light_ids]]
    else:
        luv = np.broadcast_to(light_uv, (len(light_ids), 3, 2))
    lights = emissive_records(verts_c, idx, light_ids, tri_material, np.ascontiguousarray(luv, np.float32))
    # textures
    descs = np.zeros(len(tex_list), TEXTURE_DTYPE)
    blob, offs, avgs = [], 0, []
    for i, t in enumerate(tex_list):
        descs[i]["width"], descs[i]["height"], descs[i]["depth"] = t.shape[1], t.shape[0], 1
        descs[i]["format"], descs[i]["srgb"], descs[i]["addr"] = 0, 1, (0, 0, 0)
        descs[i]["filter"], descs[i]["offset"] = 1, offs
        blob.append(t.reshape(-1))
        offs += t.size
        avgs.append(uv_footprint_average(t))
    texels = np.concatenate(blob) if blob else np.zeros(0, np.uint8)
    flux = triangle_flux(lights, materials, np.array(avgs))
    sflux = None
    if shapes is not None:
        shapes = np.array(shapes, SHAPE_LIGHT_DTYPE).reshape(-1)
        shapes["position"] = _transform(shapes["position"], cam["pos"])
        sflux = shape_flux(shapes, materials)
    return Scene(verts=verts_c, indices=idx, attribs=pack_attribs(norms, uvs), tri_material=tri_material,
                 tri_light=tri_light, materials=materials,
                 spectra=np.ascontiguousarray(np.array(spectra, np.float32).reshape(-1, 471)),
                 textures=descs, texels=texels, lights=lights, flux=flux, env=env, camera=cam, width=width,
                 height=height, depth=depth, spp=spp, tables=tables, name=name, shapes=shapes, shape_flux=sflux)


# ----------------------------------------------------------------- configs
def cornell(seed=0x5EED0001, width=256, height=256, spp=16, depth=8, light_up=False, occluder=False, bounce=False,
            shapes=None):
    """Seeded Cornell box. With light_up the emitter faces the ceiling from
    just below it, so every other surface is lit only indirectly. With
    occluder a plate under the ceiling light hides it from the room, which is
    then lit only through the gap between the plate and the walls (a scene
    where BSDF sampling rarely finds the light's path and NEE is occluded).
    With bounce a small emitter 0.5 in front of the back wall faces it: the
    room is lit by the bright patch it throws on the wall, which NEE cannot
    target and BSDF sampling finds only by chance, while the patch's own NEE
    is well conditioned (no emitter within 0.4 of a lit surface); the spheres
    stay in the lower half of the box, out of the emitter's view, and every
    surface is at least 0.4 rough, so that no caustic path adds its own
    variance.

    shapes="mixed" adds a disk light under the ceiling (facing down,
    blackbody) and a sphere light (RGB emission) to the ceiling quad;
    shapes="only" replaces the quad with them."""
    rng = np.random.default_rng(seed)
    tables = load_tables()
    mats, parts = [], []
    spectra = [tables["blackbody4000"]]

    def gd():
        rough = float(rng.uniform(0.05, 1.0))
        if bounce:
            rough = 0.4 + 0.6 * rough   # no sharp highlights: the variance is in the indirect light
        return material(base=rng.uniform(0.1, 0.9, 3), metal=0.0, rough=rough, ior=1.5)

    walls = [
        ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)),          # floor (faces +y)
        ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),          # ceiling (faces -y)
        ((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),        # back wall (faces +z)
        ((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),        # left wall (faces +x)
        ((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)),            # right wall (faces -x)
    ]
    for w in walls:
        v, i, n, uv = quad(*w)
        if np.dot(n[0], np.array([0.0, 1.0, 0.0]) - v.mean(axis=0)) < 0:   # face into the box
            v, i, n, uv = quad(*w[::-1])
        mats.append(gd())
        parts.append((v, i, n, uv, len(mats) - 1))
    # emissive quad just below the ceiling, facing down, blackbody 4000 K
    if bounce:
        v, i, n, uv = quad((-0.2, 1.4, -0.5), (-0.2, 1.8, -0.5), (0.2, 1.8, -0.5), (0.2, 1.4, -0.5))
        power = 60.0
    elif light_up:
        v, i, n, uv = quad((-0.4, 1.9, -0.4), (-0.4, 1.9, 0.4), (0.4, 1.9, 0.4), (0.4, 1.9, -0.4))
        power = 60.0
    else:
        v, i, n, uv = quad((-0.3, 1.98, -0.3), (0.3, 1.98, -0.3), (0.3, 1.98, 0.3), (-0.3, 1.98, 0.3))
        power = 12.0
    mats.append(material(base=(0.8, 0.8, 0.8), metal=0.0, rough=1.0, spectral_emission=(0, power)))
    if shapes != "only":
        mats[-1]["factors"] |= FC_EMISSION
        parts.append((v, i, n, uv, len(mats) - 1))
    shape_recs = None
    if shapes is not None:
        mats.append(material(base=(0.0, 0.0, 0.0), metal=0.0, rough=1.0, spectral_emission=(0, 10.0)))
        mats.append(material(base=(0.0, 0.0, 0.0), metal=0.0, rough=1.0, emissive=(3.0, 2.4, 1.6)))
        shape_recs = [disk_light((0.55, 1.97, -0.45), (0.0, -1.0, 0.0), 0.22, 1.5, len(mats) - 2),
                      sphere_light((-0.55, 1.55, 0.25), 0.1, 2.0, len(mats) - 1)]
    if occluder:
        # grey plate 0.12 below the light (intersection is two-sided)
        mats.append(material(base=(0.6, 0.6, 0.6), metal=0.0, rough=0.8, ior=1.5))
        v, i, n, uv = quad((-0.75, 1.86, -0.75), (0.75, 1.86, -0.75), (0.75, 1.86, 0.75), (-0.75, 1.86, 0.75))
        parts.append((v, i, n, uv, len(mats) - 1))
    for _ in range(8):
        r = float(rng.uniform(0.1, 0.3))
        c = rng.uniform([-1 + r, r, -1 + r], [1 - r, 2 - r, 1 - r])
        if bounce:
            # lower half, behind the emitter's plane: no sphere sees its face
            c[1] = min(c[1], 1.0)
            c[2] = max(c[2], -0.4 + r)
        v, i, n, uv = uv_sphere(c, r)
        mats.append(gd())
        parts.append((v, i, n, uv, len(mats) - 1))
    cam = camera_basis((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(45.0)
    return _finish(parts, mats, spectra, [], cam, None, width, height, depth, spp, "cornell", tables,
                   shapes=shape_recs)


def shape_lit(kind="disk", triangles=False, width=48, height=48, spp=256, depth=4):
    """A floor and a back wall lit by one shape light: a tilted disk (kind
    "disk", blackbody emission) or a sphere (kind "sphere", RGB emission), or
    with triangles=True the same emitter as emissive triangles: a 256-triangle
    fan of the disk (area ratio 0.9999) or a level-4 icosphere of the sphere
    (outward faces). Both versions have the same expectation; the camera sees
    the light too."""
    tables = load_tables()
    spectra = [tables["blackbody4000"]]
    floor = quad((-2, 0, -2), (-2, 0, 2), (2, 0, 2), (2, 0, -2))
    wall = quad((-2, 0, -2), (2, 0, -2), (2, 3, -2), (-2, 3, -2))
    mats = [material(base=(0.7, 0.7, 0.7), metal=0.0, rough=0.8, ior=1.5),
            material(base=(0.5, 0.6, 0.7), metal=0.0, rough=0.6, ior=1.5)]
    parts = [floor + (0,), wall + (1,)]
    shapes = None
    if kind == "disk":
        c, n, r, power = np.array([0.2, 1.2, -0.3]), np.array([0.3, -1.0, 0.25]), 0.45, 1.5
        rec = disk_light(c, n, r, power, 2)
        if triangles:
            mats.append(material(base=(0, 0, 0), metal=0.0, rough=1.0, spectral_emission=(0, 4.0 * power)))
            mats[-1]["factors"] |= FC_EMISSION
            q = rec["frame"].astype(np.float64)
            x, y, z, w = q
            t = np.array([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)])
            b = np.array([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)])
            k = 256
            ang = 2 * np.pi * np.arange(k) / k
            rim = c + r * (np.cos(ang)[:, None] * t + np.sin(ang)[:, None] * b)
            v = np.concatenate([c[None], rim])
            idx = np.stack([np.zeros(k), 1 + np.arange(k), 1 + (np.arange(k) + 1) % k], 1).astype(np.uint32)
            nn = np.repeat((n / np.linalg.norm(n))[None], len(v), 0)
            parts.append((v, idx, nn, np.zeros((len(v), 2), np.float32), 2))
        else:
            mats.append(material(base=(0, 0, 0), metal=0.0, rough=1.0, spectral_emission=(0, 4.0)))
            shapes = [rec]
    else:
        c, r, power = np.array([0.0, 0.9, -0.4]), 0.35, 2.0
        if triangles:
            mats.append(material(base=(0, 0, 0), metal=0.0, rough=1.0,
                                 emissive=(2.0 * power, 1.6 * power, 1.2 * power)))
            mats[-1]["factors"] |= FC_EMISSION
            iv, ii = icosphere(4)
            parts.append((c + iv * r, ii.astype(np.uint32), iv, np.zeros((len(iv), 2), np.float32), 2))
        else:
            mats.append(material(base=(0, 0, 0), metal=0.0, rough=1.0, emissive=(2.0, 1.6, 1.2)))
            shapes = [sphere_light(c, r, power, 2)]
    cam = camera_basis((0.0, 1.6, 3.2), (0.0, 0.5, -0.5), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(50.0)
    return _finish(parts, mats, spectra, [], cam, None, width, height, depth, spp, "shape-" + kind, tables,
                   light_uv=np.zeros((3, 2), np.float32), shapes=shapes)


def furnace(inner="metal", rough=0.3, width=64, height=64, spp=64, depth=8, level=3, radius=1.5):
    """An emissive enclosure around a sphere: the inside of a radius-10
    icosphere whose triangles all emit white (1, 1, 1) and reflect almost
    nothing (ComplexSurface metal with a black base colour: Schlick with
    f0 = 0, seen near normal incidence from the centre), and a level-`level`
    icosphere of `radius` at the centre with a white, non-absorbing material:
    "metal" (metalness 1, base white, multiple-scattering compensation) or
    "dielectric" (metalness 0, base white, ior 1.5) or "glass" (a PureDielectric
    with a dispersive index of refraction). Every direction seen from
    the inner sphere carries the wall radiance Le, so its pixels converge to
    Le times its directional albedo (one within the LUTs' accuracy), like the
    pixels that see the wall directly. The camera looks at the centre from
    distance 5."""
    tables = load_tables()
    ev, ei = icosphere(3)
    ev = ev * 10.0
    ei = ei[:, [0, 2, 1]].copy()                 # normals towards the centre
    en = -ev / 10.0
    euv = np.zeros((len(ev), 2), np.float32)
    iv, ii = icosphere(level)
    iv = iv * radius
    inn = vertex_normals(iv, ii)
    iuv = np.zeros((len(iv), 2), np.float32)
    mats = [material(base=(0.0, 0.0, 0.0), metal=1.0, rough=0.5, emissive=(1.0, 1.0, 1.0))]
    spectra = []
    if inner == "metal":
        mats.append(material(base=(1.0, 1.0, 1.0), metal=1.0, rough=rough))
    elif inner == "glass":
        # PureDielectric with a dispersive index of refraction (Cauchy b = 0.03 um^2)
        mats.append(material(bxdf=BXDF_DIELECTRIC, rough=rough))
        mats[-1]["iorSpectrum"] = 0
        spectra = [dispersive_ior(1.45, 0.03)]
    else:
        mats.append(material(base=(1.0, 1.0, 1.0), metal=0.0, rough=rough, ior=1.5))
    parts = [(ev, ei, en, euv, 0), (iv, ii, inn, iuv, 1)]
    cam = camera_basis((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(60.0)
    return _finish(parts, mats, spectra, [], cam, None, width, height, depth, spp, "furnace-" + inner, tables,
                   light_uv=np.zeros((3, 2), np.float32))


def quat_from_frame(t, b, n):
    """Unit quaternions (x, y, z, w) of the rotations whose matrix columns are
    the orthonormal frames (t, b, n), rows of (k, 3) arrays (Shepperd's
    method, the largest of the four candidates as pivot)."""
    t, b, n = (np.asarray(a, np.float64) for a in (t, b, n))
    m00, m10, m20 = t[:, 0], t[:, 1], t[:, 2]
    m01, m11, m21 = b[:, 0], b[:, 1], b[:, 2]
    m02, m12, m22 = n[:, 0], n[:, 1], n[:, 2]
    tr = m00 + m11 + m22
    cand = np.stack([tr, m00, m11, m22], 1)
    k = cand.argmax(axis=1)
    q = np.zeros((len(t), 4))
    for case in range(4):
        sel = k == case
        if not sel.any():
            continue
        a00, a11, a22 = m00[sel], m11[sel], m22[sel]
        a01, a10, a02, a20, a12, a21 = m01[sel], m10[sel], m02[sel], m20[sel], m12[sel], m21[sel]
        if case == 0:
            r = np.sqrt(1.0 + a00 + a11 + a22) * 2.0
            q[sel] = np.stack([(a21 - a12) / r, (a02 - a20) / r, (a10 - a01) / r, 0.25 * r], 1)
        elif case == 1:
            r = np.sqrt(1.0 + a00 - a11 - a22) * 2.0
            q[sel] = np.stack([0.25 * r, (a01 + a10) / r, (a02 + a20) / r, (a21 - a12) / r], 1)
        elif case == 2:
            r = np.sqrt(1.0 + a11 - a00 - a22) * 2.0
            q[sel] = np.stack([(a01 + a10) / r, 0.25 * r, (a12 + a21) / r, (a02 - a20) / r], 1)
        else:
            r = np.sqrt(1.0 + a22 - a00 - a11) * 2.0
            q[sel] = np.stack([(a02 + a20) / r, (a12 + a21) / r, 0.25 * r, (a10 - a01) / r], 1)
    # w >= 0, so that the sign of w is free to carry the bitangent flip
    q *= np.where(q[:, 3:4] < 0, -1.0, 1.0)
    q[:, 3] = np.maximum(q[:, 3], 1e-6)
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def sphere_frames(normals):
    """Tangent (along increasing u of uv_sphere), bitangent n x t, normal."""
    n = np.asarray(normals, np.float64)
    t = np.cross(n, np.array([0.0, 1.0, 0.0]))
    pole = np.linalg.norm(t, axis=1) < 1e-6
    t[pole] = (1.0, 0.0, 0.0)
    t = -t / np.linalg.norm(t, axis=1, keepdims=True)
    return t, np.cross(n, t), n


def _noise_tex(rng, size, cells, channels):
    return np.stack([_value_noise_2d(rng, size, cells) for _ in range(channels)], -1)


def features(seed=0x5EED0009, width=96, height=96, spp=4, depth=8):
    """The Cornell box with objects that exercise every surface-attribute and
    texture branch of the shading code (pathtrace.cu:1279-1567), which the
    benchmark scenes leave unused:
      - a sphere whose tangent frames are per-vertex quaternions (AT_FRAME),
        with a normal map read through TEX_COORD_2 and a base-colour texture;
      - the same with the bitangent flipped (negative quaternion w) and an
        RGB32F normal map on TEX_COORD_1 with mirror addressing;
      - a metal-roughness texture (TEX_COORD_2, nearest filtering, clamp);
      - a transmission texture (R8, ior 1.4);
      - sheen colour and sheen roughness textures;
      - an RGBA8 base-colour texture with checkered alpha and an alpha
        cutoff, and a quad with alpha blending (RGBA32F texture);
      - a PureDielectric (bxdf 1) sphere, a wall without vertex attributes
        (the geometric normal) and a PureConductor sphere.
    Every vertex holds 32 bytes: a frame quaternion or an octahedral normal
    at 0, TEX_COORD_1 at 16 and TEX_COORD_2 at 24."""
    rng = np.random.default_rng(seed)
    tables = load_tables()
    spectra = [tables["blackbody4000"]]
    mats, parts, frame_parts = [], [], set()
    walls = [
        ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)),
        ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),
        ((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)),
        ((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)),
        ((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)),
    ]
    for w in walls:
        v, i, n, uv = quad(*w)
        if np.dot(n[0], np.array([0.0, 1.0, 0.0]) - v.mean(axis=0)) < 0:
            v, i, n, uv = quad(*w[::-1])
        mats.append(material(base=rng.uniform(0.2, 0.9, 3), rough=float(rng.uniform(0.3, 1.0))))
        parts.append((v, i, n, uv, len(mats) - 1))
    no_attrib_mat = 2          # the back wall shades with its geometric normal
    v, i, n, uv = quad((-0.3, 1.98, -0.3), (0.3, 1.98, -0.3), (0.3, 1.98, 0.3), (-0.3, 1.98, 0.3))
    mats.append(material(base=(0.8, 0.8, 0.8), rough=1.0, spectral_emission=(0, 12.0)))
    mats[-1]["factors"] |= FC_EMISSION
    parts.append((v, i, n, uv, len(mats) - 1))

    # textures: (array, format, srgb, addr, filter)
    tex = []

    def add_tex(a, fmt, srgb, addr=(0, 0, 0), filt=1):
        tex.append((a, fmt, srgb, addr, filt))
        return len(tex) - 1

    def normal_map(size, amp):
        h = _value_noise_2d(rng, size, 4)
        gy, gx = np.gradient(h * amp * size / 4.0)
        nm = np.stack([-gx, -gy, np.ones_like(h)], -1)
        return nm / np.linalg.norm(nm, axis=-1, keepdims=True)

    nm8 = normal_map(64, 0.6)
    t_nm8 = add_tex(np.concatenate([np.round((nm8 * 0.5 + 0.5) * 255), np.full((64, 64, 1), 255)], -1)
                    .astype(np.uint8), 0, 0)
    t_nm32 = add_tex(np.ascontiguousarray(normal_map(32, 0.8), np.float32), 3, 0, addr=(2, 2, 2))
    base = np.round(_noise_tex(rng, 64, 4, 3) * 255)
    checker = ((np.arange(64)[:, None] // 8 + np.arange(64)[None, :] // 8) % 2).astype(np.float64)
    alpha = np.where(checker > 0, 255, 40)[..., None]
    t_base = add_tex(np.concatenate([base, alpha], -1).astype(np.uint8), 0, 1)
    mr = np.zeros((32, 32, 4))
    mr[..., 1] = np.round(_value_noise_2d(rng, 32, 4) * 255)
    mr[..., 2] = np.where(_value_noise_2d(rng, 32, 2) > 0.5, 255, 0)
    mr[..., 3] = 255
    t_mr = add_tex(mr.astype(np.uint8), 0, 0, addr=(1, 1, 1), filt=0)
    t_tr = add_tex(np.round(_value_noise_2d(rng, 32, 4) * 255).astype(np.uint8), 1, 0)
    t_shc = add_tex(np.concatenate([np.round(_noise_tex(rng, 32, 4, 3) * 255), np.full((32, 32, 1), 255)], -1)
                    .astype(np.uint8), 0, 1)
    shr = np.full((32, 32, 4), 255.0)
    shr[..., 3] = np.round((0.2 + 0.8 * _value_noise_2d(rng, 32, 4)) * 255)
    t_shr = add_tex(shr.astype(np.uint8), 0, 0)
    blend = np.ones((16, 16, 4), np.float32)
    blend[..., :3] = rng.uniform(0.2, 0.9, 3)
    blend[..., 3] = np.linspace(0.1, 0.9, 16)[None, :]
    t_blend = add_tex(blend, 2, 0, addr=(1, 1, 1))

    def sphere(c, r, m, frame=False):
        v, i, n, uv = uv_sphere(c, r)
        mats.append(m)
        parts.append((v, i, n, uv, len(mats) - 1))
        if frame:
            frame_parts.add(len(parts) - 1)
        return mats[-1]

    m = material(base=(1, 1, 1), rough=0.35)
    m["textures"] = TX_NORMAL | TX_BASE_COLOR
    m["tex"][1], m["tex"][0] = t_nm8, t_base
    m["texCoordSets"] = 1 << 1                    # the normal map (and so the base colour) read TEX_COORD_2
    sphere((-0.55, 0.3, -0.4), 0.28, m, frame=True)
    m = material(base=(0.9, 0.6, 0.3), metal=0.0, rough=0.25)
    m["textures"] = TX_NORMAL
    m["tex"][1] = t_nm32
    sphere((0.55, 0.3, -0.4), 0.28, m, frame=True)   # flipped below
    m = material(base=(0.8, 0.8, 0.85), metal=1.0, rough=0.8)
    m["textures"] = TX_MET_ROUGH
    m["tex"][2] = t_mr
    m["texCoordSets"] = 1 << 2
    sphere((0.0, 0.25, 0.1), 0.25, m)
    m = material(base=(0.9, 0.95, 1.0), rough=0.15, ior=1.4, transmission=1.0)
    m["textures"] = TX_TRANSMISSION
    m["tex"][3] = t_tr
    sphere((-0.5, 1.0, 0.2), 0.22, m)
    m = material(base=(0.3, 0.2, 0.4), rough=0.6, sheen=((1.0, 1.0, 1.0), 1.0))
    m["textures"] = TX_SHEEN_COLOR | TX_SHEEN_ROUGH
    m["tex"][4], m["tex"][5] = t_shc, t_shr
    sphere((0.5, 1.0, 0.2), 0.22, m)
    eta, kk = cauchy_spectra(rng)
    spectra += [eta, kk, dispersive_ior(1.45, 0.03)]
    # PureDielectric spheres with a dispersive index of refraction: rough
    # (alpha 0.01) and smooth
    m = material(bxdf=BXDF_DIELECTRIC, base=(1, 1, 1), rough=0.1, ior=1.5)
    m["iorSpectrum"] = 3
    sphere((0.0, 1.2, -0.5), 0.2, m)
    m = material(bxdf=BXDF_DIELECTRIC, base=(1, 1, 1), rough=0.0, ior=1.5)
    m["iorSpectrum"] = 3
    sphere((0.45, 1.55, -0.3), 0.15, m)
    sphere((0.0, 0.7, 0.45), 0.15, material(bxdf=BXDF_CONDUCTOR, rough=0.3, eta_k=(1, 2)))
    # alpha-cutoff quad (base-colour alpha 40/255 or 1 in a checker) and an
    # alpha-blended quad in front of the back half of the box
    m = material(base=(1, 1, 1), rough=0.7)
    m["textures"] = TX_BASE_COLOR
    m["tex"][0] = t_base
    m["factors"] |= FC_CUTOFF
    m["alphaCutoff"] = 0.5
    mats.append(m)
    v, i, n, uv = quad((-0.9, 0.05, -0.75), (-0.1, 0.05, -0.75), (-0.1, 0.85, -0.75), (-0.9, 0.85, -0.75))
    parts.append((v, i, n, uv * 2.0, len(mats) - 1))
    m = material(base=(1, 1, 1), rough=0.5)
    m["textures"] = TX_BASE_COLOR
    m["tex"][0] = t_blend
    m["factors"] = (int(m["factors"]) & ~FC_BASE_COLOR & 0xFFFFFFFF) | FC_BLEND
    mats.append(m)
    v, i, n, uv = quad((0.1, 0.05, -0.7), (0.9, 0.05, -0.7), (0.9, 0.85, -0.7), (0.1, 0.85, -0.7))
    parts.append((v, i, n, uv, len(mats) - 1))

    cam = camera_basis((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(45.0)
    sc = _finish(parts, mats, spectra, [], cam, None, width, height, depth, spp, "features", tables)

    # 32-byte vertex attributes
    stride = 32
    nv = sum(len(p[0]) for p in parts)
    attr = np.zeros((nv, 8), np.float32)
    off = 0
    for pi, pt in enumerate(parts):
        k = len(pt[0])
        uv1 = np.asarray(pt[3], np.float32)
        attr[off:off + k, 4:6] = uv1
        attr[off:off + k, 6:8] = uv1 * np.float32(3.0) + np.float32(0.25)      # TEX_COORD_2
        if pi in frame_parts:
            t, b, nn = sphere_frames(pt[2])
            attr[off:off + k, 0:4] = quat_from_frame(t, b, nn)
        else:
            attr[off:off + k, 0:2] = oct_encode(pt[2])
        off += k
    # the second frame sphere stores its bitangent flipped (w < 0)
    fparts = sorted(frame_parts)
    off = sum(len(parts[j][0]) for j in range(fparts[1]))
    k = len(parts[fparts[1]][0])
    attr[off:off + k, 0:4] *= -1.0
    sc.attribs = attr.view(np.uint8).reshape(-1).copy()
    mt = sc.materials
    mt["attributeStride"] = stride
    mt["normalOffset"] = 0
    mt["frameOffset"] = 0
    mt["texCoord1Offset"] = 16
    mt["texCoord2Offset"] = 24
    mt["attributes"] = AT_NORMAL | AT_TEX1 | AT_TEX2
    for j in fparts:
        mt["attributes"][parts[j][4]] = AT_FRAME | AT_TEX1 | AT_TEX2
    mt["attributes"][no_attrib_mat] = 0
    # textures
    descs = np.zeros(len(tex), TEXTURE_DTYPE)
    blob, offs = [], 0
    for ti, (a, fmt, srgb, addr, filt) in enumerate(tex):
        raw = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
        descs[ti]["width"], descs[ti]["height"], descs[ti]["depth"] = a.shape[1], a.shape[0], 1
        descs[ti]["format"], descs[ti]["srgb"], descs[ti]["addr"] = fmt, srgb, addr
        descs[ti]["filter"], descs[ti]["offset"] = filt, offs
        pad = (-len(raw)) % 256
        blob += [raw, np.zeros(pad, np.uint8)]
        offs += len(raw) + pad
    sc.textures = descs
    sc.texels = np.concatenate(blob)
    return sc


def batch(n=64, seed=0x5EED0003):
    """The batched uneven workload: n seeded spheres scenes with triangle
    counts log-uniform in [1k, 500k] (level-2 or level-3 icospheres), films
    drawn from 64^2 .. 1024^2, emissive-triangle counts log-uniform in
    [1, 5000], the four-way material mix, 32 spp, depth 12; every other
    scene has a 256x128 log-normal env map."""
    rng = np.random.default_rng(seed)
    sides = [64, 128, 256, 512, 1024]
    out = []
    for i in range(n):
        tris = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        level = 3 if tris > 2e5 else 2
        per = 20 * 4 ** level
        n_sph = max(1, int(round(tris / per)))
        side = int(rng.choice(sides))
        n_lights = int(np.exp(rng.uniform(0.0, np.log(5000.0))))
        env = None if i % 2 else (256, 128)
        out.append(spheres(seed=seed + i, n_spheres=n_sph, level=level, n_lights=n_lights, width=side,
                           height=side, spp=32, depth=12, env_size=env))
    return out


def dispersive_ior(a, b):
    """A dense 1 nm index-of-refraction spectrum n = a + b / lambda^2
    (lambda in um): Cauchy's dispersion law for a glass."""
    lam = np.arange(360, 831, dtype=np.float64) * 1e-3
    return (a + b / lam ** 2).astype(np.float32)


def cauchy_spectra(rng):
    """Seeded Cauchy fits n = A + B / l^2, k = C + D / l^2 (l in um) at 1 nm."""
    lam = np.arange(360, 831, dtype=np.float64) * 1e-3
    A, B = rng.uniform(0.2, 1.5), rng.uniform(0.0, 0.15)
    C, D = rng.uniform(1.5, 4.0), rng.uniform(0.0, 0.6)
    return (A + B / lam ** 2).astype(np.float32), (C + D / lam ** 2).astype(np.float32)


def spheres(seed=0x5EED0002, n_spheres=200, level=5, n_lights=10000, n_light_tex=16, width=1920, height=1080,
            spp=64, depth=16, env_size=(2048, 1024), env_sigma=2.0):
    rng = np.random.default_rng(seed)
    tables = load_tables()
    base_v, base_i = icosphere(level)
    mats, parts, spectra = [], [], []
    for s in range(n_spheres):
        R = random_rotation(rng)
        scale = float(np.exp(rng.uniform(np.log(0.2), np.log(5.0))))
        center = np.array([rng.uniform(-60, 60), rng.uniform(-8, 25), rng.uniform(-120, -15)])
        # radial displacement from smooth noise on the unit sphere
        k = rng.normal(size=(4, 3))
        ph = rng.uniform(0, 2 * np.pi, 4)
        disp = 1.0 + 0.08 * np.sum(np.sin(base_v @ k.T * 3.0 + ph), axis=1) / 4.0
        v = (base_v * disp[:, None]) @ R.T * scale + center
        n = vertex_normals(v, base_i)
        uv = np.stack([0.5 + np.arctan2(base_v[:, 2], base_v[:, 0]) / (2 * np.pi),
                       np.arccos(np.clip(base_v[:, 1], -1, 1)) / np.pi], axis=1).astype(np.float32)
        r = rng.random()
        rough = float(rng.uniform(0.02, 1.0))
        if r < 0.40:
            m = material(base=rng.uniform(0.1, 0.9, 3), metal=0.0, rough=rough)
        elif r < 0.65:
            eta, kk = cauchy_spectra(rng)
            spectra += [eta, kk]
            m = material(bxdf=BXDF_CONDUCTOR, rough=rough, eta_k=(len(spectra) - 2, len(spectra) - 1))
        elif r < 0.85:
            m = material(base=rng.uniform(0.5, 1.0, 3), metal=0.0, rough=rough, ior=float(rng.uniform(1.3, 1.8)),
                         transmission=1.0)
        else:
            m = material(base=rng.uniform(0.1, 0.9, 3), metal=0.0, rough=rough,
                         sheen=(rng.uniform(0.3, 1.0, 3), float(rng.uniform(0.05, 1.0))))
        mats.append(m)
        parts.append((v, base_i, n, uv, len(mats) - 1))
    # textured emissive triangles scattered through the volume
    tex_list = [value_noise_texture(rng) for _ in range(n_light_tex)]
    em_first = len(mats)
    for t in range(n_light_tex):
        mats.append(material(base=(0.8, 0.8, 0.8), rough=1.0, emissive=rng.uniform(2.0, 8.0, 3), emission_tex=t))
    c = rng.uniform([-70, -10, -130], [70, 30, -10], size=(n_lights, 3))
    size = np.exp(rng.uniform(np.log(0.3), np.log(1.5), n_lights))
    lv = np.zeros((n_lights * 3, 3))
    for k in range(3):
        d = rng.normal(size=(n_lights, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        lv[k::3] = c + d * size[:, None]
    li = np.arange(n_lights * 3, dtype=np.uint32).reshape(-1, 3)
    ln = vertex_normals(lv, li)
    luv = np.tile(np.array([[0, 0], [1, 0], [0, 1]], np.float32), (n_lights, 1))
    lmat = (em_first + rng.integers(0, n_light_tex, n_lights)).astype(np.uint32)
    parts.append((lv, li, ln, luv, lmat))
    if not spectra:
        spectra = [np.ones(471, np.float32)]
    cam = camera_basis((0.0, 5.0, 10.0), (0.0, 4.0, -60.0), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(60.0)
    env = lognormal_envmap(env_size[0], env_size[1], sigma=env_sigma, seed=seed, sun=True) if env_size else None
    return _finish(parts, mats, spectra, tex_list, cam, env, width, height, depth, spp, "spheres", tables)


def _value_noise_2d(rng, n, cells):
    """n x n smooth value noise in [0, 1] from a (cells+1)^2 random lattice."""
    g = rng.random((cells + 1, cells + 1))
    t = np.arange(n) / max(1, n - 1) * cells
    i = np.minimum(np.floor(t).astype(int), cells - 1)
    f = t - i
    f = f * f * (3 - 2 * f)
    fy, fx = f[:, None], f[None, :]
    a, b = g[i][:, i], g[i + 1][:, i]
    c, d = g[i][:, i + 1], g[i + 1][:, i + 1]
    return (a * (1 - fy) + b * fy) * (1 - fx) + (c * (1 - fy) + d * fy) * fx


def heightfield(rng, n, x_range, z_range, base, amplitude, octaves=6):
    """Seeded fBm terrain of n x n vertices (2 (n-1)^2 triangles) with
    analytic normals from the height gradient."""
    h = np.zeros((n, n))
    amp, cells, norm = 1.0, 4, 0.0
    for _ in range(octaves):
        h += amp * _value_noise_2d(rng, n, cells)
        norm += amp
        amp *= 0.5
        cells *= 2
    h = base + amplitude * h / norm
    xs = np.linspace(x_range[0], x_range[1], n)
    zs = np.linspace(z_range[0], z_range[1], n)
    X, Z = np.meshgrid(xs, zs)
    verts = np.stack([X, h, Z], -1).reshape(-1, 3)
    dz, dx = np.gradient(h, zs, xs)
    nrm = np.stack([-dx, np.ones_like(h), -dz], -1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r = np.arange(n - 1, dtype=np.uint32)
    a = (r[:, None] * n + r[None, :]).ravel()
    idx = np.empty((2 * len(a), 3), np.uint32)
    idx[0::2] = np.stack([a, a + n, a + 1], 1)
    idx[1::2] = np.stack([a + 1, a + n, a + n + 1], 1)
    uv = np.stack([(X - x_range[0]) / (x_range[1] - x_range[0]), (Z - z_range[0]) / (z_range[1] - z_range[0])],
                  -1).reshape(-1, 2).astype(np.float32)
    return verts, idx, nrm, uv


def large(seed=0x5EED0005, n_spheres=800, level=4, terrain=1449, n_lights=100000, n_light_tex=16, width=1920,
          height=1080, spp=256, depth=24, env_size=(8192, 4096), env_sigma=3.0):
    """The large precision-split scene: 800 displaced level-4 icospheres
    (5,120 triangles each, 4.10M) over a seeded fBm heightfield terrain of
    1449^2 vertices (4.19M triangles), 8.29M triangles in total; 60% of the
    spheres translucent or metal (35% translucent, 25% Cauchy metal), the rest
    glossy-diffuse (25%) and fuzz (15%); 100,000 textured emissive triangles;
    an 8192x4096 log-normal env map with sigma 3 and no sun; 256 spp,
    depth 24."""
    rng = np.random.default_rng(seed)
    tables = load_tables()
    base_v, base_i = icosphere(level)
    mats, parts, spectra = [], [], []
    x_range, z_range = (-170.0, 170.0), (-300.0, 15.0)
    tv, ti, tn, tuv = heightfield(rng, terrain, x_range, z_range, base=-14.0, amplitude=14.0)
    mats.append(material(base=(0.35, 0.32, 0.28), metal=0.0, rough=0.85))
    parts.append((tv, ti, tn, tuv, 0))
    for s in range(n_spheres):
        R = random_rotation(rng)
        scale = float(np.exp(rng.uniform(np.log(0.2), np.log(5.0))))
        center = np.array([rng.uniform(-120, 120), rng.uniform(-2, 35), rng.uniform(-240, -15)])
        k = rng.normal(size=(4, 3))
        ph = rng.uniform(0, 2 * np.pi, 4)
        disp = 1.0 + 0.08 * np.sum(np.sin(base_v @ k.T * 3.0 + ph), axis=1) / 4.0
        v = (base_v * disp[:, None]) @ R.T * scale + center
        n = vertex_normals(v, base_i)
        uv = np.stack([0.5 + np.arctan2(base_v[:, 2], base_v[:, 0]) / (2 * np.pi),
                       np.arccos(np.clip(base_v[:, 1], -1, 1)) / np.pi], axis=1).astype(np.float32)
        r = rng.random()
        rough = float(rng.uniform(0.02, 1.0))
        if r < 0.35:
            m = material(base=rng.uniform(0.5, 1.0, 3), metal=0.0, rough=rough, ior=float(rng.uniform(1.3, 1.8)),
                         transmission=1.0)
        elif r < 0.60:
            eta, kk = cauchy_spectra(rng)
            spectra += [eta, kk]
            m = material(bxdf=BXDF_CONDUCTOR, rough=rough, eta_k=(len(spectra) - 2, len(spectra) - 1))
        elif r < 0.85:
            m = material(base=rng.uniform(0.1, 0.9, 3), metal=0.0, rough=rough)
        else:
            m = material(base=rng.uniform(0.1, 0.9, 3), metal=0.0, rough=rough,
                         sheen=(rng.uniform(0.3, 1.0, 3), float(rng.uniform(0.05, 1.0))))
        mats.append(m)
        parts.append((v, base_i, n, uv, len(mats) - 1))
    tex_list = [value_noise_texture(rng) for _ in range(n_light_tex)]
    em_first = len(mats)
    for t in range(n_light_tex):
        mats.append(material(base=(0.8, 0.8, 0.8), rough=1.0, emissive=rng.uniform(0.5, 2.0, 3), emission_tex=t))
    c = rng.uniform([-160, 0, -290], [160, 45, -10], size=(n_lights, 3))
    size = np.exp(rng.uniform(np.log(0.2), np.log(1.0), n_lights))
    lv = np.zeros((n_lights * 3, 3))
    for k in range(3):
        d = rng.normal(size=(n_lights, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        lv[k::3] = c + d * size[:, None]
    li = np.arange(n_lights * 3, dtype=np.uint32).reshape(-1, 3)
    ln = vertex_normals(lv, li)
    luv = np.tile(np.array([[0, 0], [1, 0], [0, 1]], np.float32), (n_lights, 1))
    lmat = (em_first + rng.integers(0, n_light_tex, n_lights)).astype(np.uint32)
    parts.append((lv, li, ln, luv, lmat))
    cam = camera_basis((0.0, 12.0, 10.0), (0.0, 6.0, -80.0), (0.0, 1.0, 0.0))
    cam["fov"] = math.radians(60.0)
    env = lognormal_envmap(env_size[0], env_size[1], sigma=env_sigma, seed=seed, sun=False) if env_size else None
    return _finish(parts, mats, spectra, tex_list, cam, env, width, height, depth, spp, "large", tables)
