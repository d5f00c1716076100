"""Python handle over one rbh_ctx: uploads, env-map tables, rendering.

Each method maps one-to-one onto an rbh_* C entry point and raises RbhError
with the library's message on a non-zero status.
"""
import ctypes

import numpy as np

from . import _lib as L


class Context:
    def __init__(self, device=0):
        self._l = L.lib()
        h = ctypes.c_void_p()
        rc = self._l.rbh_create(device, ctypes.byref(h))
        if rc != 0:
            raise L.RbhError(rc, "rbh_create(device=%d) failed" % device)
        self.h = h
        self.width = 0
        self.height = 0

    # --------------------------------------------------------------- basics
    def close(self):
        if self.h:
            self._l.rbh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc):
        if rc != 0:
            raise L.RbhError(rc, self._l.rbh_last_error(self.h).decode())
        return rc

    def call(self, name, *args):
        return self.check(getattr(self._l, name)(self.h, *args))

    def synchronize(self):
        self.call("rbh_synchronize")

    def query(self):
        rc = self._l.rbh_query(self.h)
        if rc < 0:
            self.check(rc)
        return rc == 1

    def stats(self):
        s = L.Stats()
        self.call("rbh_stats_get", ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in L.Stats._fields_}

    # -------------------------------------------------------------- uploads
    def upload_spectra(self, dense):
        dense = np.ascontiguousarray(dense, np.float32).reshape(-1, 471)
        self.call("rbh_upload_spectra", L.ptr(dense), dense.shape[0])

    def upload_color_tables(self, xyz, basis):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(3, 471)
        basis = np.ascontiguousarray(basis, np.float32).reshape(3, 471)
        self.call("rbh_upload_color_tables", L.ptr(xyz), L.ptr(basis))

    def upload_luts(self, luts):
        arrs = [np.ascontiguousarray(luts[k], np.float32) for k in
                ("conductor", "diel_outer", "diel_inner", "refl_outer", "refl_inner", "sheen")]
        self.call("rbh_upload_luts", *[L.ptr(a) for a in arrs])

    def upload_textures(self, descs, blob):
        descs = np.ascontiguousarray(descs, L.TEXTURE_DTYPE)
        blob = np.ascontiguousarray(blob, np.uint8)
        self.call("rbh_upload_textures", L.ptr(descs) if len(descs) else None, len(descs),
                  L.ptr(blob) if blob.size else None, blob.size)

    def upload_materials(self, mats):
        mats = np.ascontiguousarray(mats, L.MATERIAL_DTYPE)
        self.call("rbh_upload_materials", L.ptr(mats), len(mats))

    def upload_geometry(self, verts, indices, attribs, tri_material, tri_light):
        """indices: uint32, or uint16 (uploaded as a 16-bit index buffer)."""
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        fmt = L.RBH_INDEX_U16 if np.asarray(indices).dtype == np.uint16 else L.RBH_INDEX_U32
        indices = np.ascontiguousarray(indices, np.uint16 if fmt == L.RBH_INDEX_U16 else np.uint32).reshape(-1, 3)
        attribs = np.ascontiguousarray(attribs, np.uint8).reshape(-1)
        tri_material = np.ascontiguousarray(tri_material, np.uint32)
        tri_light = np.ascontiguousarray(tri_light, np.uint32)
        self.call("rbh_upload_geometry_ex", L.ptr(verts), len(verts), L.ptr(indices), fmt, len(indices),
                  L.ptr(attribs) if attribs.size else None, attribs.size, L.ptr(tri_material), L.ptr(tri_light))

    def build_accel(self):
        self.call("rbh_build_accel")

    def update_vertices(self, first, verts):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        self.call("rbh_update_vertices", int(first), len(verts), L.ptr(verts))

    def refit_accel(self):
        self.call("rbh_refit_accel")

    def translate_scene(self, delta):
        d = np.ascontiguousarray(delta, np.float32).reshape(3)
        self.call("rbh_translate_scene", L.ptr(d))

    def update_materials(self, first, mats):
        mats = np.ascontiguousarray(mats, L.MATERIAL_DTYPE)
        self.call("rbh_update_materials", int(first), len(mats), L.ptr(mats))

    def build_light_tree(self, tris, flux, fp64=False):
        tris = np.ascontiguousarray(tris, L.EMISSIVE_TRI_DTYPE)
        flux = np.ascontiguousarray(flux, np.float32)
        self.call("rbh_build_light_tree_ex", L.ptr(tris) if len(tris) else None, L.ptr(flux) if len(flux) else None,
                  len(tris), L.RBH_LIGHT_TREE_FP64 if fp64 else 0)

    def set_light_tree(self, nodes, pointers, masks, tris):
        """Uploads a light tree built by the caller (checked before use)."""
        nodes = np.ascontiguousarray(nodes, L.PACKED_NODE_DTYPE)
        pointers = np.ascontiguousarray(pointers, np.uint32)
        masks = np.ascontiguousarray(masks, np.uint64)
        tris = np.ascontiguousarray(tris, L.EMISSIVE_TRI_DTYPE)
        n = len(tris)
        self.call("rbh_set_light_tree", L.ptr(nodes) if len(nodes) else None, len(nodes),
                  L.ptr(pointers) if len(pointers) else None, L.ptr(masks) if len(masks) else None,
                  L.ptr(tris) if n else None, n)

    def upload_shape_lights(self, shapes, flux):
        """Disk and sphere lights (SHAPE_LIGHT_DTYPE) with their flux; they
        join the light tree at its next build."""
        shapes = np.ascontiguousarray(shapes, L.SHAPE_LIGHT_DTYPE)
        flux = np.ascontiguousarray(flux, np.float32)
        if len(flux) != len(shapes):
            raise ValueError("one flux per shape light")
        self.call("rbh_upload_shape_lights", L.ptr(shapes) if len(shapes) else None,
                  L.ptr(flux) if len(flux) else None, len(shapes))

    def get_light_tree(self, nlights):
        nodes = np.zeros(max(1, 2 * nlights), L.PACKED_NODE_DTYPE)
        lp = np.zeros(max(1, nlights), np.uint32)
        bm = np.zeros(max(1, nlights), np.uint64)
        n = ctypes.c_uint32(0)
        self.call("rbh_get_light_tree", L.ptr(nodes), ctypes.byref(n), L.ptr(lp), L.ptr(bm))
        return nodes[:n.value], lp[:nlights], bm[:nlights]

    # -------------------------------------------------------------- env map
    def envmap_build(self, texels, fp64=False):
        texels = np.ascontiguousarray(texels, np.float32)
        h, w, ch = texels.shape
        self.call("rbh_envmap_build", L.ptr(texels), w, h, ch, L.RBH_CDF_FP64 if fp64 else 0)
        self._env = (w, h)

    def envmap_rebuild(self, fp64=False, iterations=1):
        self.call("rbh_envmap_rebuild", L.RBH_CDF_FP64 if fp64 else 0, iterations)

    def envmap_read(self):
        w, h = self._env
        cond = np.zeros((h, w), np.uint16)
        marg = np.zeros(h, np.uint16)
        cdfy = np.zeros(h, np.float32)
        integral = ctypes.c_double(0)
        self.call("rbh_envmap_read", L.ptr(cond), L.ptr(marg), L.ptr(cdfy), ctypes.byref(integral))
        return cond, marg, cdfy, integral.value

    def envmap_disable(self):
        self.call("rbh_envmap_disable")

    # --------------------------------------------------------------- render
    def render(self, params):
        self.width, self.height = params.width, params.height
        self.call("rbh_render", ctypes.byref(params))

    def set_path_slots(self, slots):
        """Path slots of one wavefront pass (samples per pass = slots // pixels)."""
        self.call("rbh_set_path_slots", int(slots))

    def read_accum(self):
        a = np.zeros((self.height, self.width, 4), np.float64)
        self.call("rbh_read_accum", L.ptr(a))
        return a

    def write_accum(self, a):
        a = np.ascontiguousarray(a, np.float64)
        self.call("rbh_write_accum", L.ptr(a), a.shape[1], a.shape[0])

    def resolve(self, colorspace, exposure_scale):
        m = np.ascontiguousarray(colorspace, np.float32).reshape(9)
        out = np.zeros((self.height, self.width, 4), np.uint8)
        self.call("rbh_resolve", L.ptr(m), float(exposure_scale), L.ptr(out))
        return out
