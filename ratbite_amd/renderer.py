"""Host-side renderer over one rbh context.

Mirrors the orchestration of the reference's RenderingInterface
(src/core/rendering_interface.cpp): upload spectra, colour tables, LUTs,
textures, materials and geometry; build the acceleration structure, the light
tree (buildLightTree, 1065-1186) and the env-map tables (loadEnvironmentMap,
1339-1432); then render sub-launches into the FP64 accumulation buffer and
resolve (renderResolve, 1512-1533).
"""
import numpy as np

from .context import Context


class Renderer:
    def __init__(self, scene, device=0, ctx=None, fp64_cdf=False):
        self.scene = scene
        self.ctx = ctx or Context(device)
        t = scene.tables
        c = self.ctx
        c.upload_spectra(scene.spectra if len(scene.spectra) else np.ones((1, 471), np.float32))
        c.upload_color_tables(t["xyz"], t["basis"])
        c.upload_luts(t)
        c.upload_textures(scene.textures, scene.texels)
        c.upload_materials(scene.materials)
        c.upload_geometry(scene.verts, scene.indices, scene.attribs, scene.tri_material, scene.tri_light)
        c.build_accel()
        c.build_light_tree(scene.lights, scene.flux)
        if scene.env is not None:
            c.envmap_build(scene.env, fp64=fp64_cdf)
        else:
            c.envmap_disable()
        self.samples = 0

    def light_tree(self):
        return self.ctx.get_light_tree(len(self.scene.lights))

    def render(self, spp=None, offset=None, **kw):
        spp = self.scene.spp if spp is None else spp
        offset = self.samples if offset is None else offset
        p = self.scene.launch_params(offset=offset, count=spp, **kw)
        self.ctx.render(p)
        self.samples = offset + spp
        return self.ctx.stats()

    def accum(self):
        return self.ctx.read_accum()

    def xyz(self):
        a = self.accum()
        return a[..., :3] / a[..., 3:4]

    def image(self, exposure=0.0):
        m = np.asarray(self.scene.tables["xyz_to_srgb"], np.float32)
        return self.ctx.resolve(np.ascontiguousarray(m.T), 2.0 ** (exposure - 7.0))

    def close(self):
        self.ctx.close()
