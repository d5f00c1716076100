"""Host-side renderer over one rbh context.

Mirrors the orchestration of the reference's RenderingInterface
(src/core/rendering_interface.cpp): upload spectra, colour tables, LUTs,
textures, materials and geometry; build the acceleration structure, the light
tree (buildLightTree, 1065-1186) and the env-map tables (loadEnvironmentMap,
1339-1432); then render sub-launches into the FP64 accumulation buffer and
resolve (renderResolve, 1512-1533).
"""
import ctypes

import numpy as np

from . import _lib as L
from .context import Context


class Renderer:
    def __init__(self, scene, device=0, ctx=None, fp64_cdf=False, fp64_light_tree=False):
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
        # (an empty set too: a reused context drops the previous scene's shape lights)
        if scene.nshapes:
            c.upload_shape_lights(scene.shapes, scene.shape_flux)
        else:
            c.upload_shape_lights(np.zeros(0, L.SHAPE_LIGHT_DTYPE), np.zeros(0, np.float32))
        c.build_light_tree(scene.lights, scene.flux, fp64=fp64_light_tree)
        if scene.env is not None:
            c.envmap_build(scene.env, fp64=fp64_cdf)
        else:
            c.envmap_disable()
        self.samples = 0

    def light_tree(self):
        return self.ctx.get_light_tree(len(self.scene.lights) + self.scene.nshapes)

    def render(self, spp=None, offset=None, guided=False, collect=False, no_drop=False, flags=0, fresh=False, **kw):
        """Renders spp samples starting at Sobol index `offset`. A launch with
        offset != 0 continues the accumulation (the reference's progressive
        sub-launches); fresh=True clears it first, for an independent image
        from a later part of the sample sequence."""
        spp = self.scene.spp if spp is None else spp
        offset = self.samples if offset is None else offset
        if fresh and offset != 0:
            self.ctx.write_accum(np.zeros((self.scene.height, self.scene.width, 4), np.float64))
        flags |= (L.RBH_RENDER_GUIDED if guided else 0) | (L.RBH_RENDER_COLLECT if collect else 0)
        flags |= L.RBH_RENDER_NO_DROP if no_drop else 0
        p = self.scene.launch_params(offset=offset, count=spp, flags=flags, **kw)
        self.ctx.render(p)
        self.samples = offset + spp
        return self.ctx.stats()

    def autotune(self, trial_spp=4, **kw):
        """Picks slots per pass and persistent grid sizes by timing short renders."""
        t = L.Tuning()
        p = self.scene.launch_params(offset=0, count=trial_spp, **kw)
        self.ctx.call("rbh_autotune", ctypes.byref(p), trial_spp, ctypes.byref(t))
        self.samples = 0
        return {"pathSlots": t.pathSlots, "traceBlocksPerCU": t.traceBlocksPerCU,
                "shadeBlocksPerCU": t.shadeBlocksPerCU, "useGraph": t.useGraph, "shadeWaves": t.shadeWaves,
                "traceFetch": t.traceFetch, "traceWaves": t.traceWaves, "quantNodes": t.quantNodes,
                "sortHits": t.sortHits, "samplesPerSecond": t.samplesPerSecond}

    def tuning(self):
        """The launch shape this context renders with."""
        t = L.Tuning()
        self.ctx.call("rbh_get_tuning", ctypes.byref(t))
        return {"pathSlots": t.pathSlots, "traceBlocksPerCU": t.traceBlocksPerCU,
                "shadeBlocksPerCU": t.shadeBlocksPerCU, "useGraph": t.useGraph, "shadeWaves": t.shadeWaves,
                "traceFetch": t.traceFetch, "traceWaves": t.traceWaves, "quantNodes": t.quantNodes,
                "sortHits": t.sortHits}

    def set_tuning(self, path_slots=0, trace_blocks=256, shade_blocks=256, use_graph=False, shade_waves=4,
                   trace_fetch=0, trace_waves=7, quant_nodes=False, sort_hits=True):
        """Sets the launch shape; the defaults are the library's own."""
        t = L.Tuning(path_slots, trace_blocks, shade_blocks, int(use_graph), shade_waves, int(trace_fetch),
                     trace_waves, int(quant_nodes), int(sort_hits), 0.0)
        self.ctx.call("rbh_set_tuning", ctypes.byref(t))

    # ------------------------------------------------------------ guiding
    def configure_guide(self, seed=1, lr=1e-3, capacity=1 << 22, fixed_selection=False):
        """Sets up the guiding network (random initialisation) with the
        scene's bounding box as the position normalisation. With
        fixed_selection the guided render draws from the lobes with
        probability 0.5 and only the 40 lobe outputs are trained."""
        v = np.asarray(self.scene.verts, np.float64)
        lo, hi = v.min(axis=0), v.max(axis=0)
        pad = 1e-3 * max(1e-6, float((hi - lo).max()))
        cfg = L.GuideConfig()
        cfg.lr, cfg.beta1, cfg.beta2, cfg.eps = lr, 0.9, 0.999, 1e-8
        cfg.sceneMin[:] = [float(x) for x in lo - pad]
        cfg.sceneMax[:] = [float(x) for x in hi + pad]
        cfg.seed = seed
        cfg.recordCapacity = capacity
        cfg.fixedSelection = int(bool(fixed_selection))
        self.ctx.call("rbh_guide_configure", ctypes.byref(cfg))

    def train_guide(self, batch=1 << 18, steps=1):
        loss = ctypes.c_float()
        self.ctx.call("rbh_guide_train_collected", batch, steps, ctypes.byref(loss))
        return loss.value

    def guide_records(self, max_records=None):
        n = ctypes.c_uint32()
        self.ctx.call("rbh_guide_records", None, 0, ctypes.byref(n))
        k = n.value if max_records is None else min(n.value, max_records)
        out = np.zeros((k, L.GUIDE_REC), np.float32)
        if k:
            self.ctx.call("rbh_guide_records", L.ptr(out), k, ctypes.byref(n))
        return out

    def accum(self):
        return self.ctx.read_accum()

    def xyz(self):
        a = self.accum()
        return a[..., :3] / a[..., 3:4]

    def image(self, exposure=0.0):
        m = np.asarray(self.scene.tables["xyz_to_srgb"], np.float32)
        return self.ctx.resolve(np.ascontiguousarray(m.T), 2.0 ** (exposure - 7.0))

    def save_png(self, path, exposure=0.0):
        """Resolves the accumulation and writes it as a PNG, bottom film row
        last (src/core/ui.cpp:894-912)."""
        from .export import write_png
        write_png(path, self.image(exposure))

    def close(self):
        self.ctx.close()


def render_batch(renderers, spp=None, flags=0):
    """Renders several scenes (one Renderer each, same device) in one batched
    launch sequence; the first renderer's context owns the pass state.
    Returns the per-scene completion times in ms."""
    n = len(renderers)
    ctxs = (ctypes.c_void_p * n)(*[r.ctx.h for r in renderers])
    lps = (L.LaunchParams * n)()
    for i, r in enumerate(renderers):
        lps[i] = r.scene.launch_params(offset=0, count=r.scene.spp if spp is None else spp, flags=flags)
        r.ctx.width, r.ctx.height = lps[i].width, lps[i].height
        r.samples = lps[i].sampleCount
    done = np.zeros(n, np.float64)
    owner = renderers[0].ctx
    owner.call("rbh_render_batch", ctypes.cast(ctxs, ctypes.c_void_p), ctypes.cast(lps, ctypes.c_void_p), n, L.ptr(done))
    return done
