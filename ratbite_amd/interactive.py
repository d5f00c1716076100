"""The progressive / interactive loop around rbh_render, as the reference's
RenderingInterface runs it (src/core/rendering_interface.cpp:1534-1879,
src/core/command.h):

  * a command buffer (a stack: the last command pushed is processed first,
    as CommandBuffer::pullCommand pops) of the reference's command types,
    with the reference's enum values (src/core/command.h:9-43); the light
    commands act on the disk and sphere lights (whose handling the reference
    has commented out, rendering_interface.cpp:1667-1740): each re-uploads
    the shape lights and rebuilds the light tree;
    processing sets restart / re-resolve / rebuild flags, then applies the
    rebuilds once: the light tree (full rebuild), the light records and
    light-tree spatial data after a camera move, the acceleration structure
    (a refit here: one flattened mesh replaces the instance structure), and
    restarts the accumulation (offset 0, next sub-launch of 1 sample);
  * sub-launch sizing (updateSamplingState): in GRADUAL mode the next
    sub-launch holds min(floor(max(3, c)^1.5), 64) samples, c the previous
    sub-launch's count, in IMMEDIATE mode 1 sample, always clipped to the
    samples left;
  * step(): one sub-launch (rbh_render at the current offset) and the
    sampling-state update; finished once the sample count is reached.
"""
import enum
import math

import numpy as np

from . import _lib as L


class Mode(enum.IntEnum):
    IMMEDIATE = 0
    GRADUAL = 1


class Cmd(enum.IntEnum):
    CHANGE_RENDER_MODE = 1
    CHANGE_SAMPLE_COUNT = 2
    CHANGE_PATH_DEPTH = 3
    CHANGE_IMAGE_EXPOSURE = 4
    CHANGE_DEPTH_OF_FIELD_SETTINGS = 5
    CHANGE_RENDER_RESOLUTION = 6
    CHANGE_CAMERA_POS = 7
    CHANGE_CAMERA_ORIENTATION = 8
    CHANGE_CAMERA_FIELD_OF_VIEW = 9
    ADD_ENVIRONMENT_MAP = 10
    REMOVE_ENVIRONMENT_MAP = 11
    ADD_LIGHT = 12
    CHANGE_LIGHT_POSITION = 13
    CHANGE_LIGHT_SIZE = 14
    CHANGE_LIGHT_ORIENTATION = 15
    CHANGE_LIGHT_POWER = 16
    CHANGE_LIGHT_EMISSION_SPECTRUM = 17
    REMOVE_LIGHT = 18
    CHANGE_MODEL_MATERIAL = 20
    CHANGE_MODEL_TRANSFORM = 21


MAX_SUBLAUNCH = 64


def next_sublaunch(mode, current_count, sample_count, next_offset):
    """updateSamplingState: the size of the sub-launch after one of
    current_count samples, given the total sample_count and the offset the
    next sub-launch starts at."""
    if mode == Mode.GRADUAL:
        size = min(int(math.pow(max(3, current_count), 1.5)), MAX_SUBLAUNCH)
    else:
        size = 1
    return min(max(0, sample_count - next_offset), size)


def sublaunch_schedule(mode, sample_count):
    """(offset, count) of every sub-launch of a restarted render."""
    out, offset, count = [], 0, 1
    count = min(count, sample_count)
    while count > 0:
        out.append((offset, count))
        offset += count
        count = next_sublaunch(mode, count, sample_count, offset)
    return out


class RenderSession:
    """One scene on one context rendered progressively. `renderer` is a
    ratbite_amd.renderer.Renderer; the camera lives in the scene's launch
    parameters (scene.camera: pos, u, v, w, fov)."""

    def __init__(self, renderer, sample_count=None, mode=Mode.GRADUAL, exposure=0.0):
        self.r = renderer
        self.scene = renderer.scene
        self.mode = Mode(mode)
        self.sample_count = int(sample_count if sample_count is not None else self.scene.spp)
        self.exposure = exposure
        self.depth = (self.scene.depth, self.scene.depth, self.scene.depth)
        self.dof = (0, 0.0, 1.0)
        self.commands = []
        self.restart()
        self.image = None

    # ------------------------------------------------------------- commands
    def push(self, cmd, payload=None):
        self.commands.append((Cmd(cmd), payload))

    def restart(self):
        self.offset = 0
        self.count = min(1, self.sample_count)
        self.processed = 0
        self.finished = self.sample_count == 0

    def process_commands(self):
        """Pops every command (last in, first out), then applies the
        rebuilds the commands asked for once, in the reference's order."""
        restart = reresolve = rebuild_tree = refit = translate = reshape = False
        delta = np.zeros(3)
        while self.commands:
            cmd, payload = self.commands.pop()
            if cmd == Cmd.CHANGE_RENDER_MODE:
                self.mode = Mode(payload)
                restart = True
            elif cmd == Cmd.CHANGE_SAMPLE_COUNT:
                self.sample_count = int(payload)
                restart = True
            elif cmd == Cmd.CHANGE_PATH_DEPTH:
                d, refl, trans = payload
                d = max(1, int(d))
                self.depth = (d, min(d, int(refl)), min(d, int(trans)))
                restart = True
            elif cmd == Cmd.CHANGE_IMAGE_EXPOSURE:
                self.exposure = float(payload)
                reresolve = True
            elif cmd == Cmd.CHANGE_DEPTH_OF_FIELD_SETTINGS:
                self.dof = tuple(payload)
                restart = True
            elif cmd == Cmd.CHANGE_RENDER_RESOLUTION:
                self.scene.width, self.scene.height = int(payload[0]), int(payload[1])
                restart = True
            elif cmd == Cmd.CHANGE_CAMERA_POS:
                # payload: the new camera position (world space)
                new = np.asarray(payload, np.float64)
                delta += new - np.asarray(self.scene.camera["pos"], np.float64)
                self.scene.camera["pos"] = new
                translate = restart = True
            elif cmd == Cmd.CHANGE_CAMERA_ORIENTATION:
                u, v, w = payload
                self.scene.camera["u"], self.scene.camera["v"], self.scene.camera["w"] = (
                    np.asarray(u, np.float32), np.asarray(v, np.float32), np.asarray(w, np.float32))
                restart = True
            elif cmd == Cmd.CHANGE_CAMERA_FIELD_OF_VIEW:
                self.scene.camera["fov"] = float(payload)
                restart = True
            elif cmd == Cmd.ADD_ENVIRONMENT_MAP:
                self.scene.env = payload
                self.r.ctx.envmap_build(payload)
                restart = True
            elif cmd == Cmd.REMOVE_ENVIRONMENT_MAP:
                self.scene.env = None
                self.r.ctx.envmap_disable()
                restart = True
            elif cmd == Cmd.CHANGE_LIGHT_POWER:
                if isinstance(payload, tuple):
                    # (shape light, new power scale)
                    i, power = payload
                    self._shapes()[int(i)]["powerScale"] = float(power)
                    reshape = True
                else:
                    # new flux per emissive triangle
                    self.scene.flux = np.ascontiguousarray(payload, np.float32)
                rebuild_tree = restart = True
            elif cmd == Cmd.ADD_LIGHT:
                # payload: a shape-light record (camera-relative)
                rec = np.asarray(payload, L.SHAPE_LIGHT_DTYPE).reshape(1)
                sh = self._shapes()
                self.scene.shapes = np.concatenate([sh, rec]) if sh is not None else rec.copy()
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.REMOVE_LIGHT:
                self.scene.shapes = np.delete(self._shapes(), int(payload))
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.CHANGE_LIGHT_POSITION:
                # (shape light, new camera-relative centre)
                i, pos = payload
                self._shapes()[int(i)]["position"] = np.asarray(pos, np.float32)
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.CHANGE_LIGHT_SIZE:
                i, radius = payload
                self._shapes()[int(i)]["radius"] = float(radius)
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.CHANGE_LIGHT_ORIENTATION:
                # (disk light, new normal): the disk frame of scene.h:386-392
                from .scenes import disk_light
                i, normal = payload
                r = self._shapes()[int(i)]
                r["frame"] = disk_light(r["position"], normal, r["radius"], r["powerScale"],
                                        r["materialIndex"])["frame"]
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.CHANGE_LIGHT_EMISSION_SPECTRUM:
                # (shape light, spectrum index): its material emits that spectrum
                i, spec = payload
                mi = int(self._shapes()[int(i)]["materialIndex"])
                m = self.scene.materials[mi:mi + 1].copy()
                m["flags"] |= 4
                m["emissionSpectrum"] = int(spec)
                self.scene.materials[mi:mi + 1] = m
                self.r.ctx.update_materials(mi, m)
                reshape = rebuild_tree = restart = True
            elif cmd == Cmd.CHANGE_MODEL_MATERIAL:
                first, mats = payload
                self.scene.materials[first:first + len(mats)] = mats
                self.r.ctx.update_materials(first, mats)
                restart = True
            elif cmd == Cmd.CHANGE_MODEL_TRANSFORM:
                # payload: (first vertex, new camera-relative positions)
                first, verts = payload
                verts = np.asarray(verts, np.float32).reshape(-1, 3)
                self.scene.verts[first:first + len(verts)] = verts
                self.r.ctx.update_vertices(first, verts)
                lights = self.scene.lights
                if len(lights):
                    # emitters of the moved vertices follow them
                    tri = self.scene.indices[lights["primitiveDataIndex"]]
                    moved = ((tri >= first) & (tri < first + len(verts))).any(axis=1)
                    if moved.any():
                        lights["vertices"] = self.scene.verts[tri].reshape(-1, 9)
                        rebuild_tree = True
                refit = restart = True
            else:
                raise ValueError("unknown command %r" % (cmd,))
        if reshape:
            from .scenes import shape_flux
            sh = self._shapes()
            if sh is None or len(sh) == 0:
                self.scene.shapes = self.scene.shape_flux = None
                self.r.ctx.upload_shape_lights(np.zeros(0, L.SHAPE_LIGHT_DTYPE), np.zeros(0, np.float32))
            else:
                self.scene.shape_flux = shape_flux(sh, self.scene.materials)
                self.r.ctx.upload_shape_lights(sh, self.scene.shape_flux)
        if rebuild_tree:
            self.r.ctx.build_light_tree(self.scene.lights, self.scene.flux)
        if translate:
            d = delta.astype(np.float32)
            self.r.ctx.translate_scene(d)                     # light records, tree means, vertices + refit
            self.scene.verts = (self.scene.verts - d).astype(np.float32)
            self.scene.lights["vertices"] = (self.scene.lights["vertices"].reshape(-1, 3, 3) - d).reshape(-1, 9)
            if self.scene.nshapes:
                self.scene.shapes["position"] -= d
        elif refit:
            self.r.ctx.refit_accel()
        if restart:
            self.restart()
        elif reresolve and self.processed:
            self.image = self.resolve()
        return restart

    def _shapes(self):
        return self.scene.shapes

    # ------------------------------------------------------------ rendering
    def launch_params(self):
        p = self.scene.launch_params(offset=self.offset, count=self.count)
        p.maxPathDepth, p.maxReflectedPathDepth, p.maxTransmittedPathDepth = self.depth
        p.depthOfField, p.aperture, p.focusDistance = int(self.dof[0]), float(self.dof[1]), float(self.dof[2])
        return p

    def step(self):
        """One sub-launch; returns True once the sample count is reached."""
        self.process_commands()
        if self.finished:
            return True
        self.r.ctx.render(self.launch_params())
        self.processed = self.offset + self.count
        self.offset += self.count
        self.count = next_sublaunch(self.mode, self.count, self.sample_count, self.offset)
        self.finished = self.count == 0
        return self.finished

    def run(self):
        while not self.step():
            pass
        return self.processed

    def resolve(self):
        xyz_to_srgb = np.asarray(self.scene.tables["xyz_to_srgb"], np.float32).reshape(3, 3)
        return self.r.ctx.resolve(xyz_to_srgb, float(2.0 ** self.exposure))
