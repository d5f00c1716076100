"""ctypes binding of libratbite_hip.so (include/ratbite_hip.h).

The library is the product: every compute call goes through it. Loading fails
loudly when it has not been built; there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libratbite_hip.so")

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_uint32 = ctypes.c_uint32
c_uint64 = ctypes.c_uint64
c_float = ctypes.c_float
c_double = ctypes.c_double

RBH_OK = 0
RBH_ERR_INVALID = -1
RBH_ERR_HIP = -2
RBH_ERR_STATE = -3
RBH_ERR_RANGE = -4
RBH_ERR_NOMEM = -5

RBH_RENDER_GUIDED = 1
RBH_RENDER_COLLECT = 2
RBH_RENDER_GUIDE_FP8 = 4
RBH_RENDER_TIMING = 8
RBH_RENDER_NO_DROP = 16
RBH_CDF_FP64 = 1
RBH_ENVMAP_NOALLOC_MAX_ROWS = 32768
RBH_LIGHT_TREE_FP64 = 1

# Record layouts (byte-identical to include/ratbite_hip.h)
MATERIAL_DTYPE = np.dtype([
    ("bxdf", "<u4"), ("flags", "<u4"), ("attributes", "<u4"), ("attributeStride", "<u4"),
    ("normalOffset", "<u4"), ("frameOffset", "<u4"), ("texCoord1Offset", "<u4"), ("texCoord2Offset", "<u4"),
    ("iorSpectrum", "<u4"), ("acSpectrum", "<u4"), ("emissionSpectrum", "<u4"),
    ("mfRoughness", "<f4"), ("ior", "<f4"), ("alphaCutoff", "<f4"),
    ("textures", "<u4"), ("texCoordSets", "<u4"), ("tex", "<i4", (7,)), ("factors", "<u4"),
    ("baseColor", "<f4", (4,)), ("metalness", "<f4"), ("roughness", "<f4"), ("transmission", "<f4"),
    ("sheenColor", "<f4", (3,)), ("sheenRoughness", "<f4"), ("emissive", "<f4", (3,)),
])
assert MATERIAL_DTYPE.itemsize == 152

TEXTURE_DTYPE = np.dtype([
    ("width", "<u4"), ("height", "<u4"), ("depth", "<u4"), ("format", "<u4"), ("srgb", "<u4"),
    ("addr", "<u4", (3,)), ("filter", "<u4"), ("pad", "<u4"), ("offset", "<u8"),
])
assert TEXTURE_DTYPE.itemsize == 48

EMISSIVE_TRI_DTYPE = np.dtype([
    ("vertices", "<f4", (9,)), ("uvs", "<f4", (6,)), ("materialIndex", "<u4"), ("primitiveDataIndex", "<u4"),
])
assert EMISSIVE_TRI_DTYPE.itemsize == 68

# rbh_shape_light: a disk (type 0) or sphere (type 1) light, 48 bytes
SHAPE_LIGHT_DTYPE = np.dtype([
    ("position", "<f4", (3,)), ("powerScale", "<f4"), ("frame", "<f4", (4,)), ("radius", "<f4"),
    ("materialIndex", "<u4"), ("type", "<u4"), ("pad", "<u4"),
])
assert SHAPE_LIGHT_DTYPE.itemsize == 48
SHAPE_DISK, SHAPE_SPHERE = 0, 1

PACKED_NODE_DTYPE = np.dtype([
    ("spatialMean", "<f4", (3,)), ("spatialVariance", "<f4"), ("packedAverageDirection", "<u4"),
    ("sharpness", "<f4"), ("flux", "<f4"), ("coreData", "<u4"),
])
assert PACKED_NODE_DTYPE.itemsize == 32


class LaunchParams(ctypes.Structure):
    _fields_ = [
        ("width", c_uint32), ("height", c_uint32),
        ("maxPathDepth", c_uint32), ("maxReflectedPathDepth", c_uint32), ("maxTransmittedPathDepth", c_uint32),
        ("sampleOffset", c_uint32), ("sampleCount", c_uint32),
        ("camU", c_float * 3), ("camV", c_float * 3), ("camW", c_float * 3),
        ("perspectiveScaleW", c_float), ("perspectiveScaleH", c_float),
        ("depthOfField", c_uint32), ("aperture", c_float), ("focusDistance", c_float),
        ("flags", c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("closestRays", c_uint64), ("shadowRays", c_uint64), ("pathVertices", c_uint64), ("samples", c_uint64),
        ("renderMs", c_double), ("traceMs", c_double), ("shadowMs", c_double), ("shadeMs", c_double),
        ("cdfMs", c_double), ("bvhMs", c_double), ("traceLaunches", c_uint32), ("bvhDepth", c_uint32),
        ("guideMs", c_double), ("nonFiniteSamples", c_uint64),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("pathSlots", c_uint64), ("traceBlocksPerCU", ctypes.c_int32), ("shadeBlocksPerCU", ctypes.c_int32),
                ("useGraph", ctypes.c_int32), ("shadeWaves", ctypes.c_int32), ("traceFetch", ctypes.c_int32), ("traceWaves", ctypes.c_int32),
                ("quantNodes", ctypes.c_int32), ("sortHits", ctypes.c_int32),
                ("samplesPerSecond", c_double)]


class GuideConfig(ctypes.Structure):
    _fields_ = [
        ("lr", c_float), ("beta1", c_float), ("beta2", c_float), ("eps", c_float),
        ("sceneMin", c_float * 3), ("sceneMax", c_float * 3), ("seed", c_uint32), ("recordCapacity", c_uint32),
        ("fixedSelection", c_uint32),
    ]


GUIDE_PARAMS = 20329
GUIDE_REC = 16
GUIDE_OUTPUTS = 41
RBH_INDEX_U32 = 0
RBH_INDEX_U16 = 1


# name -> (restype, argtypes)
_SIGS = {
    "rbh_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "rbh_destroy": (c_int, [c_void_p]),
    "rbh_last_error": (ctypes.c_char_p, [c_void_p]),
    "rbh_query": (c_int, [c_void_p]),
    "rbh_synchronize": (c_int, [c_void_p]),
    "rbh_upload_spectra": (c_int, [c_void_p, c_void_p, c_uint32]),
    "rbh_upload_color_tables": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rbh_upload_luts": (c_int, [c_void_p] + [c_void_p] * 6),
    "rbh_upload_textures": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint64]),
    "rbh_upload_materials": (c_int, [c_void_p, c_void_p, c_uint32]),
    "rbh_upload_geometry": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_uint64, c_void_p, c_void_p]),
    "rbh_build_accel": (c_int, [c_void_p]),
    "rbh_upload_geometry_ex": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, c_void_p, c_uint64,
                                       c_void_p, c_void_p]),
    "rbh_set_light_tree": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_build_light_tree": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_build_light_tree_ex": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32]),
    "rbh_light_tree_build_host": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p]),
    "rbh_light_tree_build_host_ex": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_emissive_flux": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32,
                                  c_uint32, c_void_p]),
    "rbh_rasterise_uv_triangle": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p]),
    "rbh_update_vertices": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p]),
    "rbh_refit_accel": (c_int, [c_void_p]),
    "rbh_translate_scene": (c_int, [c_void_p, c_void_p]),
    "rbh_update_materials": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p]),
    "rbh_get_light_tree": (c_int, [c_void_p, c_void_p, ctypes.POINTER(c_uint32), c_void_p, c_void_p]),
    "rbh_upload_shape_lights": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_light_tree_build_host_shapes": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32,
                                                 c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_envmap_build": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_uint32]),
    "rbh_envmap_rebuild": (c_int, [c_void_p, c_uint32, c_int]),
    "rbh_envmap_read": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double)]),
    "rbh_envmap_disable": (c_int, [c_void_p]),
    "rbh_envmap_build_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_uint32, c_void_p, c_void_p, c_void_p,
                                         c_void_p]),
    "rbh_render": (c_int, [c_void_p, ctypes.POINTER(LaunchParams)]),
    "rbh_set_path_slots": (c_int, [c_void_p, c_uint64]),
    "rbh_render_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p]),
    "rbh_autotune": (c_int, [c_void_p, ctypes.POINTER(LaunchParams), c_uint32, ctypes.POINTER(Tuning)]),
    "rbh_set_tuning": (c_int, [c_void_p, ctypes.POINTER(Tuning)]),
    "rbh_get_tuning": (c_int, [c_void_p, ctypes.POINTER(Tuning)]),
    "rbh_guide_configure": (c_int, [c_void_p, ctypes.POINTER(GuideConfig)]),
    "rbh_guide_set_params": (c_int, [c_void_p, c_void_p, c_uint32]),
    "rbh_guide_get_params": (c_int, [c_void_p, c_void_p, c_uint32]),
    "rbh_guide_gradient": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "rbh_guide_train": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, ctypes.POINTER(ctypes.c_float)]),
    "rbh_guide_infer": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "rbh_guide_infer_fp8": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "rbh_guide_records": (c_int, [c_void_p, c_void_p, c_uint32, ctypes.POINTER(c_uint32)]),
    "rbh_guide_train_collected": (c_int, [c_void_p, c_uint32, c_uint32, ctypes.POINTER(ctypes.c_float)]),
    "rbh_read_accum": (c_int, [c_void_p, c_void_p]),
    "rbh_write_accum": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32]),
    "rbh_resolve": (c_int, [c_void_p, c_void_p, c_float, c_void_p]),
    "rbh_stats_get": (c_int, [c_void_p, ctypes.POINTER(Stats)]),
}

# Optional groups, declared in the other include/ headers
_OPTIONAL_SIGS = {}

_lib = None


def declared_symbols():
    return list(_SIGS) + list(_OPTIONAL_SIGS)


def register(sigs):
    """Adds signatures of another header's entry points."""
    _OPTIONAL_SIGS.update(sigs)
    if _lib is not None:
        _apply(_lib, sigs)


MISSING = []


def _apply(lib, sigs):
    for name, (res, args) in sigs.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            MISSING.append(name)
            continue
        fn.restype = res
        fn.argtypes = args


def lib():
    """Loads libratbite_hip.so; raises when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libratbite_hip.so is not built (run __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        _apply(l, _SIGS)
        _apply(l, _OPTIONAL_SIGS)
        _lib = l
    return _lib


def ptr(a):
    """Raw data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(c_void_p)


class RbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rbh error %d: %s" % (code, msg))
        self.code = code


# include/ratbite_hip_test.h
register({
    "rbh_test_sobol": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_pixel_hash": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_wavelengths": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_rgb_to_spectrum": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_oct": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_copy": (c_int, [c_void_p, c_uint64, c_int, ctypes.POINTER(ctypes.c_float)]),
    "rbh_test_lbvh_read": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_uint32)]),
    "rbh_test_radix_sort": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_int, c_int, c_void_p, c_void_p]),
    "rbh_test_select_positive": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, ctypes.POINTER(c_uint32)]),
    "rbh_test_trace": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_int]),
    "rbh_test_light_pdf": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_light_sample": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32]),
    "rbh_test_bxdf": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32]),
})
