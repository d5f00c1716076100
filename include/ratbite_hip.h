/*
 * ratbite_hip.h — C ABI of the MI355X (gfx950) spectral path-tracing hot path.
 *
 * The reference drives its device code through the C++ class
 * RenderingInterface (src/core/rendering_interface.h:19-226) with a
 * LaunchParameters block (src/core/launch_parameters.h:16-96), OptiX program
 * groups and three PTX kernels. This header replaces that boundary with plain
 * C functions over plain pointers and sizes. Every entry point names the
 * reference interface it replaces.
 *
 * Conventions
 *  - Every function returns an int: RBH_OK (0) or a negative RBH_ERR_* code;
 *    rbh_last_error(ctx) then holds a message. Nothing aborts the process.
 *  - One host thread per context; calls are stream ordered on the context's
 *    stream. Host buffers are caller owned and copied during the call.
 *    Device memory is context owned.
 *  - Geometry and lights are in camera-relative world space (the reference
 *    translates instances, light triangles and tree nodes by -cameraPos on
 *    the host, rendering_interface.cpp:986-994, 1116-1125, 1449-1457,
 *    1484-1486).
 */
#ifndef RATBITE_HIP_H
#define RATBITE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbh_context* rbh_ctx;

enum {
    RBH_OK = 0,
    RBH_ERR_INVALID = -1,   /* bad argument / size / layout */
    RBH_ERR_HIP = -2,       /* HIP runtime failure */
    RBH_ERR_STATE = -3,     /* call out of order (e.g. render before geometry) */
    RBH_ERR_RANGE = -4,     /* size above a documented limit */
    RBH_ERR_NOMEM = -5,     /* allocation failed */
};

/* ------------------------------------------------------------------ records
 * Byte layouts are fixed (asserted in the implementation) so that numpy
 * structured arrays and ctypes can fill them directly. */

/* MaterialData (src/core/material.h:17-107) without raw device pointers:
 * index/attribute buffers are the context's geometry buffers, textures are
 * indices into the rbh_upload_textures table. 152 bytes. */
typedef struct rbh_material {
    uint32_t bxdf;            /* bxdfIndexSBT: 0 PureConductor, 1 PureDielectric, 2 ComplexSurface */
    uint32_t flags;           /* 1 doubleSided, 2 sheenPresent, 4 spectral emission (see below) */
    uint32_t attributes;      /* AttributeTypeBitfield: 1 NORMAL, 2 FRAME, 4 COLOR, 8 TEX_COORD_1, 16 TEX_COORD_2 */
    uint32_t attributeStride; /* bytes per vertex in the attribute buffer */
    uint32_t normalOffset, frameOffset, texCoord1Offset, texCoord2Offset;
    uint32_t iorSpectrum;     /* indexOfRefractSpectrumDataIndex */
    uint32_t acSpectrum;      /* absorpCoefSpectrumDataIndex */
    uint32_t emissionSpectrum;/* emissionSpectrumDataIndex */
    float mfRoughness, ior, alphaCutoff;
    uint32_t textures;        /* TextureTypeBitfield */
    uint32_t texCoordSets;    /* bit i: texture slot i reads TEX_COORD_2 (the *TexCoordSetIndex flags) */
    int32_t tex[7];           /* baseColor, normal, metRough, transmission, sheenColor, sheenRough, emission */
    uint32_t factors;         /* FactorTypeBitfield */
    float baseColor[4];
    float metalness, roughness, transmission;
    float sheenColor[3];
    float sheenRoughness;
    float emissive[3];        /* emissiveFactor; with flag 4: emissive[0] scales spectra[emissionSpectrum] */
} rbh_material;

/* A texture (replaces CudaCombinedTexture, src/core/texture.h:237-500).
 * Filtering is done in software with exact FP32 weights. 48 bytes. */
typedef struct rbh_texture_desc {
    uint32_t width, height, depth;
    uint32_t format;          /* 0 RGBA8, 1 R8, 2 RGBA32F, 3 RGB32F, 4 R32F */
    uint32_t srgb;            /* decode RGB from sRGB before filtering */
    uint32_t addr[3];         /* 0 wrap, 1 clamp, 2 mirror */
    uint32_t filter;          /* 0 nearest, 1 linear */
    uint32_t pad;
    uint64_t offset;          /* byte offset of the texels in the blob */
} rbh_texture_desc;

/* EmissiveTriangleLightData (src/core/light_tree_types.h:40-46), 68 bytes. */
typedef struct rbh_emissive_tri {
    float vertices[9];
    float uvs[6];
    uint32_t materialIndex;
    uint32_t primitiveDataIndex;
} rbh_emissive_tri;

/* DiskLightData / SphereLightData (src/core/light_tree_types.h:24-39,
 * src/core/scene.h:319-492) as one 48-byte record. The reference declares
 * these lights but leaves their programs commented out
 * (src/kernels/pathtrace.cu:150-228). Here a disk emits from the side its
 * normal (the frame's third column) faces and occludes from both; a sphere
 * emits outwards and gives nothing to points inside it; a path that hits
 * either ends there (pure emitter). Emission = the material's emission
 * (emissionSpectrum x emissive[0] with material flag 4, else the RGB
 * emissive colour; no texture) x powerScale. */
typedef struct rbh_shape_light {
    float position[3];
    float powerScale;
    float frame[4];           /* unit quaternion x, y, z, w (the vertex-frame order) */
    float radius;
    uint32_t materialIndex;
    uint32_t type;            /* RBH_SHAPE_DISK or RBH_SHAPE_SPHERE */
    uint32_t pad;
} rbh_shape_light;
enum { RBH_SHAPE_DISK = 0, RBH_SHAPE_SPHERE = 1 };

/* LightTree::PackedNode (src/core/light_tree_types.h:119-162), 32 bytes. */
typedef struct rbh_packed_node {
    float spatialMean[3];
    float spatialVariance;
    uint32_t packedAverageDirection;
    float sharpness;
    float flux;
    uint32_t coreData;
} rbh_packed_node;

/* LaunchParameters' per-launch state (launch_parameters.h:18-50). */
typedef struct rbh_launch_params {
    uint32_t width, height;                      /* ResolutionState */
    uint32_t maxPathDepth, maxReflectedPathDepth, maxTransmittedPathDepth; /* PathState */
    uint32_t sampleOffset, sampleCount;          /* SamplingState */
    float camU[3], camV[3], camW[3];             /* CameraState */
    float perspectiveScaleW, perspectiveScaleH;
    uint32_t depthOfField;
    float aperture, focusDistance;
    uint32_t flags;                              /* RBH_RENDER_* */
} rbh_launch_params;

enum {
    RBH_RENDER_GUIDED = 1,        /* one-sample MIS with the guiding MLP (rbh_guide_configure first) */
    RBH_RENDER_COLLECT = 2,       /* with GUIDED: record training samples (rbh_guide_train_collected) */
    RBH_RENDER_GUIDE_FP8 = 4,     /* FP8 E4M3 guiding inference */
    RBH_RENDER_TIMING = 8,        /* per-kernel HIP-event timing into rbh_stats */
    RBH_RENDER_NO_DROP = 16,      /* an invalid BxDF sample ends the path (FINISHED) instead of
                                     discarding the whole sample as the reference does (implied by GUIDED) */
};

enum {
    RBH_CDF_FP64 = 1,             /* FP64 prefix sums and normalisation */
};

typedef struct rbh_stats {
    uint64_t closestRays;         /* closest-hit queries of the last rbh_render */
    uint64_t shadowRays;          /* any-hit queries issued (occluded-flag culled rays excluded) */
    uint64_t pathVertices;        /* shaded path vertices */
    uint64_t samples;             /* W*H*sampleCount */
    double renderMs;              /* device time of the last rbh_render */
    double traceMs;               /* time in closest-hit kernels */
    double shadowMs;              /* time in any-hit kernels */
    double shadeMs;               /* time in shading kernels */
    double cdfMs;                 /* time of the last env-map CDF build kernel */
    double bvhMs;                 /* time of the last LBVH build */
    uint32_t traceLaunches;       /* closest-hit kernel launches of the last render */
    uint32_t bvhDepth;            /* max depth of the LBVH */
    double guideMs;               /* time in guiding inference kernels */
    uint64_t nonFiniteSamples;    /* samples whose radiance was not finite (added as black; the reference would
                                     accumulate the NaN/inf into the pixel) */
} rbh_stats;

/* ----------------------------------------------------------------- context
 * Replaces RenderingInterface::RenderingInterface / cleanup
 * (rendering_interface.cpp:1952-1978, 1880-1950). */
int rbh_create(int device, rbh_ctx* out);
int rbh_destroy(rbh_ctx ctx);
const char* rbh_last_error(rbh_ctx ctx);
/* cudaEventQuery polling (rendering_interface.cpp:1982-1986): 1 idle, 0 busy */
int rbh_query(rbh_ctx ctx);
int rbh_synchronize(rbh_ctx ctx);

/* ----------------------------------------------------------- scene uploads */
/* LaunchParameters::spectra (n dense 471-sample spectra, spectral.cpp:1277-1286) */
int rbh_upload_spectra(rbh_ctx ctx, const float* dense, uint32_t n);
/* sensorSpectralCurveA/B/C and spectralBasisR/G/B
 * (rendering_interface.cpp:518-527): 3 x 471 floats each */
int rbh_upload_color_tables(rbh_ctx ctx, const float* xyz, const float* basis);
/* loadLookUpTables (rendering_interface.cpp:528-622) */
int rbh_upload_luts(rbh_ctx ctx, const float* conductor32x32, const float* dielOuter32x32x32,
                    const float* dielInner32x32x32, const float* reflOuter32x32x32,
                    const float* reflInner32x32x32, const float* sheen32x32x4);
int rbh_upload_textures(rbh_ctx ctx, const rbh_texture_desc* descs, uint32_t n, const void* blob, uint64_t blobBytes);
/* fillModelMaterials / addMaterial (rendering_interface.cpp:47-307, 1187-1290) */
int rbh_upload_materials(rbh_ctx ctx, const rbh_material* mats, uint32_t n);
/* Vertex, index and attribute buffers (replaces the GAS inputs and the
 * per-material index/attribute pointers). triLight[i] is the light index of
 * an emissive triangle (LightPointer TRIANGLE) or 0xFFFFFFFF. */
int rbh_upload_geometry(rbh_ctx ctx, const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntris,
                        const void* attribs, uint64_t attribBytes, const uint32_t* triMaterial,
                        const uint32_t* triLight);
/* The same with 16-bit index buffers as well (the reference's per-geometry
 * index type, read at pathtrace.cu:1295-1299 and set up at
 * rendering_interface.cpp:201-301): indexData holds 3*ntris indices of
 * indexFormat. */
enum {
    RBH_INDEX_U32 = 0,
    RBH_INDEX_U16 = 1,
};
int rbh_upload_geometry_ex(rbh_ctx ctx, const float* verts, uint32_t nverts, const void* indexData, uint32_t indexFormat,
                           uint32_t ntris, const void* attribs, uint64_t attribBytes, const uint32_t* triMaterial,
                           const uint32_t* triLight);
/* optixAccelBuild GAS/IAS (rendering_interface.cpp:733-1064): builds the LBVH */
int rbh_build_accel(rbh_ctx ctx);
/* ------------------------------------------------------ partial updates
 * The interactive loop's partial rebuilds (rendering_interface.cpp:1534-1845:
 * CHANGE_MODEL_TRANSFORM / CHANGE_CAMERA_POS rebuild the instance structure,
 * refill the light records and shift the light-tree spatial data;
 * CHANGE_MODEL_MATERIAL rewrites material records). With one flattened mesh
 * the instance rebuild becomes a refit of the LBVH. */
/* Rewrites count vertex positions from `first` (float3 each). Rendering is
 * refused until rbh_refit_accel or rbh_build_accel has run. */
int rbh_update_vertices(rbh_ctx ctx, uint32_t first, uint32_t count, const float* verts);
/* Recomputes every box, triangle record and four-wide node for the current
 * vertices over the topology of the last rbh_build_accel (same triangles);
 * the traversal stays exact, its cost grows with the drift. */
int rbh_refit_accel(rbh_ctx ctx);
/* Camera-relative coordinates after a camera move by delta (world units):
 * subtracts delta from every vertex, emitter record corner and light-tree
 * spatial mean, then refits (CHANGE_CAMERA_POS). */
int rbh_translate_scene(rbh_ctx ctx, const float* delta);
/* Rewrites count material records from `first`, with rbh_upload_materials'
 * checks (CHANGE_MODEL_MATERIAL). */
int rbh_update_materials(rbh_ctx ctx, uint32_t first, uint32_t count, const rbh_material* mats);

/* ------------------------------------------------------------ light tree */
/* RenderingInterface::buildLightTree + LightTree::Builder::build
 * (rendering_interface.cpp:1065-1186, light_tree.cpp:47-79): builds the tree
 * on the host from n emissive triangles with their flux and the context's
 * shape lights (rbh_upload_shape_lights), uploads nodes, light pointers,
 * bitmasks and the triangle records. */
int rbh_build_light_tree(rbh_ctx ctx, const rbh_emissive_tri* tris, const float* flux, uint32_t n);
enum {
    RBH_LIGHT_TREE_FP64 = 1,      /* FP64 power mode: flux sums, attribute merges and split costs in double */
};
/* rbh_build_light_tree with flags (RBH_LIGHT_TREE_*); flags 0 is the reference's
 * FP32 builder. In the FP64 mode each node's flux is the float rounding of a
 * double sum of its lights' fluxes; the topology may differ from FP32 mode. */
int rbh_build_light_tree_ex(rbh_ctx ctx, const rbh_emissive_tri* tris, const float* flux, uint32_t n,
                            uint32_t flags);
/* The same builder without a context or device (host only): fills nodes
 * (capacity 2n), lightPointers (n) and bitmasks (n); returns the node count
 * or a negative RBH_ERR_* code. */
int rbh_light_tree_build_host(const rbh_emissive_tri* tris, const float* flux, uint32_t n, rbh_packed_node* nodes,
                              uint32_t* lightPointers, uint64_t* bitmasks);
int rbh_light_tree_build_host_ex(const rbh_emissive_tri* tris, const float* flux, uint32_t n, rbh_packed_node* nodes,
                                 uint32_t* lightPointers, uint64_t* bitmasks, uint32_t flags);
/* fillLightTreeDataBuffers (rendering_interface.cpp:1470-1510): uploads a
 * tree the caller built (nnodes packed nodes in pre-order: the left child of
 * an inner node follows it, its right child index is in coreData; leaves of
 * 1..15 lights), n + m light pointers (type 0 for the n triangle lights, type
 * 3 for the m shape lights uploaded with rbh_upload_shape_lights), n + m
 * bitmasks (the path of each light from the root) and the n
 * emissive-triangle records. The tree is checked (every walk ends; every
 * light's bitmask reaches the leaf holding it) before the kernels use it. */
int rbh_set_light_tree(rbh_ctx ctx, const rbh_packed_node* nodes, uint32_t nnodes, const uint32_t* lightPointers,
                       const uint64_t* bitmasks, const rbh_emissive_tri* tris, uint32_t n);
/* Emitted flux of n emissive triangles of one material, as the glTF loader
 * computes it when it cuts emitters out of a mesh (scene.cpp:485-632,
 * sw_rast.h): mean(emissiveFactor) x the mean (r+g+b)a/3 of the RGBA8
 * emission texels whose centres the triangle's UV footprint covers (the
 * texels under its corners when it covers none) x area x pi. verts: n x 9
 * floats (model-space corners), uvs: n x 6 (emission texture coordinates),
 * rgba8: texW x texH texels or NULL for an untextured emitter, addrX/addrY:
 * 0 wrap, 1 clamp, 2 mirror (the loader's mirror rule). Host only. */
int rbh_emissive_flux(const float* verts, const float* uvs, uint32_t n, const float* emissiveFactor,
                      const uint8_t* rgba8, uint32_t texW, uint32_t texH, uint32_t addrX, uint32_t addrY,
                      float* flux);
/* The texel set of that rasterisation for one UV-space triangle (corners in
 * texel units): up to capacity (x, y) pairs, *count = the full count
 * (RBH_ERR_RANGE when it exceeds capacity). Host only. */
int rbh_rasterise_uv_triangle(const float* a, const float* b, const float* c, int* xy, uint32_t capacity,
                              uint32_t* count);
/* Read the built tree back: nodes (<= 2(n+m)), lightPointers (n+m), bitmasks
 * (n+m), for n triangle lights and m shape lights. */
int rbh_get_light_tree(rbh_ctx ctx, rbh_packed_node* nodes, uint32_t* nnodes, uint32_t* lightPointers,
                       uint64_t* bitmasks);
/* Disk and sphere lights (scene.h:319-492; the ADD_LIGHT / CHANGE_LIGHT_*
 * commands of rendering_interface.cpp:1667-1740 re-upload them): replaces
 * the context's m shape lights (m <= 4096; 0 removes them) and their flux
 * (positive; the light tree's importance, e.g. pi x area x emitted radiance).
 * Shape light i has light index n + i after the n triangle lights and light
 * pointers of type 3 with that index. They join the light tree at the next
 * rbh_build_light_tree / rbh_set_light_tree; a render refuses to run until
 * then. Records are checked: finite position and frame, unit quaternion
 * (|q| within 1e-3 of 1), radius > 0, type 0 or 1, materialIndex below the
 * material count. */
int rbh_upload_shape_lights(rbh_ctx ctx, const rbh_shape_light* lights, const float* flux, uint32_t m);
/* rbh_light_tree_build_host_ex over n triangle lights and m shape lights:
 * nodes (capacity 2(n+m)), lightPointers and bitmasks (n+m). Host only. */
int rbh_light_tree_build_host_shapes(const rbh_emissive_tri* tris, const float* flux, uint32_t n,
                                     const rbh_shape_light* shapes, const float* shapeFlux, uint32_t m,
                                     rbh_packed_node* nodes, uint32_t* lightPointers, uint64_t* bitmasks,
                                     uint32_t flags);

/* ---------------------------------------------------------- environment map
 * loadEnvironmentMap + buildCDFLatLongRGBUpTo{4,8,16,32}k + invertCDFToIndices
 * (rendering_interface.cpp:1339-1432, cdf.cu:468-535). texels: w*h*channels
 * floats (channels 3 or 4), w and h multiples of 4, w <= 32768.
 * The integral is read back after a stream sync. */
int rbh_envmap_build(rbh_ctx ctx, const float* texels, int w, int h, int channels, uint32_t flags);
/* Re-run the CDF build on the resident env map (device-resident input). */
int rbh_envmap_rebuild(rbh_ctx ctx, uint32_t flags, int iterations);
int rbh_envmap_read(rbh_ctx ctx, uint16_t* condIdx, uint16_t* margIdx, float* cdfY, double* integral);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The build and inversion of rbh_envmap_build, from and into
 * caller-owned DEVICE buffers: texels (w*h*channels floats), condIdx (w*h),
 * margIdx (h), cdfY (h) and integralOut (one double: the integral divided by
 * w*h, as rbh_envmap_read returns it). The build's scratch (h row terms
 * and the raw integral) is a fixed array of the library on each device, so
 * h <= RBH_ENVMAP_NOALLOC_MAX_ROWS; the call allocates nothing and only
 * launches kernels on the context's stream (no host synchronisation: the
 * outputs are ready when the stream reaches them). Calls that overlap on one
 * GPU (several contexts or threads) race on that array. The context's own
 * env map is not touched. */
enum { RBH_ENVMAP_NOALLOC_MAX_ROWS = 32768 };
int rbh_envmap_build_noalloc(rbh_ctx ctx, const float* texels, int w, int h, int channels, uint32_t flags,
                             uint16_t* condIdx, uint16_t* margIdx, float* cdfY, double* integralOut);
int rbh_envmap_disable(rbh_ctx ctx);

/* ------------------------------------------------------------------ render
 * optixLaunch(__raygen__main) (rendering_interface.cpp:1873-1879): renders
 * sampleCount samples per pixel starting at sampleOffset into the FP64
 * accumulation buffer (re-read when sampleOffset != 0). Returns when the
 * work has finished; rbh_stats_get then describes this launch. */
int rbh_render(rbh_ctx ctx, const rbh_launch_params* p);
/* Path slots of one wavefront pass: a pass renders max(1, slots / (W*H))
 * samples of every pixel at once, about 200 B of device memory per slot.
 * 0 (the default) takes all samples of a launch in one pass when a quarter of
 * the free device memory holds them. Results do not depend on it. No
 * reference counterpart: the OptiX megakernel keeps path state in registers. */
int rbh_set_path_slots(rbh_ctx ctx, uint64_t slots);
/* Many scenes in one wavefront launch sequence (no reference counterpart;
 * the reference renders one scene per optixLaunch). Each scene is a context
 * prepared with the upload/build calls above, all on owner's device; lps[i]
 * are its launch parameters (flags: RBH_RENDER_TIMING, RBH_RENDER_NO_DROP).
 * Every pass holds the next samples of every unfinished scene; each scene
 * accumulates into its own context. completionMs (optional, n entries)
 * receives the time from the start of the call to the end of the pass in
 * which each scene finished. Statistics go to owner. */
int rbh_render_batch(rbh_ctx owner, const rbh_ctx* scenes, const rbh_launch_params* lps, uint32_t n,
                     double* completionMs);
/* Launch shape of the wavefront passes (no reference counterpart). */
typedef struct rbh_tuning {
    uint64_t pathSlots;           /* slots per pass (see rbh_set_path_slots; 0 automatic) */
    int32_t traceBlocksPerCU;     /* persistent grid of the trace / shadow kernels */
    int32_t shadeBlocksPerCU;     /* persistent grid of the shade / miss kernels */
    int32_t useGraph;             /* replay each pass as a captured HIP graph (ignored with RBH_RENDER_TIMING) */
    int32_t shadeWaves;           /* shade-kernel register budget: 1 (unconstrained), 3 or 4 waves per SIMD */
    int32_t traceFetch;           /* 1: a trace/shadow lane takes a new ray as soon as its query completes; 0: waves take 64 rays at a time */
    int32_t traceWaves;           /* trace/shadow register budget (wave-synchronous kernels): 1 (unconstrained), 7 or 8 waves per SIMD */
    int32_t quantNodes;           /* 1: the wave-synchronous trace/shadow kernels read the four-wide nodes quantised to 64 bytes */
    int32_t sortHits;             /* 1: material-sorted shading (hits queued by shading class: plain, transmissive and sheen
                                     ComplexSurface, PureConductor, PureDielectric, shape light) */
    double samplesPerSecond;      /* measured by rbh_autotune */
} rbh_tuning;
/* Times trialSpp-sample renders of lp over a grid of slot budgets and grid
 * sizes, keeps the fastest on the context and returns it. Overwrites the
 * accumulation buffer. */
int rbh_autotune(rbh_ctx ctx, const rbh_launch_params* lp, uint32_t trialSpp, rbh_tuning* best);
int rbh_set_tuning(rbh_ctx ctx, const rbh_tuning* t);
/* The launch shape the context renders with (samplesPerSecond 0). */
int rbh_get_tuning(rbh_ctx ctx, rbh_tuning* t);
/* Copy the FP64 (x, y, z, w) accumulation buffer (W*H*4 doubles). */
int rbh_read_accum(rbh_ctx ctx, double* rgbw);
int rbh_write_accum(rbh_ctx ctx, const double* rgbw, uint32_t width, uint32_t height);
/* renderResolve (image.cu:92-121): colour-space matrix (column-major 3x3),
 * exposure scale 2^(e-7), AgX, RGBA8 out (W*H*4 bytes). */
int rbh_resolve(rbh_ctx ctx, const float* colorspace, float exposureScale, uint8_t* rgba8);
int rbh_stats_get(rbh_ctx ctx, rbh_stats* out);

/* ------------------------------------------------------------- path guiding
 * No reference counterpart: a learned proposal that extends the BSDF
 * sampling of src/kernels/pathtrace.cu:548-905 through one-sample MIS.
 * An MLP 80 -> 64 x4 (ReLU) -> 41 maps a shading point to 8 von Mises-Fisher
 * lobes (per lobe: mean xyz, log sharpness, mixture logit; outputs 0..39) and
 * a selection logit c (output 40): a renderer draws a direction from the lobes
 * with probability a = 0.1 + 0.8 sigmoid(c), otherwise from the BSDF (or with
 * a = 0.5 when the guide is configured with fixedSelection). BF16 weights
 * with FP32 accumulation and FP32 master weights, Adam.
 * Records are RBH_GUIDE_RECORD floats: position normalised to [0,1]^3 by the
 * configured scene bounds [0:3], shading normal [3:6], outgoing direction
 * [6:9], roughness [9], wavelength normalised to [0,1] over 360..830 nm [10]
 * (inference reads these 11), then for training the sampled incident
 * direction [11:14], weight w = f |cos| L_i / pdf of the sampled contribution
 * [14] and the BSDF sampler's density of that direction p_bsdf [15].
 * The loss is sum(w * -log(a q(d) + (1 - a) p_bsdf(d))) / sum(w), q the lobe
 * mixture's density. */
enum {
    RBH_GUIDE_RECORD = 16,        /* floats per record */
    RBH_GUIDE_OUTPUTS = 41,       /* floats per rbh_guide_infer output row */
    RBH_GUIDE_PARAMS = 20329,     /* FP32 master parameters: 20032 weights + 297 biases */
};
typedef struct rbh_guide_config {
    float lr, beta1, beta2, eps;      /* Adam (1e-3, 0.9, 0.999, 1e-8) */
    float sceneMin[3], sceneMax[3];   /* position normalisation */
    uint32_t seed;                    /* He-uniform initialisation */
    uint32_t recordCapacity;          /* records kept per RBH_RENDER_COLLECT render */
    uint32_t fixedSelection;          /* 1: one-sample MIS at a fixed selection probability a = 0.5; the
                                         selection output's weights and bias are held at zero (set_params
                                         zeroes them too), so only the 40 lobe outputs are trained */
} rbh_guide_config;
int rbh_guide_configure(rbh_ctx ctx, const rbh_guide_config* cfg);
/* FP32 master parameters, per layer W[out][in] then b[out]
 * (n = RBH_GUIDE_PARAMS floats); setting them resets the Adam moments */
int rbh_guide_set_params(rbh_ctx ctx, const float* params, uint32_t n);
int rbh_guide_get_params(rbh_ctx ctx, float* params, uint32_t n);
/* gradient of the loss over n records (no update; grad holds
 * RBH_GUIDE_PARAMS floats) */
int rbh_guide_gradient(rbh_ctx ctx, const float* records, uint32_t n, float* grad, float* loss);
/* steps Adam steps over consecutive batches of the n records */
int rbh_guide_train(rbh_ctx ctx, const float* records, uint32_t n, uint32_t batch, uint32_t steps, float* lastLoss);
/* raw network outputs, n x RBH_GUIDE_OUTPUTS floats: per lobe mean xyz, log
 * sharpness, logit, then the selection logit */
int rbh_guide_infer(rbh_ctx ctx, const float* records, uint32_t n, float* out);
/* the same with FP8 E4M3 matrix products: per-layer weight scales and
 * per-layer, per-wave activation scales */
int rbh_guide_infer_fp8(rbh_ctx ctx, const float* records, uint32_t n, float* out);
/* Training records gathered by the last RBH_RENDER_COLLECT render (count in
 * *n; up to max records of RBH_GUIDE_RECORD floats copied when out is not
 * null). */
int rbh_guide_records(rbh_ctx ctx, float* out, uint32_t max, uint32_t* n);
/* steps Adam steps over consecutive batches of the gathered records */
int rbh_guide_train_collected(rbh_ctx ctx, uint32_t batch, uint32_t steps, float* lastLoss);

#ifdef __cplusplus
}
#endif
#endif
