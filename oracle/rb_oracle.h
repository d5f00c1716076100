/* TEST INFRASTRUCTURE — C interface of the CPU oracle (liboracle.so).
 *
 * Loaded through ctypes by tests/, by the smoke check and by the bench's
 * CPU-baseline leg only. The scene records below have the same byte layout as
 * the product's include/ratbite_hip.h records so that one set of numpy
 * buffers feeds both sides; the oracle keeps its own copies of the structs so
 * it does not depend on product headers.
 */
#ifndef RB_ORACLE_H
#define RB_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_material {
    uint32_t bxdf;            /* 0 pure conductor, 1 pure dielectric, 2 complex surface */
    uint32_t flags;           /* 1 double sided, 2 sheen present, 4 spectral emission */
    uint32_t attributes;      /* 1 NORMAL, 2 FRAME, 4 COLOR, 8 TEX_COORD_1, 16 TEX_COORD_2 */
    uint32_t attributeStride; /* bytes per vertex in the attribute buffer */
    uint32_t normalOffset, frameOffset, texCoord1Offset, texCoord2Offset;
    uint32_t iorSpectrum, acSpectrum, emissionSpectrum;
    float mfRoughness, ior, alphaCutoff;
    uint32_t textures;        /* TextureTypeBitfield */
    uint32_t texCoordSets;    /* bit i: texture slot i reads TEX_COORD_2 */
    int32_t tex[7];           /* baseColor, normal, metRough, transmission, sheenColor, sheenRough, emission */
    uint32_t factors;         /* FactorTypeBitfield */
    float baseColor[4];
    float metalness, roughness, transmission;
    float sheenColor[3];
    float sheenRoughness;
    float emissive[3];
} orc_material;

typedef struct orc_texture {
    uint32_t width, height, depth;
    uint32_t format;          /* 0 RGBA8, 1 R8, 2 RGBA32F, 3 RGB32F, 4 R32F */
    uint32_t srgb;            /* decode RGB from sRGB before filtering */
    uint32_t addr[3];         /* 0 wrap, 1 clamp, 2 mirror */
    uint32_t filter;          /* 0 nearest, 1 linear */
    uint32_t pad;
    uint64_t offset;          /* byte offset into the texel blob */
} orc_texture;

typedef struct orc_emissive_tri { /* EmissiveTriangleLightData, 68 bytes */
    float vertices[9];
    float uvs[6];
    uint32_t materialIndex;
    uint32_t primitiveDataIndex;
} orc_emissive_tri;

typedef struct orc_shape_light { /* DiskLightData / SphereLightData, 48 bytes */
    float position[3];
    float powerScale;
    float frame[4];           /* unit quaternion x, y, z, w; a disk's normal is its third column */
    float radius;
    uint32_t materialIndex;
    uint32_t type;            /* 0 disk, 1 sphere */
    uint32_t pad;
} orc_shape_light;

typedef struct orc_scene {
    const float* verts; uint32_t nverts;
    const uint32_t* indices; uint32_t ntris;
    const uint8_t* attribs;               /* per-vertex packed attributes */
    const uint32_t* triMaterial;
    const uint32_t* triLight;             /* ~0u = not emissive */
    const orc_material* materials; uint32_t nmaterials;
    const orc_texture* textures; uint32_t ntextures;
    const uint8_t* texels;
    const float* spectra; uint32_t nspectra;          /* n x 471 */
    const float* xyz;                                  /* 3 x 471 sensor curves */
    const float* basis;                                /* 3 x 471 RGB basis */
    const float* lutConductor;                         /* 32 x 32 */
    const float* lutDielOuter; const float* lutDielInner;      /* 32^3 */
    const float* lutReflOuter; const float* lutReflInner;      /* 32^3 */
    const float* lutSheen;                             /* 32 x 32 x 4 */
    const orc_emissive_tri* lights; uint32_t nlights;
    const void* nodes; uint32_t nnodes;                /* PackedNode, 32 B */
    const uint32_t* lightPointers;
    const uint64_t* bitmasks;
    uint32_t envEnabled; uint32_t envWidth, envHeight, envChannels;
    const float* envTexels;
    const uint16_t* envCond; const uint16_t* envMarg;
    float envIntegral;
    const orc_shape_light* shapes; uint32_t nshapes;   /* light indices nlights + i */
} orc_scene;

typedef struct orc_params {
    uint32_t width, height;
    uint32_t maxPathDepth, maxReflectedPathDepth, maxTransmittedPathDepth;
    uint32_t sampleOffset, sampleCount;
    float camU[3], camV[3], camW[3];
    float perspectiveScaleW, perspectiveScaleH;
    uint32_t depthOfField;
    float aperture, focusDistance;
    uint32_t flags;
    /* optional crop: render only [x0,x0+cw) x [y0,y0+ch) of the film */
    uint32_t cropX, cropY, cropW, cropH;
} orc_params;

typedef struct orc_render_stats {
    uint64_t closestRays, shadowRays, pathVertices;
    double bvhSeconds;        /* CPU BVH build inside orc_render */
    double renderSeconds;     /* path tracing only (after the BVH build) */
} orc_render_stats;

/* QRNG / spectral unit functions (n independent inputs) */
int orc_sobol(const uint32_t* in /* n x {sample, hash, offset, dim, kind} */, float* out /* n x 3 */, uint32_t n);
int orc_sobol_table(uint32_t* out /* 4 x 32 */);
int orc_sg_erf(const float* x, float* out, uint32_t n);
int orc_pixel_hash(const uint32_t* xy, uint32_t* out, uint32_t n);
int orc_wavelengths(const float* u, float* lambda /* n x 4 */, float* pdf /* n x 4 */, uint32_t n);

/* env-map CDF */
int orc_cdf_build(const float* texels, int w, int h, int channels, float* cdf_x, float* cdf_y, float* integral);
int orc_cdf_build_f64(const float* texels, int w, int h, int channels, double* cdf_x, double* cdf_y, double* integral);
void orc_cdf_invert(const float* cdf_x, const float* cdf_y, int w, int h, uint16_t* cond, uint16_t* marg);
int orc_cdf_block_size(int w);

/* light tree */
int orc_light_tree_build(const float* tri_verts, const float* flux, uint32_t n,
                         void* nodes /* >= 2n x 32 B */, uint32_t* light_ptrs, uint64_t* bitmasks);
/* the same over n triangle lights and m shape lights (light indices n + i) */
int orc_light_tree_build_shapes(const float* tri_verts, const float* flux, uint32_t n, const orc_shape_light* shapes,
                                const float* shape_flux, uint32_t m, void* nodes /* >= 2(n+m) x 32 B */,
                                uint32_t* light_ptrs, uint64_t* bitmasks);
int orc_light_tree_pdf(const void* nodes, const uint32_t* lptrs, const uint64_t* masks, int env,
                       const uint32_t* lights, const float* pos_nrm /* n x 6 */, float* pdf, uint32_t n);
int orc_light_tree_sample(const void* nodes, const uint32_t* lptrs, const uint64_t* masks, int env,
                          const float* rnd_pos_nrm /* n x 7 */, uint32_t* type_index /* n x 2 */, float* pdf, uint32_t n);

/* geometry */
int orc_oct_encode_u32(const float* v, uint32_t* out, uint32_t n);
int orc_oct_decode_u32(const uint32_t* e, float* out, uint32_t n);
int orc_trace_brute(const float* verts, const uint32_t* idx, uint32_t ntris, const float* rays /* n x 8: o, tmin, d, tmax */,
                    float* hits /* n x 4: t, prim(as uint bits), b1, b2 */, uint32_t n, int anyhit);

/* per-sample BxDF check: evaluates the reference BxDF callable body for
 * given inputs; see oracle/bxdf_unit.cpp for the record layouts */
int orc_bxdf_unit(const orc_scene* sc, const float* in, float* out, uint32_t n);

/* single-function evaluations for the oracle's own pins (8 floats in,
 * 4 out per record; function table in capi.cpp) */
int orc_math_unit(int fn, const float* in, float* out, uint32_t n);

/* full CPU path tracer (the reference's __raygen__main restated) */
int orc_render(const orc_scene* sc, const orc_params* p, double* accum /* W*H*4 */, int threads, orc_render_stats* stats);

/* resolve: FP64 accumulation -> exposure -> colour-space matrix -> AgX -> RGBA8 */
int orc_resolve(const double* accum, uint32_t w, uint32_t h, const float* m3 /* column-major */, float exposureScale,
                uint8_t* rgba);

#ifdef __cplusplus
}
#endif
#endif
