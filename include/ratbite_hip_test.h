/*
 * ratbite_hip_test.h — test-only C ABI: runs single device functions of the
 * hot path over batches of inputs so each can be compared with a CPU
 * restatement. Inputs and outputs are host buffers; n is the batch size.
 * The reference has no such interface (its device functions are only
 * reachable inside the OptiX pipeline); each entry point names the
 * reference function it exercises.
 */
#ifndef RATBITE_HIP_TEST_H
#define RATBITE_HIP_TEST_H
#include "ratbite_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* QRNG::Sobol::sample1D/2D/3D (src/device/quasi_random.h:76-99).
 * in: n x {sample, pixelHash, dimensionOffset, dimension, kind(1|2|3)}; out: n x 3 floats */
int rbh_test_sobol(rbh_ctx ctx, const uint32_t* in, float* out, uint32_t n);
/* QRNG::getPixelHash (quasi_random.h:30-37). in: n x {x, y}; out: n */
int rbh_test_pixel_hash(rbh_ctx ctx, const uint32_t* xy, uint32_t* out, uint32_t n);
/* SampledWavelengths::sampleVisible (src/device/spectral.h:227-241). out: n x 4 each */
int rbh_test_wavelengths(rbh_ctx ctx, const float* u, float* lambda, float* pdf, uint32_t n);
/* color::RGBtoSpectrum (src/device/color.h:369-381) with the uploaded basis.
 * in: n x {r, g, b, l0, l1, l2, l3}; out: n x 4 */
int rbh_test_rgb_to_spectrum(rbh_ctx ctx, const float* in, float* out, uint32_t n);
/* Octohedral::encodeU32 / decodeU32 (src/core/math_util.h:199-219) */
int rbh_test_oct(rbh_ctx ctx, const float* v, uint32_t* enc, float* dec, uint32_t n);
/* Closest-hit / any-hit query on the built LBVH (optixTrace / optixTraverse,
 * src/kernels/pathtrace.cu:1162-1177, 1704-1720). rays: n x {o.xyz, tmin
 * (ignored, 0), d.xyz, tmax}; hits: n x {t, prim (uint bits, 0xFFFFFFFF =
 * miss), b1, b2}. mode bit 0: any hit; bit 1: a 2-entry LDS stack so that the
 * global spill part of the traversal stack is exercised; bit 2: traverse the
 * binary tree instead of the four-wide one the renderer uses; bit 3: the
 * four-wide nodes quantised to 64 bytes. */
int rbh_test_trace(rbh_ctx ctx, const float* rays, float* hits, uint32_t n, int mode);
/* The built LBVH (replaces optixAccelBuild, rendering_interface.cpp:733-1064):
 * sorted 64-bit keys (Morton code << b | triangle, b the bits that hold n - 1,
 * the Morton code min(21, (64 - b) / 3) bits per axis), n entries; 64-byte nodes
 * (nnodes entries); 64-byte triangle records in leaf order (vertices,
 * triangle, material, light, vertex indices; n entries). Any
 * output pointer may be null. */
int rbh_test_lbvh_read(rbh_ctx ctx, uint64_t* keys, void* nodes, void* tris, uint32_t* nnodes);
/* A plain device copy of `bytes` bytes, `repeats` times (no reference
 * counterpart): calibrates HBM byte counters and measures copy bandwidth;
 * *ms = average time per copy. */
int rbh_test_copy(rbh_ctx ctx, uint64_t bytes, int repeats, float* ms);
/* The LBVH build's stable radix sort (replaces the sort inside
 * optixAccelBuild; no reference source): n 64-bit keys (and, when vals is
 * non-null, n 32-bit values) sorted by bits [beginBit, endBit) into
 * outKeys / outVals. Host buffers. */
int rbh_test_radix_sort(rbh_ctx ctx, const uint64_t* keys, const uint32_t* vals, uint32_t n, int beginBit, int endBit,
                        uint64_t* outKeys, uint32_t* outVals);
/* The guide's training-set compaction (no reference counterpart): the
 * indices i < n with weights[i] > 0, in increasing order, into out; their
 * number into *count. Host buffers. */
int rbh_test_select_positive(rbh_ctx ctx, const float* weights, uint32_t n, uint32_t* out, uint32_t* count);
/* traverseLightTree (src/kernels/pathtrace.cu:1002-1049) on the built tree.
 * lights: n light indices; posNrm: n x 6; out: n pdfs */
int rbh_test_light_pdf(rbh_ctx ctx, const uint32_t* lights, const float* posNrm, float* pdf, uint32_t n);
/* sampleLightTree (pathtrace.cu:1050-1112). in: n x {rand, pos.xyz, nrm.xyz};
 * typeIndex: n x 2; pdf: n */
int rbh_test_light_sample(rbh_ctx ctx, const float* in, uint32_t* typeIndex, float* pdf, uint32_t n);
/* BxDF callables (pathtrace.cu:464-547, 548-905) for one shading point each.
 * Record layouts: 64-word input, 16-word output, documented in
 * ratbite_amd/csrc/kernels/unit.hip. */
int rbh_test_bxdf(rbh_ctx ctx, const float* in, float* out, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
