// Spherical-Gaussian light-tree importance, sampling and PDF
// (src/kernels/pathtrace.cu:974-1112, src/core/math_util.h:229-588,
// src/core/light_tree_types.h:165-206). The SG formulas are Tokuyoshi's
// published ones (VSGL; "Hierarchical Light Sampling with Accurate Spherical
// Gaussian Lighting", 2024); erf is the project's fitted polynomial below.
#pragma once
#include "geom.hpp"
#include "types.hpp"

namespace rbh {
namespace dev {

// Light-pointer types (light_tree_types.h:10-18, 191-207); LT_SHAPE, the free
// fourth value of the 2-bit type field, marks a disk or sphere light whose
// pointer index is its light index (after the triangle lights).
constexpr uint32_t LT_TRIANGLE = 0, LT_SKY = 1, LT_NONE = 2, LT_SHAPE = 3;
constexpr float kEnvImportance = 0.3f;

__device__ __forceinline__ float sg_expm1(float x) {
    const float u = expf(x);
    if (u == 1.0f) return x;
    const float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return y * x / logf(u);
    return y;
}
__device__ __forceinline__ float sg_expm1_over_x(float x) {
    const float u = expf(x);
    if (u == 1.0f) return 1.0f;
    const float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return y / logf(u);
    return y / x;
}
// erf for the spherical-Gaussian integrals: a piecewise Chebyshev fit made
// for this project (erf(x)/x as a degree-7 polynomial in x^2 on [0, 1];
// erf(x) as a degree-16 polynomial in x - 2.5 on [1, 4]; +-1 beyond 4, where
// the reference also saturates). Max abs error 1.7e-7 against the exact erf.
// Horner steps are explicit FMAs, so every implementation evaluates it to the
// same bits for the same input, which keeps the cancellation in
// erfc = 1 - erf (math_util.h:326-329) reproducible.
__device__ __forceinline__ float sg_erf(float x) {
    const float ax = fabsf(x);
    if (ax >= 4.0f) return copysignf(1.0f, x);
    float r;
    if (ax <= 1.0f) {
        const float u = ax * ax;
        const float A[8] = {1.128379225730896f, -0.37612637877464294f, 0.11283782869577408f, -0.026865437626838684f, 0.005221053492277861f, -0.0008484396385028958f, 0.00011268253729213029f, -9.673591193859465e-06f};
        float h = A[7];
#pragma unroll
        for (int k = 6; k >= 0; --k) h = fmaf(h, u, A[k]);
        r = ax * h;
    } else {
        const float y = ax - 2.5f;
        const float B[17] = {0.9995930194854736f, 0.0021782785188406706f, -0.005445701535791159f, 0.008350212126970291f, -0.008622484281659126f, 0.006116587668657303f, -0.002797981258481741f, 0.0005445338319987059f, 0.00025947490939870477f, -0.0002534241648390889f, 8.039429667405784e-05f, 7.756445484119467e-06f, -1.5348350643762387e-05f, 3.462653694441542e-06f, 7.886890216468601e-07f, -3.5129914977005683e-07f, 1.1272015321139861e-08f};
        float h = B[16];
#pragma unroll
        for (int k = 15; k >= 0; --k) h = fmaf(h, y, B[k]);
        r = h;
    }
    return copysignf(r, x);
}
__device__ __forceinline__ float sg_erfc(float x) { return 1.0f - sg_erf(x); }

__device__ __forceinline__ float sg_integral(float s) { return 4.0f * 3.14159265359f * sg_expm1_over_x(-2.0f * s); }

__device__ __forceinline__ float sg_upper(float s) {
    if (s <= 0.5f)
        return (((((((-1.0f / 362880.0f) * s + 1.0f / 40320.0f) * s - 1.0f / 5040.0f) * s + 1.0f / 720.0f) * s -
                  1.0f / 120.0f) * s + 1.0f / 24.0f) * s - 1.0f / 6.0f) * s + 0.5f;
    return (sg_expm1(-s) + s) / (s * s);
}
__device__ __forceinline__ float sg_lower(float s) {
    const float e = expf(-s);
    if (s <= 0.5f)
        return e * (((((((((1.0f / 403200.0f) * s - 1.0f / 45360.0f) * s + 1.0f / 5760.0f) * s - 1.0f / 840.0f) * s +
                        1.0f / 144.0f) * s - 1.0f / 30.0f) * s + 1.0f / 8.0f) * s - 1.0f / 3.0f) * s + 0.5f);
    return e * (-sg_expm1(-s) - s * e) / (s * s);
}
// Listing 7 of the 2024 supplement (math_util.h:514-538)
__device__ __forceinline__ float sg_clamped_cos_integral_over_pi(float cosine, float s) {
    const float A = 2.7360831611272558028247203765204f, B = 17.02129778174187535455530451145f;
    const float C = 4.0100826728510421403939290030394f, D = 15.219156263147210594866010069381f;
    const float E = 76.087896272360737270901154261082f;
    const float t = s * sqrtf(0.5f * ((s + A) * s + B) / (((s + C) * s + D) * s + E));
    const float tz = t * cosine;
    const float inv_sqrt_pi = 0.56418958354775628694807945156077f;
    const float lerp = clampf(fmaxf(0.5f * (cosine * sg_erfc(-tz) + sg_erfc(t)) -
                                        0.5f * inv_sqrt_pi * expf(-tz * tz) * sg_expm1(t * t * (cosine * cosine - 1.0f)) / t,
                                    0.5f * 1.19209290e-07f),
                              0.0f, 1.0f);
    return 2.0f * lerpf(sg_lower(s), sg_upper(s), lerp);
}

struct Node {
    V3 mean, dir;
    float variance, flux, sharpness;
    float integral;          // sg_integral(sharpness)
    bool leaf;
    uint32_t right, count, offset;
};

// A packed node as the traversal reads it: the decoded direction and the
// node's spherical-Gaussian integral computed once when the tree is built
// (by the same device functions, so importances are bit-identical), 48 B.
struct TreeNodeX {
    float4 meanVar;          // mean, variance
    float4 dirSharp;         // decoded average direction, sharpness
    float4 fluxIntCore;      // flux, sg_integral(sharpness), core bits, 0
};
static_assert(sizeof(TreeNodeX) == 48, "TreeNodeX is 48 bytes");

// light_tree_types.h:165-189; the direction is decoded from [0,1]-scaled
// octahedral components without the [-1,1] remap, as the reference does.
__device__ __forceinline__ Node load_node(const PackedNode* __restrict__ nodes, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    const float4 a = p[0], b = p[1];
    Node n;
    n.mean = {a.x, a.y, a.z};
    n.variance = a.w;
    const uint32_t d = __float_as_uint(b.x);
    n.dir = oct_dec((float)(d & 0xFFFFu) / 65535.0f, (float)(d >> 16) / 65535.0f);
    n.sharpness = b.y;
    n.flux = b.z;
    const uint32_t core = __float_as_uint(b.w);
    n.leaf = (core >> 31) != 0u;
    n.right = core;
    n.count = (core >> 27) & 15u;
    n.offset = core & ((1u << 27) - 1u);
    n.integral = sg_integral(n.sharpness);
    return n;
}
__device__ __forceinline__ TreeNodeX expand_node(const Node& n) {
    TreeNodeX x;
    x.meanVar = make_float4(n.mean.x, n.mean.y, n.mean.z, n.variance);
    x.dirSharp = make_float4(n.dir.x, n.dir.y, n.dir.z, n.sharpness);
    x.fluxIntCore = make_float4(n.flux, n.integral, __uint_as_float(n.right), 0.0f);
    return x;
}
__device__ __forceinline__ Node load_node(const TreeNodeX* __restrict__ nodes, uint32_t i) {
    const TreeNodeX x = nodes[i];
    Node n;
    n.mean = {x.meanVar.x, x.meanVar.y, x.meanVar.z};
    n.variance = x.meanVar.w;
    n.dir = {x.dirSharp.x, x.dirSharp.y, x.dirSharp.z};
    n.sharpness = x.dirSharp.w;
    n.flux = x.fluxIntCore.x;
    n.integral = x.fluxIntCore.y;
    const uint32_t core = __float_as_uint(x.fluxIntCore.z);
    n.leaf = (core >> 31) != 0u;
    n.right = core;
    n.count = (core >> 27) & 15u;
    n.offset = core & ((1u << 27) - 1u);
    return n;
}

// pathtrace.cu:974-1001
__device__ __forceinline__ float node_importance(const Node& n, V3 pos, V3 nrmv) {
    V3 ld = n.mean - pos;
    const float d2 = dot(ld, ld);
    ld = ld * (1.0f / sqrtf(d2));
    const float var = fmaxf(n.variance, 0x1.0p-31f * d2);
    const float ls = d2 / var;
    // SG product of (dir, sharpness) and (ld, ls), math_util.h:362-375
    const V3 ax = n.dir * n.sharpness + ld * ls;
    const float sh = len(ax);
    const float cosine = clampf(dot(n.dir, ld), -1.0f, 1.0f);
    const float smin = fminf(n.sharpness, ls);
    const float ratio = smin / fmaxf(fmaxf(n.sharpness, ls), 1.17549435e-38f);
    const float logA = 2.0f * smin * (cosine - 1.0f) / (sqrtf(2.0f * ratio * cosine + ratio * ratio + 1.0f) + ratio + 1.0f);
    const V3 axis = ax / fmaxf(sh, 1.17549435e-38f);
    const float amp = expf(logA);
    const float c = clampf(dot(axis, nrmv), -1.0f, 1.0f);
    return (n.flux / (var * n.integral)) * amp * sg_clamped_cos_integral_over_pi(c, sh);
}

struct LightTree {
    const TreeNodeX* __restrict__ nodes;    // nullptr: no tree
    const uint32_t* __restrict__ pointers;
    const uint64_t* __restrict__ masks;
    bool env;

    // pathtrace.cu:1002-1049. Inlined into the shading kernels: as a call it
    // made them spill the registers live across it (measured 1.0 ms per
    // 1080p frame of the shade kernel).
    __device__ __forceinline__ float pdf(uint32_t type, uint32_t index, V3 pos, V3 n) const {
        const float envImp = env ? kEnvImportance : 0.0f;
        if (type == LT_SKY) return envImp;
        float p = 1.0f - envImp;
        const uint64_t mask = masks[index];
        uint32_t cur = 0;
        Node c = load_node(nodes, 0);
        for (int depth = 0; depth < 64 && !c.leaf; ++depth) {
            const Node l = load_node(nodes, cur + 1);
            const Node r = load_node(nodes, c.right);
            const float li = node_importance(l, pos, n);
            const float ri = node_importance(r, pos, n);
            const bool right = (mask >> depth) & 1ull;
            cur = right ? c.right : cur + 1;
            p *= right ? ri / (li + ri) : li / (li + ri);
            c = right ? r : l;
        }
        return p * (1.0f / (float)c.count);
    }

    // pathtrace.cu:1050-1112 (inlined like pdf)
    __device__ __forceinline__ void sample(float rnd, V3 pos, V3 n, uint32_t& type, uint32_t& index, float& p) const {
        const float envImp = env ? (nodes ? kEnvImportance : 1.0f) : 0.0f;
        if (rnd < envImp) { type = LT_SKY; index = 0; p = envImp; return; }
        rnd = clampf((rnd - envImp) / (1.0f - envImp), 0.0f, 1.0f);
        p = 1.0f - envImp;
        uint32_t cur = 0;
        Node c = load_node(nodes, 0);
        while (!c.leaf) {
            const Node l = load_node(nodes, cur + 1);
            const Node r = load_node(nodes, c.right);
            const float li = node_importance(l, pos, n);
            const float ri = node_importance(r, pos, n);
            const float sep = li / (li + ri);
            const bool right = rnd > sep;
            cur = right ? c.right : cur + 1;
            p *= right ? 1.0f - sep : sep;
            rnd = clampf(right ? (rnd - sep) / (1.0f - sep) : rnd / sep, 0.0f, 1.0f);
            c = right ? r : l;
        }
        const uint32_t k = (uint32_t)((float)c.count * rnd);
        const uint32_t slot = c.offset + (k == c.count ? c.count - 1u : k);
        p *= 1.0f / (float)c.count;
        const uint32_t lp = pointers[slot];
        type = lp >> 30;
        index = lp & ((1u << 30) - 1u);
        if (isnan(p) || isinf(p) || p == 0.0f) type = LT_NONE;
    }
};

}  // namespace dev
}  // namespace rbh
