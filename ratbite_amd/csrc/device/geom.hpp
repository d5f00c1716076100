// Geometric helpers of the hot path: octahedral coding
// (src/core/math_util.h:145-225), ray-origin offsetting
// (src/device/util.h:143-158), sampling routines (src/device/sampling.h,
// src/device/filter.h, src/device/dir_gen.h).
#pragma once
#include "vec.hpp"

namespace rbh {
namespace dev {

__device__ __forceinline__ V2 oct_enc(V3 v) {
    const float inv = 1.0f / (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
    const float px = v.x * inv, py = v.y * inv;
    if (v.z <= 0.0f)
        return {(1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f), (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f)};
    return {px, py};
}
__device__ __forceinline__ V3 oct_dec(float u, float v) {
    V3 r{u, v, 1.0f - fabsf(u) - fabsf(v)};
    const float t = fmaxf(0.0f, -r.z);
    r.x += r.x >= 0.0f ? -t : t;
    r.y += r.y >= 0.0f ? -t : t;
    return nrm(r);
}
// 16+16-bit quantisation of the geometric normal (pathtrace.cu:112, 911)
__device__ __forceinline__ uint32_t oct_enc_u32(V3 v) {
    const V2 p = oct_enc(v);
    const uint32_t x = (uint32_t)(clampf(p.x * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.99f);
    const uint32_t y = (uint32_t)(clampf(p.y * 0.5f + 0.5f, 0.0f, 1.0f) * 65535.99f);
    return (y << 16) | x;
}
__device__ __forceinline__ V3 oct_dec_u32(uint32_t e) {
    const float n = 1.0f / 65535.0f;
    return oct_dec((float)(e & 0xFFFFu) * n * 2.0f - 1.0f, (float)(e >> 16) * n * 2.0f - 1.0f);
}

// Integer-ulp origin offset with the reference's constants (1/32 origin
// band, 1/65536 float scale, 256 int scale).
__device__ __forceinline__ V3 offset_point(V3 p, V3 n) {
    const int ix = (int)(256.0f * n.x), iy = (int)(256.0f * n.y), iz = (int)(256.0f * n.z);
    const float ox = __int_as_float(__float_as_int(p.x) + (p.x < 0.0f ? -ix : ix));
    const float oy = __int_as_float(__float_as_int(p.y) + (p.y < 0.0f ? -iy : iy));
    const float oz = __int_as_float(__float_as_int(p.z) + (p.z < 0.0f ? -iz : iz));
    const float band = 1.0f / 32.0f, fs = 1.0f / 65536.0f;
    return {fabsf(p.x) < band ? p.x + fs * n.x : ox, fabsf(p.y) < band ? p.y + fs * n.y : oy,
            fabsf(p.z) < band ? p.z + fs * n.z : oz};
}

__device__ __forceinline__ V3 reflect_about(V3 v, V3 n) { return -v + 2.0f * dot(v, n) * n; }
// util.h:181-198 (returns zero on total internal reflection)
__device__ __forceinline__ V3 refract_through(V3 v, V3 n, float eta) {
    float ci = dot(v, n);
    if (ci < 0.0f) { ci = -ci; n = -n; }
    const float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    const float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) return {0.0f, 0.0f, 0.0f};
    return -v / eta + (ci / eta - sqrtf(1.0f - s2t)) * n;
}

// utility::refract with its validity flag and relative eta
// (src/device/util.h:71-96): eta is inverted when v comes from below n.
__device__ __forceinline__ V3 refract_rel(V3 v, V3 n, float eta, bool& valid, float& etaRel) {
    float ci = dot(v, n);
    if (ci < 0.0f) {
        eta = 1.0f / eta;
        ci = -ci;
        n = -n;
    }
    const float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    const float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) {
        valid = false;
        return {0.0f, 0.0f, 0.0f};
    }
    valid = true;
    etaRel = eta;
    return -v / eta + (ci / eta - sqrtf(1.0f - s2t)) * n;
}

__device__ __forceinline__ V2 disk_polar(V2 u) {
    const float r = sqrtf(u.x);
    float s, c;
    sincosf(2.0f * 3.14159265358979f * u.y, &s, &c);
    return {r * c, r * s};
}
__device__ __forceinline__ V3 cosine_hemisphere(V2 u) {
    const V2 d = disk_polar(u);
    return {d.x, d.y, sqrtf(fmaxf(0.0f, 1.0f - d.x * d.x - d.y * d.y))};
}

// filter.h:14-27: truncated Gaussian (radius 1.5, sigma 0.25) by Box-Muller
__device__ __forceinline__ V2 gaussian_offset(V2 u) {
    const float rad = 1.5f, sigma = 0.25f;
    const float uth = expf(-(rad * rad) / (2.0f * sigma * sigma));
    const float a = fmaxf(0.00001f, u.x * (1.0f - uth) + uth);
    const float m = sigma * sqrtf(-2.0f * logf(a));
    float s, c;
    sincosf(2.0f * 3.14159265358979f * u.y, &s, &c);
    return {m * c, m * s};
}

// src/device/mis.h:14-21 (nf = ng = 1), with PBRT-v4's guard: an infinite
// f^2 (a light sample seen exactly edge-on has an infinite solid-angle pdf)
// gives weight 1 instead of inf/inf = NaN, so the sample contributes
// spectrum * f * cos / inf = 0 rather than a NaN that would poison its pixel.
__device__ __forceinline__ float power_heuristic(float f, float g) {
    const float f2 = f * f, g2 = g * g;
    if (isinf(f2)) return 1.0f;
    return f2 / (f2 + g2);
}

}  // namespace dev
}  // namespace rbh
