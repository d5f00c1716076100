// Software texture filtering with exact FP32 weights. Replaces the CUDA
// texture objects of the reference (src/core/texture.h:237-500: wrap / clamp /
// mirror addressing, linear filtering, normalised coordinates, sRGB decode)
// whose hardware weights are 8-bit fixed point.
#pragma once
#include "vec.hpp"
#include "types.hpp"
#include "spectral.hpp"

namespace rbh {
namespace dev {

__device__ __forceinline__ int wrap_index(int i, int n, uint32_t mode) {
    if (mode == 0) { int r = i % n; return r < 0 ? r + n : r; }
    if (mode == 2) {
        const int p = 2 * n;
        int r = i % p;
        if (r < 0) r += p;
        return r < n ? r : p - 1 - r;
    }
    return min(max(i, 0), n - 1);
}

// 32-entry-per-axis LUTs, clamp, linear (texel centres at (i+0.5)/32)
__device__ __forceinline__ void lut_coord(float u, int& i0, int& i1, float& f) {
    const float x = u * 32.0f - 0.5f;
    const float fl = floorf(x);
    f = x - fl;
    const int i = (int)fl;
    i0 = min(max(i, 0), 31);
    i1 = min(max(i + 1, 0), 31);
}
__device__ __forceinline__ float lut2(const float* __restrict__ t, float u, float v) {
    int x0, x1, y0, y1;
    float fx, fy;
    lut_coord(u, x0, x1, fx);
    lut_coord(v, y0, y1, fy);
    const float a = t[y0 * 32 + x0] * (1.0f - fx) + t[y0 * 32 + x1] * fx;
    const float b = t[y1 * 32 + x0] * (1.0f - fx) + t[y1 * 32 + x1] * fx;
    return a * (1.0f - fy) + b * fy;
}
__device__ __forceinline__ float lut3(const float* __restrict__ t, float u, float v, float w) {
    int x0, x1, y0, y1, z0, z1;
    float fx, fy, fz;
    lut_coord(u, x0, x1, fx);
    lut_coord(v, y0, y1, fy);
    lut_coord(w, z0, z1, fz);
    const float* p0 = t + z0 * 1024;
    const float* p1 = t + z1 * 1024;
    const float a0 = p0[y0 * 32 + x0] * (1.0f - fx) + p0[y0 * 32 + x1] * fx;
    const float b0 = p0[y1 * 32 + x0] * (1.0f - fx) + p0[y1 * 32 + x1] * fx;
    const float a1 = p1[y0 * 32 + x0] * (1.0f - fx) + p1[y0 * 32 + x1] * fx;
    const float b1 = p1[y1 * 32 + x0] * (1.0f - fx) + p1[y1 * 32 + x1] * fx;
    const float r0 = a0 * (1.0f - fy) + b0 * fy;
    const float r1 = a1 * (1.0f - fy) + b1 * fy;
    return r0 * (1.0f - fz) + r1 * fz;
}
__device__ __forceinline__ float4 lut2x4(const float* __restrict__ t, float u, float v) {
    int x0, x1, y0, y1;
    float fx, fy;
    lut_coord(u, x0, x1, fx);
    lut_coord(v, y0, y1, fy);
    const float4* q = reinterpret_cast<const float4*>(t);
    const float4 c00 = q[y0 * 32 + x0], c10 = q[y0 * 32 + x1], c01 = q[y1 * 32 + x0], c11 = q[y1 * 32 + x1];
    float4 r;
    r.x = (c00.x * (1.0f - fx) + c10.x * fx) * (1.0f - fy) + (c01.x * (1.0f - fx) + c11.x * fx) * fy;
    r.y = (c00.y * (1.0f - fx) + c10.y * fx) * (1.0f - fy) + (c01.y * (1.0f - fx) + c11.y * fx) * fy;
    r.z = (c00.z * (1.0f - fx) + c10.z * fx) * (1.0f - fy) + (c01.z * (1.0f - fx) + c11.z * fx) * fy;
    r.w = (c00.w * (1.0f - fx) + c10.w * fx) * (1.0f - fy) + (c01.w * (1.0f - fx) + c11.w * fx) * fy;
    return r;
}

// 8-bit sRGB channels are decoded through a 256-entry table that the
// context fills with srgb_decode(k / 255) (the same function the float
// formats use), so the results are those of decoding each texel.
__device__ __forceinline__ float4 texel(const TextureDesc& d, const uint8_t* __restrict__ blob, int x, int y,
                                        const float* __restrict__ srgbLut) {
    const size_t i = (size_t)y * d.width + x;
    const uint8_t* base = blob + d.offset;
    float4 o;
    switch (d.format) {
        case 0: {
            const uchar4 c = reinterpret_cast<const uchar4*>(base)[i];
            if (d.srgb) return make_float4(srgbLut[c.x], srgbLut[c.y], srgbLut[c.z], c.w / 255.0f);
            o = make_float4(c.x / 255.0f, c.y / 255.0f, c.z / 255.0f, c.w / 255.0f);
            break;
        }
        case 1:
            if (d.srgb) return make_float4(srgbLut[base[i]], srgbLut[0], srgbLut[0], 1.0f);
            o = make_float4(base[i] / 255.0f, 0.0f, 0.0f, 1.0f);
            break;
        case 2: o = reinterpret_cast<const float4*>(base)[i]; break;
        case 3: {
            const float* f = reinterpret_cast<const float*>(base) + 3 * i;
            o = make_float4(f[0], f[1], f[2], 1.0f);
            break;
        }
        default: o = make_float4(reinterpret_cast<const float*>(base)[i], 0.0f, 0.0f, 1.0f); break;
    }
    if (d.srgb) {
        o.x = srgb_decode(o.x);
        o.y = srgb_decode(o.y);
        o.z = srgb_decode(o.z);
    }
    return o;
}

__device__ __forceinline__ float4 tex2d(const TextureDesc* __restrict__ descs, const uint8_t* __restrict__ blob, int ti,
                                        float u, float v, const float* __restrict__ srgbLut) {
    const TextureDesc d = descs[ti];
    const int W = (int)d.width, H = (int)d.height;
    if (d.filter == 0)
        return texel(d, blob, wrap_index((int)floorf(u * W), W, d.addr[0]), wrap_index((int)floorf(v * H), H, d.addr[1]),
                     srgbLut);
    const float x = u * W - 0.5f, y = v * H - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = wrap_index((int)fx0, W, d.addr[0]), x1 = wrap_index((int)fx0 + 1, W, d.addr[0]);
    const int y0 = wrap_index((int)fy0, H, d.addr[1]), y1 = wrap_index((int)fy0 + 1, H, d.addr[1]);
    const float4 c00 = texel(d, blob, x0, y0, srgbLut), c10 = texel(d, blob, x1, y0, srgbLut);
    const float4 c01 = texel(d, blob, x0, y1, srgbLut), c11 = texel(d, blob, x1, y1, srgbLut);
    float4 r;
    r.x = (c00.x * (1.0f - ax) + c10.x * ax) * (1.0f - ay) + (c01.x * (1.0f - ax) + c11.x * ax) * ay;
    r.y = (c00.y * (1.0f - ax) + c10.y * ax) * (1.0f - ay) + (c01.y * (1.0f - ax) + c11.y * ax) * ay;
    r.z = (c00.z * (1.0f - ax) + c10.z * ax) * (1.0f - ay) + (c01.z * (1.0f - ax) + c11.z * ax) * ay;
    r.w = (c00.w * (1.0f - ax) + c10.w * ax) * (1.0f - ay) + (c01.w * (1.0f - ax) + c11.w * ax) * ay;
    return r;
}

// Environment map (RGB or RGBA float), clamp, linear.
__device__ __forceinline__ V3 env_lookup(const float* __restrict__ tex, int W, int H, int ch, float u, float v) {
    const float x = u * W - 0.5f, y = v * H - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    const int x0 = min(max((int)fx0, 0), W - 1), x1 = min(max((int)fx0 + 1, 0), W - 1);
    const int y0 = min(max((int)fy0, 0), H - 1), y1 = min(max((int)fy0 + 1, 0), H - 1);
    const float* p00 = tex + ((size_t)y0 * W + x0) * ch;
    const float* p10 = tex + ((size_t)y0 * W + x1) * ch;
    const float* p01 = tex + ((size_t)y1 * W + x0) * ch;
    const float* p11 = tex + ((size_t)y1 * W + x1) * ch;
    V3 r;
    r.x = (p00[0] * (1.0f - ax) + p10[0] * ax) * (1.0f - ay) + (p01[0] * (1.0f - ax) + p11[0] * ax) * ay;
    r.y = (p00[1] * (1.0f - ax) + p10[1] * ax) * (1.0f - ay) + (p01[1] * (1.0f - ax) + p11[1] * ax) * ay;
    r.z = (p00[2] * (1.0f - ax) + p10[2] * ax) * (1.0f - ay) + (p01[2] * (1.0f - ax) + p11[2] * ax) * ay;
    return r;
}

}  // namespace dev
}  // namespace rbh
