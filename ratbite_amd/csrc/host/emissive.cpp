// Emitted flux of emissive triangles, as the reference's glTF loader
// computes it when it cuts emitters out of a mesh (src/core/scene.cpp:
// 485-632): the mean of the material's emissive factor, times the mean
// emission (r + g + b) a / 3 of the RGBA8 emission-texture texels whose
// centres the triangle's UV footprint covers, times the triangle's area and
// pi. The footprint is rasterised with the scanline rule of
// src/core/sw_rast.h (flat-bottom and flat-top halves, texel centres at +0.5,
// edges advanced incrementally in FP32); the texel set is part of the
// result, so the same rule and the same FP32 operation order are kept. A
// footprint that covers no texel centre falls back to the texels under its
// three corners. Texel addressing follows the loader's own mapping of wrap,
// clamp and mirror (scene.cpp:539-573), including its mirror rule, which
// also mirrors the first tile.
//
// Compiled without floating-point contraction, like the rest of host/.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "ratbite_hip.h"

namespace rbh {
namespace host {
namespace {

struct TexelSink {
    // either counts texels into a list (tests) or accumulates emission
    int* xy = nullptr;
    uint32_t cap = 0;
    uint32_t n = 0;
    const uint8_t* rgba = nullptr;
    int w = 0, h = 0;
    uint32_t addrX = 0, addrY = 0;
    double sum = 0.0;

    static int address(int x, int size, uint32_t mode) {
        if (mode == 0) {   // wrap
            x = (x >= 0) ? x : x + size * ((-x - 1) / size + 1);
            return x % size;
        }
        if (mode == 1) return x < 0 ? 0 : (x > size - 1 ? size - 1 : x);   // clamp
        if (mode == 2) {   // mirror, as the loader writes it
            x = (x >= 0) ? x : x + size * ((-x - 1) / size + 2);
            const bool even = ((x / size) % 2) == 0;
            x = x % size;
            return even ? size - 1 - x : x;
        }
        return 0;
    }

    void operator()(int x, int y) {
        if (xy) {
            if (n < cap) {
                xy[2 * n] = x;
                xy[2 * n + 1] = y;
            }
            ++n;
            return;
        }
        const int tx = address(x, w, addrX), ty = address(y, h, addrY);
        const uint8_t* t = rgba + 4 * ((size_t)w * ty + tx);
        const float r = t[0] / 255.0f, g = t[1] / 255.0f, b = t[2] / 255.0f, a = t[3] / 255.0f;
        sum += (r + g + b) * a / 3.0f;
        ++n;
    }
};

// Rows from the top vertex down to the flat bottom edge (a.y > b.y == c.y).
void flat_bottom(const float* a, const float* b, const float* c, TexelSink& out) {
    float lx = b[0], ly = b[1], rx = c[0], ry = c[1];
    if (lx > rx) {
        const float tx = lx, ty = ly;
        lx = rx; ly = ry;
        rx = tx; ry = ty;
    }
    const float slopeL = (lx - a[0]) / (a[1] - ly);
    const float slopeR = (rx - a[0]) / (a[1] - ry);
    const float toCentre = a[1] - (std::floor(a[1] - 0.5f) + 0.5f);
    float xl = a[0] + slopeL * toCentre;
    float xr = a[0] + slopeR * toCentre;
    const int first = (int)std::floor(a[1] + 0.5f);
    const int last = (int)std::floor(ly + 0.5f);
    for (int row = first - 1; row >= last; --row) {
        const int x0 = (int)std::floor(xl + 0.5f);
        const int x1 = (int)std::floor(xr + 0.5f) - 1;
        for (int x = x0; x <= x1; ++x) out(x, row);
        xl += slopeL;
        xr += slopeR;
    }
}

// Rows from the flat top edge (a.y == b.y > c.y) down to the bottom vertex,
// visited bottom-up.
void flat_top(const float* a, const float* b, const float* c, TexelSink& out) {
    float lx = a[0], ly = a[1], rx = b[0], ry = b[1];
    if (lx > rx) {
        const float tx = lx, ty = ly;
        lx = rx; ly = ry;
        rx = tx; ry = ty;
    }
    const float slopeL = (lx - c[0]) / (ly - c[1]);
    const float slopeR = (rx - c[0]) / (ry - c[1]);
    const float toCentre = (std::ceil(c[1] - 0.5f) + 0.5f) - c[1];
    float xl = c[0] + slopeL * toCentre;
    float xr = c[0] + slopeR * toCentre;
    const int end = (int)std::floor(ry + 0.5f);
    const int start = (int)std::floor(c[1] + 0.5f);
    for (int row = start; row < end; ++row) {
        const int x0 = (int)std::floor(xl + 0.5f);
        const int x1 = (int)std::floor(xr + 0.5f) - 1;
        for (int x = x0; x <= x1; ++x) out(x, row);
        xl += slopeL;
        xr += slopeR;
    }
}

void rasterise(const float* p0, const float* p1, const float* p2, TexelSink& out) {
    const float* v[3] = {p0, p1, p2};
    // descending y (three compare-swaps)
    if (v[0][1] < v[1][1]) { const float* t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1][1] < v[2][1]) { const float* t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0][1] < v[1][1]) { const float* t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1][1] == v[2][1]) {
        flat_bottom(v[0], v[1], v[2], out);
    } else if (v[0][1] == v[1][1]) {
        flat_top(v[0], v[1], v[2], out);
    } else {
        // split at the middle vertex's height on the long edge
        const float split[2] = {v[0][0] + (v[2][0] - v[0][0]) * ((v[1][1] - v[0][1]) / (v[2][1] - v[0][1])), v[1][1]};
        flat_bottom(v[0], v[1], split, out);
        flat_top(v[1], split, v[2], out);
    }
}

// glm::length(glm::cross(e1, e2)) in glm's operation order
float cross_length(const float* a, const float* b, const float* c) {
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float x = e1[1] * e2[2] - e2[1] * e1[2];
    const float y = e1[2] * e2[0] - e2[2] * e1[0];
    const float z = e1[0] * e2[1] - e2[0] * e1[1];
    const float xx = x * x, yy = y * y, zz = z * z;
    return std::sqrt(xx + yy + zz);
}

}  // namespace
}  // namespace host
}  // namespace rbh

using namespace rbh::host;

extern "C" int rbh_emissive_flux(const float* verts, const float* uvs, uint32_t n, const float* emissiveFactor,
                                 const uint8_t* rgba8, uint32_t texW, uint32_t texH, uint32_t addrX, uint32_t addrY,
                                 float* flux) {
    if (n && (!verts || !flux || !emissiveFactor)) return RBH_ERR_INVALID;
    const bool textured = rgba8 != nullptr;
    if (textured && (!uvs || texW == 0 || texH == 0 || addrX > 2 || addrY > 2)) return RBH_ERR_INVALID;
    const float pi = 3.14159265358979323846f;
    for (uint32_t k = 0; k < n; ++k) {
        float f = (emissiveFactor[0] + emissiveFactor[1] + emissiveFactor[2]) / 3.0f;
        if (textured) {
            const float* uv = uvs + 6 * (size_t)k;
            const float W = (float)texW, H = (float)texH;
            const float t0[2] = {uv[0] * W, uv[1] * H}, t1[2] = {uv[2] * W, uv[3] * H}, t2[2] = {uv[4] * W, uv[5] * H};
            TexelSink s;
            s.rgba = rgba8;
            s.w = (int)texW;
            s.h = (int)texH;
            s.addrX = addrX;
            s.addrY = addrY;
            rasterise(t0, t1, t2, s);
            if (s.n == 0) {
                // truncating float -> int, as the loader's (int, int) callback does
                s((int)t0[0], (int)t0[1]);
                s((int)t1[0], (int)t1[1]);
                s((int)t2[0], (int)t2[1]);
            }
            f *= s.sum / s.n;
        }
        const float* v = verts + 9 * (size_t)k;
        f *= cross_length(v, v + 3, v + 6) * 0.5f * pi;
        flux[k] = f;
    }
    return RBH_OK;
}

extern "C" int rbh_rasterise_uv_triangle(const float* a, const float* b, const float* c, int* xy, uint32_t capacity,
                                         uint32_t* count) {
    if (!a || !b || !c || !count || (capacity && !xy)) return RBH_ERR_INVALID;
    TexelSink s;
    s.xy = xy;
    s.cap = capacity;
    rasterise(a, b, c, s);
    *count = s.n;
    return s.n > capacity ? RBH_ERR_RANGE : RBH_OK;
}
