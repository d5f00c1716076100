// Hero-wavelength spectral sampling and RGB/XYZ conversions
// (src/device/spectral.h:196-327, src/device/color.h:10-43).
#pragma once
#include "vec.hpp"
#include "types.hpp"

namespace rbh {
namespace dev {

struct S4 { float v[kNL]; };
__device__ __forceinline__ S4 s4(float a) { return S4{{a, a, a, a}}; }
__device__ __forceinline__ S4 operator+(const S4& a, const S4& b) { return S4{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}}; }
__device__ __forceinline__ S4 operator*(const S4& a, const S4& b) { return S4{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}}; }
__device__ __forceinline__ S4 operator*(const S4& a, float s) { return S4{{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
__device__ __forceinline__ S4 operator*(float s, const S4& a) { return a * s; }
__device__ __forceinline__ S4 operator/(const S4& a, float s) { return S4{{a.v[0] / s, a.v[1] / s, a.v[2] / s, a.v[3] / s}}; }
__device__ __forceinline__ S4& operator+=(S4& a, const S4& b) { for (int i = 0; i < kNL; ++i) a.v[i] += b.v[i]; return a; }
__device__ __forceinline__ S4& operator*=(S4& a, const S4& b) { for (int i = 0; i < kNL; ++i) a.v[i] *= b.v[i]; return a; }
__device__ __forceinline__ float smax4(const S4& a) {
    float t = a.v[0];
    for (int i = 1; i < kNL; ++i) if (a.v[i] > t) t = a.v[i];
    return t;
}
__device__ __forceinline__ float savg4(const S4& a) { return (((a.v[0] + a.v[1]) + a.v[2]) + a.v[3]) / (float)kNL; }

struct Wl { S4 lambda, pdf; };

// spectral.h:289-299: sampling the visible-range density 1/cosh^2
__device__ __forceinline__ float visible_wavelength(float u) {
    return 538.0f - 138.888889f * atanhf(0.85691062f - 1.82750197f * u);
}
__device__ __forceinline__ float visible_pdf(float l) {
    if (l < 360.0f || l > 830.0f) return 0.0f;
    const float c = coshf(0.0072f * (l - 538.0f));
    return 0.0039398042f / (c * c);
}
// spectral.h:227-241: hero wavelength + 3 rotations by 1/4
__device__ __forceinline__ Wl sample_wavelengths(float u) {
    Wl w;
#pragma unroll
    for (int i = 0; i < kNL; ++i) {
        float up = u + (float)i / (float)kNL;
        if (up > 1.0f) up -= 1.0f;
        w.lambda.v[i] = visible_wavelength(up);
        w.pdf.v[i] = visible_pdf(w.lambda.v[i]);
    }
    return w;
}

// DenseSpectrum::sample (spectral.h:313-327): lroundf index into 1 nm table
__device__ __forceinline__ int dense_index(float l) { return (int)lroundf(l) - kLambdaMin; }
__device__ __forceinline__ S4 dense4(const float* __restrict__ tab, const Wl& w) {
    S4 r;
#pragma unroll
    for (int i = 0; i < kNL; ++i) r.v[i] = tab[dense_index(w.lambda.v[i])];
    return r;
}

// Colour tables: rows X, Y, Z, basis R, G, B (471 floats each).
struct ColorTables {
    const float* __restrict__ t;
    __device__ __forceinline__ S4 rgb(V3 c, const Wl& w) const {
        S4 r;
#pragma unroll
        for (int i = 0; i < kNL; ++i) {
            const int k = dense_index(w.lambda.v[i]);
            r.v[i] = t[3 * kDense + k] * c.x + t[4 * kDense + k] * c.y + t[5 * kDense + k] * c.z;
        }
        return r;
    }
    // color.h:10-16 (sensor curves = CIE X, Y, Z)
    __device__ __forceinline__ void xyz(const Wl& w, const S4& L, double out[3]) const {
        S4 a, b, c;
#pragma unroll
        for (int i = 0; i < kNL; ++i) {
            const int k = dense_index(w.lambda.v[i]);
            a.v[i] = t[k] * L.v[i];
            b.v[i] = t[kDense + k] * L.v[i];
            c.v[i] = t[2 * kDense + k] * L.v[i];
        }
        out[0] = (double)savg4(a);
        out[1] = (double)savg4(b);
        out[2] = (double)savg4(c);
    }
};

__device__ __forceinline__ float srgb_decode(float c) {
    return c < 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

}  // namespace dev
}  // namespace rbh
