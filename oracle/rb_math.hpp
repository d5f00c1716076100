// TEST INFRASTRUCTURE — CPU oracle for the spectral path-tracing hot path.
//
// Scalar C++ restatement of ratbite's device math (src/device/*.h,
// src/core/math_util.h, src/core/light_tree_types.h). Only tests/, the smoke
// check and the bench's CPU-baseline leg may load this code; the product path
// (ratbite_amd/) never links it.
//
// Every function cites the reference file:line it restates. FP32 throughout,
// like the reference; transcendental calls go to glibc (the reference's go to
// CUDA libm), so per-function agreement with the GPU product is a tolerance
// check, not a bit check. Integer functions are exact.
#pragma once

#include <cstdint>
#include <cstring>
#include <cmath>
#include <cfloat>
#include <algorithm>

namespace orc {

// ---------------------------------------------------------------- vectors --
struct V2 { float x, y; };
struct V3 { float x, y, z; };

inline V2 v2(float x, float y) { return {x, y}; }
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 v3s(float s) { return {s, s, s}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glm::normalize = v * inversesqrt(dot(v, v))
inline V3 normalize(V3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }
inline float clampf(float x, float a, float b) { return std::min(std::max(x, a), b); }
inline float mixf(float a, float b, float t) { return a + t * (b - a); }
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline int32_t f2i(float f) { int32_t u; std::memcpy(&u, &f, 4); return u; }
inline float i2f(int32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Column-major 3x3 (columns c[0..2]), matching glm::mat3 semantics.
struct M3 { V3 c[3]; };
inline V3 mul(const M3& m, V3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }
inline M3 transpose(const M3& m) {
    return M3{{{m.c[0].x, m.c[1].x, m.c[2].x}, {m.c[0].y, m.c[1].y, m.c[2].y}, {m.c[0].z, m.c[1].z, m.c[2].z}}};
}

constexpr float PI = 3.14159265358979323846f;
constexpr int NL = 4;                 // src/core/spectral_settings.h:5
constexpr int LMIN = 360, LMAX = 830; // src/core/spectral_settings.h:6-7
constexpr int NDENSE = LMAX - LMIN + 1;

// ------------------------------------------------------------------- QRNG --
// src/device/quasi_random.h:13-29
enum Dim : uint32_t {
    D_FILTER = 0, D_LENS = 1, D_WAVELENGTH = 0, D_LIGHT = 1,
    D_BXDF0 = 2, D_BXDF1 = 3, D_BXDF2 = 4, D_BXDF3 = 5,
    D_ALPHA_BLEND = 7, D_ROULETTE = 8, D_DESC = 16,
};

// src/device/quasi_random.h:30-37
inline uint32_t pixel_hash(uint32_t x, uint32_t y) {
    const uint32_t qx = 1103515245U * ((x >> 1U) ^ y);
    const uint32_t qy = 1103515245U * ((y >> 1U) ^ x);
    return 1103515245U * (qx ^ (qy >> 3U));
}

// src/device/quasi_random.h:38-64
struct QState {
    uint32_t sample, offset, hash;
    static QState make(uint32_t sample, uint32_t hash) { return {sample, D_DESC, hash}; }
    void advance_sample() { ++sample; offset = D_DESC; }
    void advance_bounce() { offset += D_DESC; }
};

inline uint32_t brev(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) { r = (r << 1) | (x & 1u); x >>= 1; }
    return r;
}
inline uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }

// src/device/quasi_random.h:101-110
inline uint32_t dim_hash(uint32_t i) {
    i ^= i >> 16; i *= 0x21f0aaadu; i ^= i >> 15; i *= 0xd35a2d97u; i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}
// src/device/quasi_random.h:119-127
inline uint32_t reversed_bit_owen(uint32_t n, uint32_t seed) {
    n ^= n * 0x3d20adeau; n += seed; n *= (seed >> 16) | 1u; n ^= n * 0x05526c56u; n ^= n * 0x53a22864u;
    return n;
}
// src/device/quasi_random.h:128-131: u32 * (1/4294967808)
inline float u2unit(uint32_t x) { return (float)x * (1.0f / 4294967808.0f); }

// Direction numbers of the first four Sobol dimensions in the bit-reversed
// ("Burley") form used by src/device/quasi_random.h:137-179. Generated from
// the Joe-Kuo primitive polynomials / initial numbers instead of transcribed:
// dim0 = identity (van der Corput), dim1: s=1 a=0 m={1}, dim2: s=2 a=1
// m={1,3}, dim3: s=3 a=1 m={1,3,1}.
struct SobolTable { uint32_t t[4][32]; };
const SobolTable& sobol_table();

// src/device/quasi_random.h:132-204
inline float sobol_burley(uint32_t rev, uint32_t dim, uint32_t seed) {
    uint32_t result = 0;
    if (dim == 0) {
        result = brev(rev);
    } else {
        const SobolTable& T = sobol_table();
        uint32_t i = 0;
        while (rev != 0) {
            uint32_t j = clz32(rev);
            result ^= T.t[dim][i + j];
            i += j + 1;
            rev <<= j;
            rev <<= 1;
        }
    }
    result = brev(reversed_bit_owen(result, seed));
    return u2unit(result);
}
// src/device/quasi_random.h:76-99
inline float sobol1(const QState& s, uint32_t d) {
    uint32_t seed = s.hash ^ dim_hash(s.offset + d);
    uint32_t idx = reversed_bit_owen(brev(s.sample), seed ^ 0xbff95bfeu);
    return sobol_burley(idx, 0, seed ^ 0x635c77bdu);
}
inline V2 sobol2(const QState& s, uint32_t d) {
    uint32_t seed = s.hash ^ dim_hash(s.offset + d);
    uint32_t idx = reversed_bit_owen(brev(s.sample), seed ^ 0xf8ade99au);
    return {sobol_burley(idx, 0, seed ^ 0xe0aaaf76u), sobol_burley(idx, 1, seed ^ 0x94964d4eu)};
}
inline V3 sobol3(const QState& s, uint32_t d) {
    uint32_t seed = s.hash ^ dim_hash(s.offset + d);
    uint32_t idx = reversed_bit_owen(brev(s.sample), seed ^ 0xcaa726acu);
    return {sobol_burley(idx, 0, seed ^ 0x9e78e391u), sobol_burley(idx, 1, seed ^ 0x67c33241u),
            sobol_burley(idx, 2, seed ^ 0x78c395c5u)};
}

// --------------------------------------------------------------- spectral --
struct Spec { float v[NL]; };
inline Spec specs(float s) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = s; return r; }
inline Spec operator+(const Spec& a, const Spec& b) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] + b.v[i]; return r; }
inline Spec operator*(const Spec& a, const Spec& b) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] * b.v[i]; return r; }
inline Spec operator/(const Spec& a, const Spec& b) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] / b.v[i]; return r; }
inline Spec operator*(const Spec& a, float s) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] * s; return r; }
inline Spec operator*(float s, const Spec& a) { return a * s; }
inline Spec operator/(const Spec& a, float s) { Spec r; for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] / s; return r; }
inline Spec& operator+=(Spec& a, const Spec& b) { for (int i = 0; i < NL; ++i) a.v[i] += b.v[i]; return a; }
inline Spec& operator*=(Spec& a, const Spec& b) { for (int i = 0; i < NL; ++i) a.v[i] *= b.v[i]; return a; }
inline Spec& operator/=(Spec& a, float s) { for (int i = 0; i < NL; ++i) a.v[i] /= s; return a; }
// src/device/spectral.h:137-175
inline float smax(const Spec& a) { float t = a.v[0]; for (int i = 1; i < NL; ++i) if (a.v[i] > t) t = a.v[i]; return t; }
inline float savg(const Spec& a) { float t = a.v[0]; for (int i = 1; i < NL; ++i) t += a.v[i]; return t / NL; }

// src/device/spectral.h:196-301
struct Wavelengths {
    Spec lambda, pdf;
    float active;
};
// src/device/spectral.h:289-299
inline float sample_visible_wl(float u) { return 538.0f - 138.888889f * std::atanh(0.85691062f - 1.82750197f * u); }
inline float visible_wl_pdf(float l) {
    if (l < 360.0f || l > 830.0f) return 0.0f;
    float c = std::cosh(0.0072f * (l - 538.0f));
    return 0.0039398042f / (c * c);
}
// src/device/spectral.h:227-241
inline Wavelengths sample_visible(float u) {
    Wavelengths w;
    for (int i = 0; i < NL; ++i) {
        float up = u + (float)i / NL;
        if (up > 1.0f) up -= 1.0f;
        w.lambda.v[i] = sample_visible_wl(up);
        w.pdf.v[i] = visible_wl_pdf(w.lambda.v[i]);
    }
    w.active = (float)NL;
    return w;
}
// SampledWavelengths::terminateSecondary / terminateAllSecondary
// (src/device/spectral.h:250-276): the remaining pdfs are rescaled by
// (N / active) * (active' / N) at each termination.
inline void wl_terminate_secondary(Wavelengths& w, int i) {
    if (w.pdf.v[i] == 0.0f) return;
    w.pdf.v[i] = 0.0f;
    float m = (float)NL / w.active;
    w.active -= 1.0f;
    m *= w.active / (float)NL;
    for (int k = 0; k < NL; ++k)
        if (w.pdf.v[k] != 0.0f) w.pdf.v[k] *= m;
}
inline bool wl_all_secondary_terminated(const Wavelengths& w) {
    for (int k = 1; k < NL; ++k)
        if (w.pdf.v[k] != 0.0f) return false;
    return true;
}
inline void wl_terminate_all_secondary(Wavelengths& w) {
    if (wl_all_secondary_terminated(w)) return;
    float m = (float)NL / w.active;
    w.active = 1.0f;
    m *= w.active / (float)NL;
    w.pdf.v[0] *= m;
    for (int k = 1; k < NL; ++k) w.pdf.v[k] = 0.0f;
}
// src/device/spectral.h:304-327 (lroundf index into the 1 nm table)
inline float dense_sample(const float* tab, float l) { return tab[std::lround(l) - LMIN]; }
inline Spec dense_sample(const float* tab, const Wavelengths& w) {
    Spec r; for (int i = 0; i < NL; ++i) r.v[i] = dense_sample(tab, w.lambda.v[i]); return r;
}
// src/device/color.h:369-381
inline Spec rgb_to_spectrum(V3 rgb, const Wavelengths& w, const float* bR, const float* bG, const float* bB) {
    Spec s;
    for (int i = 0; i < NL; ++i) {
        float l = w.lambda.v[i];
        s.v[i] = dense_sample(bR, l) * rgb.x + dense_sample(bG, l) * rgb.y + dense_sample(bB, l) * rgb.z;
    }
    return s;
}
// src/device/color.h:10-16 (sensor curves A/B/C are CIE X/Y/Z)
inline void to_xyz(const float* cA, const float* cB, const float* cC, const Wavelengths& w, const Spec& L, double out[3]) {
    out[0] = (double)savg(dense_sample(cA, w) * L);
    out[1] = (double)savg(dense_sample(cB, w) * L);
    out[2] = (double)savg(dense_sample(cC, w) * L);
}
// src/device/color.h:18-23
inline float srgb_to_linear(float c) { return c < 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f); }

// ----------------------------------------------------------- octahedral --
// src/core/math_util.h:147-162
inline V2 oct_encode(V3 v) {
    float inv = 1.0f / (std::fabs(v.x) + std::fabs(v.y) + std::fabs(v.z));
    V2 p{v.x * inv, v.y * inv};
    if (v.z <= 0.0f)
        return {(1.0f - std::fabs(p.y)) * (p.x >= 0.0f ? 1.0f : -1.0f), (1.0f - std::fabs(p.x)) * (p.y >= 0.0f ? 1.0f : -1.0f)};
    return p;
}
// src/core/math_util.h:171-187
inline V3 oct_decode(V2 e) {
    V3 v{e.x, e.y, 1.0f - std::fabs(e.x) - std::fabs(e.y)};
    float t = std::max(0.0f, -v.z);
    v.x += v.x >= 0.0f ? -t : t;
    v.y += v.y >= 0.0f ? -t : t;
    return normalize(v);
}
// src/core/math_util.h:199-207
inline uint32_t oct_encode_u32(V3 v) {
    const float m = 65535.99f;
    V2 p = oct_encode(v);
    uint32_t x = (uint32_t)(clampf(p.x * 0.5f + 0.5f, 0.0f, 1.0f) * m);
    uint32_t y = (uint32_t)(clampf(p.y * 0.5f + 0.5f, 0.0f, 1.0f) * m) << 16;
    return y | x;
}
// src/core/math_util.h:212-219
inline V3 oct_decode_u32(uint32_t e) {
    const float n = 1.0f / 65535.0f;
    V2 p{(float)(e & 0xFFFFu) * n, (float)(e >> 16) * n};
    return oct_decode({p.x * 2.0f - 1.0f, p.y * 2.0f - 1.0f});
}

// ------------------------------------------------------ spherical Gaussians --
// src/core/math_util.h:246-284 (Higham's cancellation-free forms)
inline float sg_expm1(float x) {
    float u = std::exp(x);
    if (u == 1.0f) return x;
    float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return y * x / std::log(u);
    return y;
}
inline float sg_expm1_over_x(float x) {
    float u = std::exp(x);
    if (u == 1.0f) return 1.0f;
    float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return y / std::log(u);
    return y / x;
}
// erf for the spherical-Gaussian integrals: a piecewise Chebyshev fit made
// for this project (erf(x)/x as a degree-7 polynomial in x^2 on [0, 1];
// erf(x) as a degree-16 polynomial in x - 2.5 on [1, 4]; +-1 beyond 4, where
// the reference also saturates). Max abs error 1.7e-7 against the exact erf.
// Horner steps are explicit FMAs, so every implementation evaluates it to the
// same bits for the same input, which keeps the cancellation in
// erfc = 1 - erf (math_util.h:326-329) reproducible.
inline float sg_erf(float x) {
    static const float A[8] = {1.128379225730896f, -0.37612637877464294f, 0.11283782869577408f, -0.026865437626838684f, 0.005221053492277861f, -0.0008484396385028958f, 0.00011268253729213029f, -9.673591193859465e-06f};
    static const float B[17] = {0.9995930194854736f, 0.0021782785188406706f, -0.005445701535791159f, 0.008350212126970291f, -0.008622484281659126f, 0.006116587668657303f, -0.002797981258481741f, 0.0005445338319987059f, 0.00025947490939870477f, -0.0002534241648390889f, 8.039429667405784e-05f, 7.756445484119467e-06f, -1.5348350643762387e-05f, 3.462653694441542e-06f, 7.886890216468601e-07f, -3.5129914977005683e-07f, 1.1272015321139861e-08f};
    const float ax = std::fabs(x);
    if (ax >= 4.0f) return std::copysign(1.0f, x);
    float r;
    if (ax <= 1.0f) {
        const float u = ax * ax;
        float h = A[7];
        for (int k = 6; k >= 0; --k) h = std::fma(h, u, A[k]);
        r = ax * h;
    } else {
        const float y = ax - 2.5f;
        float h = B[16];
        for (int k = 15; k >= 0; --k) h = std::fma(h, y, B[k]);
        r = h;
    }
    return std::copysign(r, x);
}
inline float sg_erfc(float x) { return 1.0f - sg_erf(x); }

struct SGLobe { V3 axis; float sharpness, logAmplitude; };
// src/core/math_util.h:349-352
inline float sg_integral(float s) { return 4.0f * 3.14159265359f * sg_expm1_over_x(-2.0f * s); }
// src/core/math_util.h:362-375
inline SGLobe sg_product(V3 a1, float s1, V3 a2, float s2) {
    V3 axis = a1 * s1 + a2 * s2;
    float sharp = length(axis);
    float cosine = clampf(dot(a1, a2), -1.0f, 1.0f);
    float smin = std::min(s1, s2);
    float ratio = smin / std::max(std::max(s1, s2), FLT_MIN);
    float logA = 2.0f * smin * (cosine - 1.0f) / (std::sqrt(2.0f * ratio * cosine + ratio * ratio + 1.0f) + ratio + 1.0f);
    return {axis / std::max(sharp, FLT_MIN), sharp, logA};
}
// src/core/math_util.h:469-487
inline float sg_upper_cc_over_2pi(float s) {
    if (s <= 0.5f)
        return (((((((-1.0f / 362880.0f) * s + 1.0f / 40320.0f) * s - 1.0f / 5040.0f) * s + 1.0f / 720.0f) * s - 1.0f / 120.0f) * s + 1.0f / 24.0f) * s - 1.0f / 6.0f) * s + 0.5f;
    return (sg_expm1(-s) + s) / (s * s);
}
// src/core/math_util.h:490-510
inline float sg_lower_cc_over_2pi(float s) {
    float e = std::exp(-s);
    if (s <= 0.5f)
        return e * (((((((((1.0f / 403200.0f) * s - 1.0f / 45360.0f) * s + 1.0f / 5760.0f) * s - 1.0f / 840.0f) * s + 1.0f / 144.0f) * s - 1.0f / 30.0f) * s + 1.0f / 8.0f) * s - 1.0f / 3.0f) * s + 0.5f);
    return e * (-sg_expm1(-s) - s * e) / (s * s);
}
// Tokuyoshi et al. 2024, supplementary Listing 7 (src/core/math_util.h:514-538)
inline float sg_clamped_cosine_product_integral_over_pi_2024(float cosine, float s) {
    const float A = 2.7360831611272558028247203765204f, B = 17.02129778174187535455530451145f;
    const float C = 4.0100826728510421403939290030394f, D = 15.219156263147210594866010069381f;
    const float E = 76.087896272360737270901154261082f;
    float t = s * std::sqrt(0.5f * ((s + A) * s + B) / (((s + C) * s + D) * s + E));
    float tz = t * cosine;
    const float INV_SQRTPI = 0.56418958354775628694807945156077f;
    const float CLAMP = 0.5f * FLT_EPSILON;
    float lerp = clampf(std::max(0.5f * (cosine * sg_erfc(-tz) + sg_erfc(t)) -
                                     0.5f * INV_SQRTPI * std::exp(-tz * tz) * sg_expm1(t * t * (cosine * cosine - 1.0f)) / t,
                                 CLAMP),
                        0.0f, 1.0f);
    float lo = sg_lower_cc_over_2pi(s), up = sg_upper_cc_over_2pi(s);
    return 2.0f * mixf(lo, up, lerp);
}
// Banerjee et al. 2005 (src/core/math_util.h:562-565)
inline float vmf_axis_length_to_sharpness(float l) { return l * (3.0f - l * l) / (1.0f - l * l); }
// Peters 2016 cubic solve (src/core/math_util.h:568-587)
inline float vmf_sharpness_to_axis_length(float s) {
    float a = s / 3.0f;
    float b = a * a * a;
    float c = std::sqrt(1.0f + 3.0f * (a * a) * (1.0f + a * a));
    float th = std::atan2(c, b) / 3.0f;
    const float SQRT3 = 1.7320508075688772935274463415059f;
    float d = std::sin(th) * SQRT3 - std::cos(th);
    return (s > 0x1.0p25f) ? 1.0f : std::sqrt(1.0f + a * a) * d + a;
}

// ------------------------------------------------------------ light tree --
// src/core/light_tree_types.h:10-18, 50-55, 191-207
// LT_SHAPE: a disk or sphere light (the free fourth value of the 2-bit type
// field); its pointer index is its light index, after the triangle lights.
enum LightType : uint32_t { LT_TRIANGLE = 0, LT_SKY = 1, LT_NONE = 2, LT_SHAPE = 3 };
constexpr uint32_t KLightCountBits = 4, KLightOffsetBits = 27;
constexpr uint32_t KMaxLeafLightCount = 15, KMaxLeafLightOffset = (1u << 27) - 1u;
constexpr float KEnvImportance = 0.3f;
inline uint32_t lptr_pack(uint32_t type, uint32_t idx) { return (type << 30) | (idx & ((1u << 30) - 1u)); }
inline void lptr_unpack(uint32_t p, uint32_t& type, uint32_t& idx) { type = p >> 30; idx = p & ((1u << 30) - 1u); }

// src/core/light_tree_types.h:119-162 (32 bytes)
struct PackedNode {
    float spatialMean[3];
    float spatialVariance;
    uint32_t packedAverageDirection;
    float sharpness;
    float flux;
    uint32_t coreData;
};
static_assert(sizeof(PackedNode) == 32, "PackedNode layout");
struct NodeAttr { float mean[3], variance, flux, dir[3], sharpness; };
struct UnpackedNode {
    NodeAttr a;
    bool leaf;
    uint32_t right, count, offset;
};
// src/core/light_tree_types.h:165-189. The packed direction is decoded from
// [0,1]-scaled components without the *2-1 remap that decodeU32 applies; that
// is what the reference does and it is kept here.
inline UnpackedNode unpack_node(const PackedNode& p) {
    UnpackedNode r{};
    r.a.mean[0] = p.spatialMean[0]; r.a.mean[1] = p.spatialMean[1]; r.a.mean[2] = p.spatialMean[2];
    r.a.variance = p.spatialVariance;
    r.a.flux = p.flux;
    r.a.sharpness = p.sharpness;
    V3 d = oct_decode({(float)(p.packedAverageDirection & 0xFFFFu) / 65535.0f, (float)(p.packedAverageDirection >> 16) / 65535.0f});
    r.a.dir[0] = d.x; r.a.dir[1] = d.y; r.a.dir[2] = d.z;
    r.leaf = (p.coreData >> 31) != 0;
    if (r.leaf) {
        r.count = (p.coreData >> KLightOffsetBits) & ((1u << KLightCountBits) - 1u);
        r.offset = p.coreData & ((1u << KLightOffsetBits) - 1u);
    } else {
        r.right = p.coreData;
    }
    return r;
}

// src/kernels/pathtrace.cu:974-1001
inline float node_importance(const UnpackedNode& n, V3 pos, V3 nrm) {
    V3 mean{n.a.mean[0], n.a.mean[1], n.a.mean[2]};
    V3 adir{n.a.dir[0], n.a.dir[1], n.a.dir[2]};
    V3 ldir = mean - pos;
    float d2 = dot(ldir, ldir);
    ldir = ldir * (1.0f / std::sqrt(d2));
    float var = std::max(n.a.variance, 0x1.0p-31f * d2);
    float lsharp = d2 / var;
    SGLobe lobe = sg_product(adir, n.a.sharpness, ldir, lsharp);
    float amp = std::exp(lobe.logAmplitude);
    float cosine = clampf(dot(lobe.axis, nrm), -1.0f, 1.0f);
    return (n.a.flux / (var * sg_integral(n.a.sharpness))) * amp *
           sg_clamped_cosine_product_integral_over_pi_2024(cosine, lobe.sharpness);
}

struct LightTreeView {
    const PackedNode* nodes = nullptr;   // nullptr = no tree
    const uint32_t* lightPointers = nullptr;
    const uint64_t* bitmasks = nullptr;
    bool envEnabled = false;
};

// src/kernels/pathtrace.cu:1002-1049
inline float light_tree_pdf(const LightTreeView& t, uint32_t type, uint32_t index, V3 pos, V3 nrm) {
    float envImp = t.envEnabled ? KEnvImportance : 0.0f;
    if (type == LT_SKY) return envImp;
    float pdf = 1.0f - envImp;
    uint64_t mask = t.bitmasks[index];
    uint32_t cur = 0;
    UnpackedNode c = unpack_node(t.nodes[0]);
    for (int depth = 0; depth < 64; ++depth) {
        if (c.leaf) break;
        UnpackedNode l = unpack_node(t.nodes[cur + 1]);
        float li = node_importance(l, pos, nrm);
        UnpackedNode r = unpack_node(t.nodes[c.right]);
        float ri = node_importance(r, pos, nrm);
        bool right = (mask & (1ull << depth)) != 0;
        cur = right ? c.right : cur + 1;
        pdf *= right ? ri / (li + ri) : li / (li + ri);
        c = right ? r : l;
    }
    pdf *= 1.0f / (float)c.count;
    return pdf;
}

// src/kernels/pathtrace.cu:1050-1112
inline void light_tree_sample(const LightTreeView& t, float rnd, V3 pos, V3 nrm, uint32_t& type, uint32_t& index, float& pdf) {
    auto normRand = [](float r, float sep, bool right) {
        float res = right ? (r - sep) / (1.0f - sep) : r / sep;
        return clampf(res, 0.0f, 1.0f);
    };
    bool envPresent = t.envEnabled;
    bool treePresent = t.nodes != nullptr;
    float envImp = envPresent ? (treePresent ? KEnvImportance : 1.0f) : 0.0f;
    if (rnd < envImp) { type = LT_SKY; index = 0; pdf = envImp; return; }
    rnd = normRand(rnd, envImp, true);
    pdf = 1.0f - envImp;
    uint32_t cur = 0;
    UnpackedNode c = unpack_node(t.nodes[0]);
    while (!c.leaf) {
        UnpackedNode l = unpack_node(t.nodes[cur + 1]);
        UnpackedNode r = unpack_node(t.nodes[c.right]);
        float li = node_importance(l, pos, nrm);
        float ri = node_importance(r, pos, nrm);
        float sep = li / (li + ri);
        bool right = rnd > sep;
        cur = right ? c.right : cur + 1;
        pdf *= right ? 1.0f - sep : sep;
        rnd = normRand(rnd, sep, right);
        c = right ? r : l;
    }
    uint32_t k = (uint32_t)((float)c.count * rnd);
    uint32_t slot = c.offset + (k == c.count ? c.count - 1 : k);
    pdf *= 1.0f / (float)c.count;
    lptr_unpack(t.lightPointers[slot], type, index);
    if (std::isnan(pdf) || std::isinf(pdf) || pdf == 0.0f) type = LT_NONE;
}

// ---------------------------------------------------------------- frames --
// src/device/local_transform.h:263-277
inline float cos2t(V3 w) { return w.z * w.z; }
inline float sin2t(V3 w) { return std::max(0.0f, 1.0f - cos2t(w)); }
inline float tan2t(V3 w) { return sin2t(w) / cos2t(w); }
inline float sint(V3 w) { return std::sqrt(std::max(0.0f, 1.0f - cos2t(w))); }
inline float cosphi(V3 w) { float s = sint(w); return s == 0.0f ? 1.0f : clampf(w.x / s, -1.0f, 1.0f); }
inline float sinphi(V3 w) { float s = sint(w); return s == 0.0f ? 0.0f : clampf(w.y / s, -1.0f, 1.0f); }

// src/device/local_transform.h:283-293: columns T, B, N
inline M3 frame_from_normal(V3 n) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return M3{{{1.0f + sign * (n.x * n.x) * a, sign * b, -sign * n.x}, {b, sign + (n.y * n.y) * a, -n.y}, n}};
}
// LocalTransform: toLocal = transpose(frame) * v, fromLocal = frame * v
struct Local {
    M3 f;  // columns T, B, N
    V3 to_local(V3 v) const { return {dot(f.c[0], v), dot(f.c[1], v), dot(f.c[2], v)}; }
    V3 from_local(V3 v) const { return mul(f, v); }
};

// ------------------------------------------------------------- sampling --
// src/device/sampling.h:16-30
inline V3 tri_sample(V2 uv, V3 a, V3 b, V3 c, float& ba, float& bb, float& bc) {
    if (uv.x + uv.y > 1.0f) uv = {1.0f - uv.x, 1.0f - uv.y};
    ba = 1.0f - uv.x - uv.y; bb = uv.x; bc = uv.y;
    return a + (b - a) * uv.x + (c - a) * uv.y;
}
// src/device/sampling.h:53-58
inline V2 disk_polar(V2 uv) {
    float r = std::sqrt(uv.x);
    float th = 2.0f * PI * uv.y;
    return {r * std::cos(th), r * std::sin(th)};
}
// src/device/util.h:143-158 (Wächter & Binder offset with the reference's constants)
inline V3 offset_point(V3 p, V3 n) {
    const float originInterval = 1.0f / 32.0f, floatScale = 1.0f / 65536.0f, intScale = 256.0f;
    int32_t ox = (int32_t)(intScale * n.x), oy = (int32_t)(intScale * n.y), oz = (int32_t)(intScale * n.z);
    V3 pi{i2f(f2i(p.x) + (p.x < 0.0f ? -ox : ox)), i2f(f2i(p.y) + (p.y < 0.0f ? -oy : oy)),
          i2f(f2i(p.z) + (p.z < 0.0f ? -oz : oz))};
    return {std::fabs(p.x) < originInterval ? p.x + floatScale * n.x : pi.x,
            std::fabs(p.y) < originInterval ? p.y + floatScale * n.y : pi.y,
            std::fabs(p.z) < originInterval ? p.z + floatScale * n.z : pi.z};
}
// src/device/util.h:177-180
inline V3 reflect(V3 v, V3 n) { return -v + 2.0f * dot(v, n) * n; }
// src/device/util.h:71-96
inline V3 refract_rel(V3 v, V3 n, float eta, bool& valid, float& etaRel) {
    float ci = dot(v, n);
    if (ci < 0.0f) { eta = 1.0f / eta; ci = -ci; n = -n; }
    float s2i = std::max(0.0f, 1.0f - ci * ci);
    float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) { valid = false; return {0.0f, 0.0f, 0.0f}; }
    valid = true;
    etaRel = eta;
    return -v / eta + (ci / eta - std::sqrt(1.0f - s2t)) * n;
}
// microfacet::FReal, src/device/surface.h:197-216
inline float fresnel_real(float c, float eta) {
    c = clampf(c, -1.0f, 1.0f);
    if (c < 0.0f) { eta = 1.0f / eta; c = -c; }
    float s2 = 1.0f - c * c;
    float s2t = s2 / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    float ct = std::sqrt(std::max(0.0f, 1.0f - s2t));
    float rParl = (eta * c - ct) / (eta * c + ct);
    float rPerp = (c - eta * ct) / (c + eta * ct);
    return (rParl * rParl + rPerp * rPerp) / 2.0f;
}
// src/device/util.h:181-198
inline V3 refract(V3 v, V3 n, float eta) {
    float ci = dot(v, n);
    if (ci < 0.0f) { ci = -ci; n = -n; }
    float s2i = std::max(0.0f, 1.0f - ci * ci);
    float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) return {0.0f, 0.0f, 0.0f};
    return -v / eta + (ci / eta - std::sqrt(1.0f - s2t)) * n;
}
// src/device/filter.h:14-27
inline V2 gaussian_filter_sample(V2 uv) {
    const float rad = 1.5f, sigma = 0.25f;
    const float uTH = std::exp(-(rad * rad) / (2.0f * sigma * sigma));
    float u = std::max(0.00001f, uv.x * (1.0f - uTH) + uTH);
    float m = sigma * std::sqrt(-2.0f * std::log(u));
    return {m * std::cos(2.0f * PI * uv.y), m * std::sin(2.0f * PI * uv.y)};
}
// src/device/mis.h:14-21, with PBRT-v4's guard against an infinite f^2
// (weight 1 instead of inf/inf = NaN; the caller then divides by the
// infinite pdf and the sample contributes 0)
inline float power_heuristic(float f, float g) {
    float f2 = f * f, g2 = g * g;
    if (std::isinf(f2)) return 1.0f;
    return f2 / (f2 + g2);
}

// ------------------------------------------------------------ microfacet --
// src/device/surface.h:125-136
struct Alpha {
    float x, y;
    void regularize() {
        if (x < 0.3f) x = clampf(2.0f * x, 0.1f, 0.3f);
        if (y < 0.3f) y = clampf(2.0f * y, 0.1f, 0.3f);
    }
};
struct CtxW { float tan2, cphi, sphi; };
struct CtxM { float cos2, tan2, cphi, sphi; };
inline CtxW ctx_w(V3 w) { return {tan2t(w), cosphi(w), sinphi(w)}; }
inline CtxM ctx_m(V3 m) { return {cos2t(m), tan2t(m), cosphi(m), sinphi(m)}; }
// src/device/surface.h:138-146
inline float lambda_g(const Alpha& a, float cphi, float sphi, float tan2) {
    if (std::isinf(tan2)) return 0.0f;
    float ax = cphi * a.x, ay = sphi * a.y;
    float a2 = ax * ax + ay * ay;
    return (std::sqrt(1.0f + a2 * tan2) - 1.0f) / 2.0f;
}
// src/device/surface.h:148-159
inline float ggx_G1(const CtxW& o, const Alpha& a) { return 1.0f / (1.0f + lambda_g(a, o.cphi, o.sphi, o.tan2)); }
inline float ggx_G(const CtxW& i, const CtxW& o, const Alpha& a) {
    return 1.0f / (1.0f + lambda_g(a, i.cphi, i.sphi, i.tan2) + lambda_g(a, o.cphi, o.sphi, o.tan2));
}
// src/device/surface.h:161-174
inline float ggx_D(const CtxM& m, const Alpha& a) {
    if (std::isinf(m.tan2)) return 0.0f;
    float c4 = m.cos2 * m.cos2;
    if (c4 < 1e-16f) return 0.0f;
    float cx = m.cphi / a.x, sy = m.sphi / a.y;
    float e = m.tan2 * (cx * cx + sy * sy);
    float ope = 1.0f + e;
    return 1.0f / (PI * a.x * a.y * c4 * ope * ope);
}
// src/device/surface.h:176-195 (complex Fresnel, per wavelength)
inline float fresnel_complex1(float ct, float er, float ei) {
    ct = clampf(ct, 0.0f, 1.0f);
    float s2 = 1.0f - ct * ct;
    // eta^2
    float e2r = er * er - ei * ei, e2i = 2.0f * er * ei;
    // sin2T = s2 / eta^2
    float den = e2r * e2r + e2i * e2i;
    float s2tr = s2 * e2r / den, s2ti = -s2 * e2i / den;
    // cosT = sqrt(1 - sin2T) (principal branch)
    float zr = 1.0f - s2tr, zi = -s2ti;
    float mag = std::sqrt(std::sqrt(zr * zr + zi * zi));
    float ang = 0.5f * std::atan2(zi, zr);
    float ctr = mag * std::cos(ang), cti = mag * std::sin(ang);
    // rParl = (eta cos - cosT) / (eta cos + cosT)
    float ar = er * ct - ctr, ai = ei * ct - cti, br = er * ct + ctr, bi = ei * ct + cti;
    float rpa = (ar * ar + ai * ai) / (br * br + bi * bi);
    // rPerp = (cos - eta cosT) / (cos + eta cosT)
    float ecr = er * ctr - ei * cti, eci = er * cti + ei * ctr;
    float cr = ct - ecr, ci = -eci, dr = ct + ecr, di = eci;
    float rpe = (cr * cr + ci * ci) / (dr * dr + di * di);
    return (rpa + rpe) / 2.0f;
}
// (1 - mu)^5 of surface.h:245-252 (the reference's pow(x, 5)), evaluated as
// a double product rounded once, the same bits on the CPU and the GPU
inline float pow5(float x) {
    const double a = (double)x * (double)x;
    return (float)(a * a * (double)x);
}
// src/device/surface.h:245-252
inline float schlick(float f0, float mu) { return f0 + (1.0f - f0) * pow5(1.0f - mu); }
inline V3 schlick3(V3 f0, float mu) { float p = pow5(1.0f - mu); return f0 + (v3s(1.0f) - f0) * p; }
// src/device/surface.h:254-264 (specularColor = 1)
inline V3 f82(V3 f0, float mu) {
    const float muD = 1.0f / 7.0f;
    V3 f = schlick3(f0, mu);
    V3 fD = schlick3(f0, muD);
    float k = (mu * std::pow(1.0f - mu, 6.0f)) / (muD * std::pow(1.0f - muD, 6.0f));
    return f - k * (fD - fD);
}
// src/device/surface.h:265-269
inline V3 f_avg_f82(V3 f0) { return 0.877105f * f0 + (v3s(0.0648148f) + 0.0752755f * f0) - v3s(0.0171958f); }

// src/device/surface.h:308-322 (Dupuy & Benyoub 2023 spherical-cap VNDF)
inline V3 vndf_sample_cap(V3 wo, const Alpha& a, V2 uv) {
    V3 ws = normalize(V3{wo.x * a.x, wo.y * a.y, wo.z});
    if (ws.z < 0.0f) ws = -ws;
    float phi = 2.0f * PI * uv.x;
    float b = ws.z;
    float z = std::fma(1.0f - uv.y, 1.0f + b, -b);
    float st = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
    V3 wi{st * std::cos(phi), st * std::sin(phi), z};
    V3 m = ws + wi;
    return normalize(V3{m.x * a.x, m.y * a.y, m.z});
}
// src/device/surface.h:323-340 (Eto & Tokuyoshi bounded VNDF)
inline V3 vndf_sample_bounded(V3 wo, const Alpha& a, V2 uv) {
    V3 ws = normalize(V3{wo.x * a.x, wo.y * a.y, wo.z});
    float phi = 2.0f * PI * uv.x;
    float a1 = std::min(a.x, a.y);
    float s = 1.0f + std::sqrt(wo.x * wo.x + wo.y * wo.y);
    float a2 = a1 * a1, s2 = s * s;
    float k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
    float b = wo.z > 0.0f ? k * ws.z : ws.z;
    float z = std::fma(1.0f - uv.y, 1.0f + b, -b);
    float st = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
    V3 wi{st * std::cos(phi), st * std::sin(phi), z};
    V3 m = ws + wi;
    return normalize(V3{m.x * a.x, m.y * a.y, m.z});
}
// src/device/surface.h:352-355
inline float vndf_pdf_cap(V3 wo, float absWoWm, const CtxW& o, const CtxM& m, const Alpha& a) {
    return ggx_G1(o, a) / std::fabs(wo.z) * ggx_D(m, a) * absWoWm;
}
// src/device/surface.h:356-372
inline float vndf_pdf_bounded(V3 wo, float absWoWm, const CtxW& o, const CtxM& m, const Alpha& a) {
    (void)o;
    float ndf = ggx_D(m, a);
    float aox = wo.x * a.x, aoy = wo.y * a.y;
    float len2 = aox * aox + aoy * aoy;
    float t = std::sqrt(len2 + wo.z * wo.z);
    if (wo.z >= 0.0f) {
        float a1 = clampf(std::min(a.x, a.y), 0.0f, 1.0f);
        float s = 1.0f + std::sqrt(wo.x * wo.x + wo.y * wo.y);
        float a2 = a1 * a1, s2 = s * s;
        float k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
        return ndf * 2.0f * absWoWm / (k * wo.z + t);
    }
    return ndf * (t - wo.z) * 2.0f * absWoWm / len2;
}

// src/device/surface.h:22-31
inline V3 cos_sample(V2 uv) {
    V2 d = disk_polar(uv);
    return {d.x, d.y, std::sqrt(std::max(0.0f, 1.0f - d.x * d.x - d.y * d.y))};
}
inline float cos_pdf(float ct) { return ct / PI; }

// src/device/surface.h:53-88 (Zeltner 2022 LTC sheen)
inline V3 ltc_sample(V3 wo, float A, float B, V2 uv) {
    V3 w0 = cos_sample(uv);
    V3 ws{w0.x / A - w0.z * B / A, w0.y / A, w0.z};
    float ct = wo.z;
    float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct));
    float cp = st == 0.0f ? 1.0f : clampf(wo.x / st, -1.0f, 1.0f);
    float sp = st == 0.0f ? 0.0f : clampf(wo.y / st, -1.0f, 1.0f);
    V3 wi{cp * ws.x - sp * ws.y, cp * ws.y + sp * ws.x, ws.z};
    return normalize(wo.z > 0.0f ? wi : -wi);
}
inline void ltc_eval(V3 wo, V3 wi, float A, float B, float& f, float& pdf) {
    float ct = wo.z;
    float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct));
    float cp = st == 0.0f ? 1.0f : clampf(wo.x / st, -1.0f, 1.0f);
    float sp = st == 0.0f ? 0.0f : clampf(wo.y / st, -1.0f, 1.0f);
    V3 ws{cp * wi.x + sp * wi.y, cp * wi.y - sp * wi.x, wi.z};
    V3 w0{A * ws.x + B * ws.z, A * ws.y, ws.z};
    float len = length(w0);
    w0 = w0 / len;
    float det = A * A;
    float jac = det / (len * len * len);
    float r = cos_pdf(w0.z) * jac;
    f = r / std::fabs(wi.z);
    pdf = r;
}

}  // namespace orc
