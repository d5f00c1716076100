// Owen-scrambled Sobol sampling (Burley 2020), as used by the reference
// (src/device/quasi_random.h:10-204, itself after Cycles).
//
// Per-path QRNG state is not stored: the sample index is the launch's
// sample number and the dimension offset is 16 * (bounce + 1)
// (State::advanceBounce, quasi_random.h:60-63), so both are recomputed from
// the path depth.
#pragma once
#include "vec.hpp"

namespace rbh {
namespace dev {

enum : uint32_t {
    QD_FILTER = 0, QD_LENS = 1, QD_WAVELENGTH = 0, QD_LIGHT = 1, QD_BXDF0 = 2, QD_BXDF1 = 3,
    QD_ALPHA_BLEND = 7, QD_ROULETTE = 8, QD_STRIDE = 16,
    QD_GUIDE = 10,  // guiding lobe choice and direction (dimensions unused by the reference)
};

// Sobol direction numbers of dimensions 0..3, bit-reversed, generated at
// compile time from the Joe-Kuo recurrence (d1: x+1; d2: x^2+x+1 m={1,3};
// d3: x^3+x+1 m={1,3,1}).
struct SobolTab { uint32_t v[4][32]; };

constexpr uint32_t brev_c(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) { r = (r << 1) | (x & 1u); x >>= 1; }
    return r;
}
constexpr SobolTab make_sobol_tab() {
    SobolTab T{};
    const int s[4] = {0, 1, 2, 3};
    const uint32_t a[4] = {0, 0, 1, 1};
    const uint32_t m[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int d = 0; d < 4; ++d) {
        uint32_t v[33] = {};
        if (d == 0) {
            for (int k = 1; k <= 32; ++k) v[k] = 1u << (32 - k);
        } else {
            for (int k = 1; k <= s[d]; ++k) v[k] = m[d][k - 1] << (32 - k);
            for (int k = s[d] + 1; k <= 32; ++k) {
                uint32_t x = v[k - s[d]] ^ (v[k - s[d]] >> s[d]);
                for (int j = 1; j < s[d]; ++j)
                    if ((a[d] >> (s[d] - 1 - j)) & 1u) x ^= v[k - j];
                v[k] = x;
            }
        }
        for (int i = 0; i < 32; ++i) T.v[d][i] = brev_c(v[i + 1]);
    }
    return T;
}
__constant__ constexpr SobolTab kSobol = make_sobol_tab();

// quasi_random.h:30-37
__device__ __forceinline__ uint32_t pixel_hash(uint32_t x, uint32_t y) {
    const uint32_t qx = 1103515245U * ((x >> 1U) ^ y);
    const uint32_t qy = 1103515245U * ((y >> 1U) ^ x);
    return 1103515245U * (qx ^ (qy >> 3U));
}
// quasi_random.h:101-110
__device__ __forceinline__ uint32_t dim_hash(uint32_t i) {
    i ^= i >> 16; i *= 0x21f0aaadu; i ^= i >> 15; i *= 0xd35a2d97u; i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}
// quasi_random.h:119-127 (Laine-Karras style hash in the bit-reversed domain)
__device__ __forceinline__ uint32_t owen_rev(uint32_t n, uint32_t seed) {
    n ^= n * 0x3d20adeau;
    n += seed;
    n *= (seed >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}
__device__ __forceinline__ float u32_to_unit(uint32_t x) { return (float)x * (1.0f / 4294967808.0f); }

// quasi_random.h:132-204
__device__ __forceinline__ float sobol_dim(uint32_t rev, uint32_t dim, uint32_t seed) {
    uint32_t r;
    if (dim == 0) {
        r = __builtin_bitreverse32(rev);
    } else {
        r = 0;
        uint32_t i = 0;
        while (rev != 0) {
            const uint32_t j = __builtin_clz(rev);
            r ^= kSobol.v[dim][i + j];
            i += j + 1;
            rev = (rev << j) << 1;
        }
    }
    return u32_to_unit(__builtin_bitreverse32(owen_rev(r, seed)));
}

struct Qrng {
    uint32_t sample, offset, hash;
    __device__ __forceinline__ float s1(uint32_t d) const {
        const uint32_t seed = hash ^ dim_hash(offset + d);
        const uint32_t idx = owen_rev(__builtin_bitreverse32(sample), seed ^ 0xbff95bfeu);
        return sobol_dim(idx, 0, seed ^ 0x635c77bdu);
    }
    __device__ __forceinline__ V2 s2(uint32_t d) const {
        const uint32_t seed = hash ^ dim_hash(offset + d);
        const uint32_t idx = owen_rev(__builtin_bitreverse32(sample), seed ^ 0xf8ade99au);
        return {sobol_dim(idx, 0, seed ^ 0xe0aaaf76u), sobol_dim(idx, 1, seed ^ 0x94964d4eu)};
    }
    __device__ __forceinline__ V3 s3(uint32_t d) const {
        const uint32_t seed = hash ^ dim_hash(offset + d);
        const uint32_t idx = owen_rev(__builtin_bitreverse32(sample), seed ^ 0xcaa726acu);
        return {sobol_dim(idx, 0, seed ^ 0x9e78e391u), sobol_dim(idx, 1, seed ^ 0x67c33241u),
                sobol_dim(idx, 2, seed ^ 0x78c395c5u)};
    }
};

__device__ __forceinline__ Qrng qrng_at(uint32_t sample, uint32_t depth, uint32_t hash) {
    return Qrng{sample, QD_STRIDE * (depth + 1u), hash};
}

}  // namespace dev
}  // namespace rbh
