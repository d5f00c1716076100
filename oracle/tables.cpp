// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// Sobol direction numbers for dimensions 0..3 in the bit-reversed layout that
// src/device/quasi_random.h:137-179 indexes. Built from the Joe-Kuo
// primitive-polynomial recurrence rather than copied; the tests check the
// result against torch.quasirandom.SobolEngine(scramble=False) and against
// the stratification property of (0,2)-sequences.
#include "rb_math.hpp"

namespace orc {

namespace {
SobolTable build_table() {
    SobolTable T{};
    struct DimSpec { int s; uint32_t a; uint32_t m[3]; };
    const DimSpec spec[4] = {{0, 0, {0, 0, 0}}, {1, 0, {1, 0, 0}}, {2, 1, {1, 3, 0}}, {3, 1, {1, 3, 1}}};
    for (int d = 0; d < 4; ++d) {
        uint32_t v[33] = {0};  // v[k], k = 1..32, MSB-aligned direction numbers
        if (d == 0) {
            for (int k = 1; k <= 32; ++k) v[k] = 1u << (32 - k);
        } else {
            const int s = spec[d].s;
            for (int k = 1; k <= s; ++k) v[k] = spec[d].m[k - 1] << (32 - k);
            for (int k = s + 1; k <= 32; ++k) {
                uint32_t x = v[k - s] ^ (v[k - s] >> s);
                for (int j = 1; j < s; ++j)
                    if ((spec[d].a >> (s - 1 - j)) & 1u) x ^= v[k - j];
                v[k] = x;
            }
        }
        for (int i = 0; i < 32; ++i) T.t[d][i] = brev(v[i + 1]);
    }
    return T;
}
}  // namespace

const SobolTable& sobol_table() {
    static const SobolTable T = build_table();
    return T;
}

}  // namespace orc
