// TEST INFRASTRUCTURE — a C entry point over the reference's own software
// rasteriser (src/core/sw_rast.h, compiled from where it lies by
// oracle/ref/Makefile into oracle/_ref/): the texel set it visits for one
// triangle, the checker of the product's restatement (host/emissive.cpp).
#include <functional>

#include "sw_rast.h"

extern "C" int ref_rasterise(const float* a, const float* b, const float* c, int* xy, int capacity) {
    int n = 0;
    const std::function<void(int, int)> visit = [&](int x, int y) {
        if (n < capacity) {
            xy[2 * n] = x;
            xy[2 * n + 1] = y;
        }
        ++n;
    };
    SoftwareRasterization::drawTriangle2D(a, b, c, visit);
    return n;
}
