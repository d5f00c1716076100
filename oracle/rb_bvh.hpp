// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header). CPU BVH used by
// the oracle's path tracer; see bvh.cpp.
#pragma once
#include "rb_math.hpp"
#include <vector>

namespace orc {

bool ray_triangle(const V3& o, const V3& d, const V3& a, const V3& b, const V3& c, float tmin, float tmax, float& t,
                  float& b1, float& b2);

struct CpuBvh {
    struct Node { float mn[3], mx[3]; uint32_t first, count; };
    struct Hit { float t; uint32_t prim; float b1, b2; };
    std::vector<Node> nodes;
    std::vector<uint32_t> prims;
    const float* verts = nullptr;
    const uint32_t* indices = nullptr;
    uint32_t ntris = 0;
    void build(const float* v, const uint32_t* idx, uint32_t nt);
    bool trace(const V3& o, const V3& d, float tmax, bool anyhit, Hit& h) const;
};

}  // namespace orc
