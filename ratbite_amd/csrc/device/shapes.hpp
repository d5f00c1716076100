// Shape lights: analytic disk and sphere emitters (DiskLightData /
// SphereLightData, src/core/light_tree_types.h:24-39, src/core/scene.h:319-492).
// The reference declares them but leaves their programs commented out
// (src/kernels/pathtrace.cu:150-228); its samplers for them are
// src/device/sampling.h:32-110. Here:
//
//   * a disk is a one-sided emitter along its frame's third column; it
//     occludes from both sides (the reference's intersection turns the
//     normal towards the ray);
//   * a sphere emits outwards; a point inside it receives nothing from it;
//   * a path that hits either ends there (the light is a pure emitter: its
//     material supplies the emission only).
//
// A scene holds few shape lights, so every ray tests all of them in a
// wave-uniform loop (the record loads are the same address on every lane)
// before the BVH walk, which then only looks for triangles nearer than the
// nearest shape.
#pragma once
#include "geom.hpp"
#include "types.hpp"
#include "vec.hpp"

namespace rbh {
namespace dev {

constexpr uint32_t kShapeHit = 0x80000000u;   // hit-record leaf: a shape light, its index in the low bits
constexpr uint32_t SHAPE_DISK = 0, SHAPE_SPHERE = 1;
constexpr float kShapePi = 3.14159265358979f;

__device__ __forceinline__ V3 shape_center(const ShapeLight& s) { return V3{s.pos[0], s.pos[1], s.pos[2]}; }

// glm::mat3_cast columns of the record's unit quaternion (x, y, z, w)
__device__ __forceinline__ Frame shape_frame(const ShapeLight& s) {
    const float x = s.q[0], y = s.q[1], z = s.q[2], w = s.q[3];
    return Frame{{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y)},
                 {2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x)},
                 {2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)}};
}
__device__ __forceinline__ V3 shape_normal(const ShapeLight& s) {
    const float x = s.q[0], y = s.q[1], z = s.q[2], w = s.q[3];
    return V3{2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)};
}

// Ray parameter of the disk hit in (0, tmax), else tmax: the plane through
// the centre, inside the radius (pathtrace.cu:150-190).
__device__ __forceinline__ float disk_t(const ShapeLight& s, V3 o, V3 d, float tmax) {
    const V3 n = shape_normal(s);
    const V3 oc = o - shape_center(s);
    const float t = -dot(n, oc) / dot(d, n);
    const V3 p = oc + d * t;
    return (t > 0.0f && t < tmax && dot(p, p) < s.radius * s.radius) ? t : tmax;
}

// Nearest sphere hit in (0, tmax), else tmax: the far root from inside. The
// discriminant is formed from the ray's closest approach to the centre and
// the roots from the non-cancelling form (Haines et al., "Precision
// Improvements for Ray/Sphere Intersection", Ray Tracing Gems, 2019).
__device__ __forceinline__ float sphere_t(const ShapeLight& s, V3 o, V3 d, float tmax) {
    const V3 oc = o - shape_center(s);
    const float b = dot(oc, d);
    const V3 qc = oc - d * b;
    const float r2 = s.radius * s.radius;
    const float disc = r2 - dot(qc, qc);
    if (!(disc >= 0.0f)) return tmax;
    const float q = -b - copysignf(sqrtf(disc), b);
    const float c = dot(oc, oc) - r2;
    float t0 = c / q, t1 = q;
    if (t0 > t1) {
        const float x = t0;
        t0 = t1;
        t1 = x;
    }
    const float t = t0 > 0.0f ? t0 : t1;
    return (t > 0.0f && t < tmax) ? t : tmax;
}

__device__ __forceinline__ float shape_t(const ShapeLight& s, V3 o, V3 d, float tmax) {
    return s.type == SHAPE_DISK ? disk_t(s, o, d, tmax) : sphere_t(s, o, d, tmax);
}

// Nearest shape hit below tmax: t and leaf = kShapeHit | index, or leaf ~0.
__device__ __forceinline__ void shape_closest(const ShapeLight* __restrict__ sh, uint32_t n, V3 o, V3 d, float tmax,
                                              float& t, uint32_t& leaf) {
    t = tmax;
    leaf = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i) {
        const float ti = shape_t(sh[i], o, d, t);
        if (ti < t) {
            t = ti;
            leaf = kShapeHit | i;
        }
    }
}
__device__ __forceinline__ bool shape_any(const ShapeLight* __restrict__ sh, uint32_t n, V3 o, V3 d, float tmax) {
    for (uint32_t i = 0; i < n; ++i)
        if (shape_t(sh[i], o, d, tmax) < tmax) return true;
    return false;
}

// Concentric square-to-disk map (sampling.h:34-51; Shirley and Chiu 1997),
// with the centre mapped to the centre instead of 0/0.
__device__ __forceinline__ V2 disk_concentric(V2 u) {
    const float a = u.x * 2.0f - 1.0f, b = u.y * 2.0f - 1.0f;
    if (a == 0.0f && b == 0.0f) return V2{0.0f, 0.0f};
    float r, th;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        th = (kShapePi / 4.0f) * (b / a);
    } else {
        r = b;
        th = (kShapePi / 2.0f) - (kShapePi / 4.0f) * (a / b);
    }
    return V2{r * cosf(th), r * sinf(th)};
}

// Uniform point on the disk (sampling.h:52-57, scaled by the radius and
// placed at the centre).
__device__ __forceinline__ V3 disk_sample(const ShapeLight& s, V2 u) {
    const Frame f = shape_frame(s);
    const V2 xy = disk_concentric(u);
    return shape_center(s) + (f.t * xy.x + f.b * xy.y) * s.radius;
}

// 1 - cos of the half-angle of the cone a sphere of radius r subtends from
// distance dist > r, with the small-angle form below sin^2 = 0.00068523
// (1.5 degrees) where 1 - sqrt(1 - x) cancels (sampling.h:73-92).
__device__ __forceinline__ float sphere_cone_omc(float dist, float r, float& sinMax, float& sin2Max, float& cosMax) {
    sinMax = r / dist;
    sin2Max = sinMax * sinMax;
    cosMax = fmaxf(0.0f, sqrtf(fmaxf(0.0f, 1.0f - sin2Max)));
    return sin2Max < 0.00068523f ? sin2Max / 2.0f : 1.0f - cosMax;
}
// Solid-angle density of the cone sampler as seen from o (outside the sphere).
__device__ __forceinline__ float sphere_cone_pdf(const ShapeLight& s, V3 o) {
    float sm, s2, cm;
    const float omc = sphere_cone_omc(len(shape_center(s) - o), s.radius, sm, s2, cm);
    return 1.0f / (2.0f * kShapePi * omc);
}

// Uniform direction in the cone the sphere subtends from o, returned as the
// point it meets on the sphere (sampling.h:71-110; Pharr, Jakob and
// Humphreys, PBRT 4th ed., 6.2.4). inside: o is in the sphere (no sample).
__device__ __forceinline__ V3 sphere_sample(const ShapeLight& s, V3 o, V2 u, float& pdf, bool& inside) {
    const V3 c = shape_center(s);
    V3 w = o - c;
    const float dist = len(w);
    inside = !(dist > s.radius);
    if (inside) {
        pdf = 0.0f;
        return c;
    }
    w = w / dist;
    float sinMax, sin2Max, cosMax;
    const float omc = sphere_cone_omc(dist, s.radius, sinMax, sin2Max, cosMax);
    float cosT = (cosMax - 1.0f) * u.x + 1.0f;
    float sin2T = 1.0f - cosT * cosT;
    if (sin2Max < 0.00068523f) {
        sin2T = sin2Max * u.x;
        cosT = sqrtf(1.0f - sin2T);
    }
    const float cosA = sin2T / sinMax + cosT * sqrtf(fmaxf(0.0f, 1.0f - sin2T / (sinMax * sinMax)));
    const float sinA = sqrtf(fmaxf(0.0f, 1.0f - cosA * cosA));
    const float phi = u.y * 2.0f * kShapePi;
    const Frame f = frame_from_normal(w);
    const V3 dir = w * cosA + f.t * (sinA * cosf(phi)) + f.b * (sinA * sinf(phi));
    pdf = 1.0f / (2.0f * kShapePi * omc);
    return c + dir * s.radius;
}

}  // namespace dev
}  // namespace rbh
