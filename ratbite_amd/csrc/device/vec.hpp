// Small vector algebra for the device code (float3-like value types with
// glm semantics where the reference relies on them).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rbh {
namespace dev {

struct V2 { float x, y; };
struct V3 { float x, y, z; };

__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
// glm::normalize: v * inversesqrt(dot(v, v))
__device__ __forceinline__ V3 nrm(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// Frame with columns t, b, n: world -> local is the transpose.
struct Frame {
    V3 t, b, n;
    __device__ __forceinline__ V3 to_local(V3 v) const { return {dot(t, v), dot(b, v), dot(n, v)}; }
    __device__ __forceinline__ V3 to_world(V3 v) const { return t * v.x + b * v.y + n * v.z; }
};

// src/device/local_transform.h:283-293 (Duff et al. branchless basis)
__device__ __forceinline__ Frame frame_from_normal(V3 n) {
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float b = n.x * n.y * a;
    return Frame{{1.0f + s * (n.x * n.x) * a, s * b, -s * n.x}, {b, s + (n.y * n.y) * a, -n.y}, n};
}

__device__ __forceinline__ V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

}  // namespace dev
}  // namespace rbh
