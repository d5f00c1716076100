// Device-side data layouts shared by the kernels and the host context.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rbh {

constexpr int kLambdaMin = 360;       // src/core/spectral_settings.h:6
constexpr int kLambdaMax = 830;       // src/core/spectral_settings.h:7
constexpr int kDense = kLambdaMax - kLambdaMin + 1;
constexpr int kNL = 4;                // hero wavelengths per path, spectral_settings.h:5

// LUT block offsets (floats) inside the context's single LUT buffer.
constexpr size_t kLutCond = 0;
constexpr size_t kLutDielOut = kLutCond + 32 * 32;
constexpr size_t kLutDielIn = kLutDielOut + 32 * 32 * 32;
constexpr size_t kLutReflOut = kLutDielIn + 32 * 32 * 32;
constexpr size_t kLutReflIn = kLutReflOut + 32 * 32 * 32;
constexpr size_t kLutSheen = kLutReflIn + 32 * 32 * 32;
constexpr size_t kLutTotal = kLutSheen + 32 * 32 * 4;

// Binary BVH node, 64 bytes: both children's boxes in one node so a
// traversal step is one 64-byte fetch (two dwordx4 pairs).
// child >= 0: inner node index; child < 0: ~(first triangle) with the
// triangle count in the low bits of cnt.
struct alignas(16) BvhNode {
    float lminx, lmaxx, lminy, lmaxy;   // left child box (x, y)
    float rminx, rmaxx, rminy, rmaxy;   // right child box (x, y)
    float lminz, lmaxz, rminz, rmaxz;   // z extents
    int32_t left, right;                // child references
    uint32_t lcnt, rcnt;                // triangle counts of leaf children (0 = inner)
};
static_assert(sizeof(BvhNode) == 64, "BvhNode is 64 bytes");

// Four-wide BVH node, 128 bytes (one 128-byte cache line): the boxes of up to
// four children as structure-of-arrays rows, then the child references
// (>= 0 inner node, < 0 ~leaf index into the triangle records, kBvh4Empty
// for an unused slot, whose box is empty: lo = +inf, hi = -inf).
constexpr int32_t kBvh4Empty = (int32_t)0x80000000;
struct alignas(16) Bvh4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node is 128 bytes");

// The same node quantised to 64 bytes: child boxes as 8-bit offsets from the
// node's box corner o in power-of-two steps s per axis, decoded as
// o + q * s (q * s is exact; one rounding in the add, the same in the builder
// and in traversal, so the builder can make every decoded box contain its
// child: lower bounds rounded down, upper bounds up).
struct alignas(64) Bvh4QNode {
    float ox, oy, oz;
    uint32_t exps;                      // biased exponents of s: x bits 0-7, y 8-15, z 16-23
    int32_t child[4];                   // as in Bvh4Node
    uint32_t lox, loy, loz, hix;        // byte k = child k
    uint32_t hiy, hiz;
    uint32_t pad[2];
};
static_assert(sizeof(Bvh4QNode) == 64, "Bvh4QNode is 64 bytes");
__host__ __device__ inline float qnode_step(uint32_t expByte) {
    union { uint32_t u; float f; } v;
    v.u = expByte << 23;
    return v.f;
}
__host__ __device__ inline float qnode_decode(float o, uint32_t q, float s) { return o + (float)q * s; }

// Triangle in leaf order, one 64-byte line: the vertices the traversal
// tests (first 48 bytes), and what shading needs besides them: the original
// triangle index, its material and light index (0xFFFFFFFF: not emissive)
// and its vertex indices (for the per-vertex attributes).
struct alignas(16) TriRecord {
    float v0x, v0y, v0z; uint32_t prim;
    float v1x, v1y, v1z; uint32_t material;
    float v2x, v2y, v2z; uint32_t light;
    uint32_t i0, i1, i2;
    uint32_t shadeClass;   // the material's shading class (SC_*), for the material-sorted hit queue
};
static_assert(sizeof(TriRecord) == 64, "TriRecord is 64 bytes");

// LightTree::PackedNode (src/core/light_tree_types.h:119-162)
struct PackedNode {
    float mean[3];
    float variance;
    uint32_t dir;
    float sharpness;
    float flux;
    uint32_t core;
};
static_assert(sizeof(PackedNode) == 32, "PackedNode is 32 bytes");

// EmissiveTriangleLightData (src/core/light_tree_types.h:40-46)
struct EmissiveTri {
    float v[9];
    float uv[6];
    uint32_t material;
    uint32_t prim;
};
static_assert(sizeof(EmissiveTri) == 68, "EmissiveTri is 68 bytes");

// DiskLightData / SphereLightData (src/core/light_tree_types.h:24-39) in one
// record: type 0 disk, 1 sphere; frame is a unit quaternion (x, y, z, w)
// whose third column is the disk's emitting normal.
struct ShapeLight {
    float pos[3];
    float power;
    float q[4];
    float radius;
    uint32_t material;
    uint32_t type;
    uint32_t pad;
};
static_assert(sizeof(ShapeLight) == 48, "ShapeLight is 48 bytes");

struct Material {
    uint32_t bxdf, flags, attributes, attributeStride;
    uint32_t normalOffset, frameOffset, texCoord1Offset, texCoord2Offset;
    uint32_t iorSpectrum, acSpectrum, emissionSpectrum;
    float mfRoughness, ior, alphaCutoff;
    uint32_t textures, texCoordSets;
    int32_t tex[7];
    uint32_t factors;
    float baseColor[4];
    float metalness, roughness, transmission;
    float sheenColor[3];
    float sheenRoughness;
    float emissive[3];
};
static_assert(sizeof(Material) == 152, "Material is 152 bytes");

// Shading classes: hits are queued by the BxDF code path their material
// takes in the shade kernel, so that a wave of the shade kernel mostly runs
// one of them.
enum : uint32_t {
    SC_SURFACE = 0,        // ComplexSurface without transmission or sheen (diffuse / glossy / metallic)
    SC_TRANSMISSIVE = 1,   // ComplexSurface with a transmission factor or texture
    SC_SHEEN = 2,          // ComplexSurface with a sheen layer
    SC_CONDUCTOR = 3,      // PureConductor
    SC_DIELECTRIC = 4,     // PureDielectric
    SC_SHAPE = 5,          // a disk or sphere light (the path ends there)
    kShadeClasses = 6
};
__host__ __device__ inline uint32_t shade_class(const Material& m) {
    if (m.bxdf == 0u) return SC_CONDUCTOR;
    if (m.bxdf == 1u) return SC_DIELECTRIC;
    if (m.flags & 2u) return SC_SHEEN;
    if ((m.textures & 8u) || ((m.factors & 8u) && m.transmission != 0.0f)) return SC_TRANSMISSIVE;
    return SC_SURFACE;
}

struct TextureDesc {
    uint32_t width, height, depth, format, srgb;
    uint32_t addr[3];
    uint32_t filter, pad;
    uint64_t offset;
};
static_assert(sizeof(TextureDesc) == 48, "TextureDesc is 48 bytes");

}  // namespace rbh
