// Launch-time view of the scene and of the wavefront path state, passed by
// value to every integrator kernel.
#pragma once
#include "device/types.hpp"
#include "device/vec.hpp"

namespace rbh {

constexpr int kTraceBlock = 128;   // lanes per block of the trace and shadow kernels
constexpr int kShadeBlock = 128;

enum : uint32_t { CNT_QUEUE0 = 0, CNT_QUEUE1 = 1, CNT_SHADOW = 2, CNT_COUNT = 4 };
enum : uint32_t { STAT_CLOSEST = 0, STAT_SHADOW = 1, STAT_VERTICES = 2, STAT_COUNT = 4 };

// Structure-of-arrays path state, one entry per path slot of the pass.
struct PathBuffers {
    float4* od;     // ray origin, bxdfPDF
    float4* dr;     // ray direction, refractionScale
    float4* lam;    // hero wavelengths
    float4* lpdf;   // wavelength pdfs
    float4* L;      // radiance estimate
    float4* T;      // throughput
    float4* nf;     // last shading normal, state flags
    uint32_t* depth;
    float4* hit;    // t, b1, b2, primitive
    float4* Lnee;   // deferred NEE term
    float4* sro;    // shadow ray origin, tmax
    float4* srd;    // shadow ray direction
    uint32_t* q0;
    uint32_t* q1;
    uint32_t* sq;
};

struct Frame1 {
    // pass: S consecutive samples of all P = W*H pixels, NP = S*P path slots;
    // slot i holds sample (sample + i / P) of pixel i % P
    uint32_t P, W, H, sample, S, NP;
    uint32_t maxDepth, maxRefl, maxTrans;
    uint32_t bounce;
    float invW, invH, psW, psH;
    dev::V3 camU, camV, camW;
    uint32_t dof;
    float aperture, focus;
    // geometry
    const BvhNode* nodes;
    const TriRecord* tris;
    uint32_t nnodes;
    const float* verts;
    const uint32_t* indices;
    const uint8_t* attribs;
    const uint32_t* triMaterial;
    const uint32_t* triLight;
    // materials and tables
    const Material* materials;
    const TextureDesc* texDescs;
    const uint8_t* texels;
    const float* spectra;
    const float* colors;
    const float* luts;
    // lights
    const PackedNode* treeNodes;
    const uint32_t* treePointers;
    const uint64_t* treeMasks;
    const EmissiveTri* lights;
    uint32_t envOn;
    uint32_t envW, envH, envC;
    const float* envTex;
    const uint16_t* envCond;
    const uint16_t* envMarg;
    float envIntegral;
    // state
    PathBuffers p;
    int* spill;              // traversal-stack spill columns of the trace/shadow grid
    uint32_t spillStride;    // lanes of the trace/shadow grid
    uint32_t* counters;
    unsigned long long* stats;
    double4* accum;
};

}  // namespace rbh
