// Launch-time view of the scene and of the wavefront path state, passed by
// value to every integrator kernel.
#pragma once
#include "device/types.hpp"
#include "device/vec.hpp"

namespace rbh {
namespace dev {
struct TreeNodeX;
}

constexpr int kTraceBlock = 128;   // lanes per block of the trace and shadow kernels
constexpr int kShadeBlock = 128;

enum : uint32_t {
    CNT_QUEUE0 = 0, CNT_QUEUE1 = 1, CNT_SHADOW = 2,
    CNT_HIT = 3,                               // hits of shading class 0 (all hits when unsorted)
    CNT_MISS = 4,
    CNT_TRACE_HEAD = 5, CNT_SHADOW_HEAD = 6,   // work-fetch positions of the trace and shadow queues
    CNT_HIT_CLASS1 = 7,                        // hits of shading classes 1 .. kShadeClasses-1
    CNT_COUNT = CNT_HIT_CLASS1 + kShadeClasses - 1
};
// counter of the hits of shading class c
__host__ __device__ inline uint32_t hit_counter(uint32_t c) { return c == 0u ? (uint32_t)CNT_HIT : CNT_HIT_CLASS1 + c - 1u; }
enum : uint32_t { STAT_CLOSEST = 0, STAT_SHADOW = 1, STAT_VERTICES = 2, STAT_NONFINITE = 3, STAT_COUNT = 4 };

// Structure-of-arrays path state, one entry per path slot of the pass.
struct PathBuffers {
    float4* od;     // ray origin, bxdfPDF
    float4* dr;     // ray direction, refractionScale
    float4* L;      // radiance estimate
    float4* T;      // throughput
    float4* nf;     // last shading normal, state flags
    uint32_t* depth;
    float4* hit;    // t, b1, b2, primitive
    float4* Lnee;   // deferred NEE term
    float4* sro;    // shadow ray origin, tmax
    float4* srd;    // shadow ray direction
    uint32_t* q0;   // paths to trace (even / odd bounces)
    uint32_t* q1;
    uint32_t* hq;   // paths whose closest hit found a surface or shape light (to shade): kShadeClasses
                    // segments of NP entries, segment c holding the hits of shading class c
                    // (class 0 only when the hits are not sorted by material)
    uint32_t* mq;   // paths whose ray left the scene
    uint32_t* sq;   // paths with a deferred NEE shadow ray
    uint32_t* prec; // guided renders that collect training data: last record of the path
};

// Everything one scene contributes to a pass: S consecutive samples of its
// P = W*H pixels occupy path slots [slot0, slot0 + S*P), slot slot0 + k*P + i
// holding sample (sample + k) of pixel i.
struct SceneView {
    uint32_t P, W, H, sample, S, slot0;
    uint32_t maxDepth, maxRefl, maxTrans;
    uint32_t first;          // the pass starts a fresh accumulation
    float invW, invH, psW, psH;
    dev::V3 camU, camV, camW;
    uint32_t dof;
    float aperture, focus;
    // geometry
    const BvhNode* nodes;
    const Bvh4Node* nodes4;
    const Bvh4QNode* nodes4q;
    const TriRecord* tris;
    uint32_t nnodes, nnodes4;
    const float* verts;
    const uint32_t* indices;
    const uint8_t* attribs;
    const uint32_t* triMaterial;
    const uint32_t* triLight;
    // materials and tables
    const Material* materials;
    const TextureDesc* texDescs;
    const uint8_t* texels;
    const float* srgbLut;    // 256 sRGB-decoded 8-bit values
    const float* spectra;
    const float* colors;
    const float* luts;
    // lights
    const dev::TreeNodeX* treeNodes;   // expanded light-tree nodes (48 B)
    const uint32_t* treePointers;
    const uint64_t* treeMasks;
    const EmissiveTri* lights;
    const ShapeLight* shapes;          // disk and sphere lights
    uint32_t nshapes;
    uint32_t shapeBase;                // light index of shapes[0] (the triangle-light count)
    uint32_t envOn;
    uint32_t envW, envH, envC;
    const float* envTex;
    const uint16_t* envCond;
    const uint16_t* envMarg;
    float envIntegral;
    double4* accum;
};

// Persistent-grid sizes of the trace/shadow and shade/miss kernels (blocks)
struct LaunchShape {
    int traceGrid, shadeGrid;
    int shadeWaves;          // register budget of the shade kernel: 1 (free), 3 or 4 waves per SIMD
    int traceFetch;          // trace/shadow lanes fetch a new ray when theirs completes (1) or wave by wave (0)
    int traceWaves;          // register budget of the wave-synchronous trace/shadow kernels: 1 (free), 7 or 8
    int quantNodes;          // wave-synchronous trace/shadow kernels read the 64-byte quantised nodes
    int sortHits;            // queue the hits by shading class (material-sorted shading)
};

struct Frame1 {
    SceneView sc;                 // single-scene passes
    const SceneView* scenes;      // batched passes: views in slot order
    uint32_t nscenes;
    uint32_t NP;                  // path slots of the pass (all scenes)
    uint32_t maxP;                // largest pixel count of a scene in the pass
    uint32_t bounce;
    uint32_t camOrigin0;          // no scene of the pass uses depth of field: camera rays start at 0
    uint32_t sortHits;            // LaunchShape::sortHits
    // state
    PathBuffers p;
    int* spill;              // traversal-stack spill columns of the trace/shadow grid
    uint32_t spillStride;    // lanes of the trace/shadow grid
    uint32_t* counters;
    unsigned long long* stats;
    // path guiding (guided renders only)
    uint32_t guided;          // one-sample MIS with the network at ComplexSurface vertices
    uint32_t noDrop;          // an invalid BxDF sample ends the path instead of discarding the sample
    const __bf16* gwf;        // network weights (BF16, forward layout) and FP32 biases
    const float* gbias;
    const uint8_t* gwq;       // FP8 E4M3 weights and per-layer scales (nullptr: BF16 inference)
    const float* gws;
    float* gfeat;             // per hit-queue entry: 16-float input record
    float* gmix;              // per hit-queue entry: 41-float mixture (world space) + selection
    float gmin[3], ginv[3];   // position normalisation
    uint32_t collect;         // record training data
    uint32_t recThreshold;    // keep a vertex when its hash is below this
    float* grec;              // pending records: 24 floats each
    float* gtrain;            // finished training records: 16 floats each
    uint32_t* grecCount;
    uint32_t grecCap;
};

// pending-record layout (floats): 0..10 input features, 11..13 sampled
// direction, 14 BSDF-sampler density of the direction, 15 previous record of
// the path (uint bits), 16..19 throughput before the vertex, 20..23 radiance
// before the next vertex
constexpr int kPendingRec = 24;
constexpr uint32_t kRecPending = 0x80000000u;   // radiance snapshot not taken yet
constexpr uint32_t kNoRec = 0xFFFFFFFFu;

}  // namespace rbh
