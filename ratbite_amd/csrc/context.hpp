// Host-side context of the HIP path tracer: device buffers, stream, events
// and error state behind the rbh_* C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>
#include "../../include/ratbite_hip.h"
#include "device/types.hpp"

namespace rbh {

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grow-only: keeps the allocation when it is large enough.
    hipError_t reserve(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = n;
        return hipSuccess;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct EnvMap {
    bool enabled = false;
    int width = 0, height = 0, channels = 0;
    double integral = 0.0;       // host copy, already divided by w*h (a float value in FP32 mode)
    DevBuf texels;               // w*h*channels floats
    DevBuf cond, marg;           // uint16 inverted-CDF indices
    DevBuf cdfY;                 // h floats (normalised marginal CDF)
    DevBuf integralDev;          // raw total (float) + double slot
};

struct LightTreeDev {
    uint32_t nnodes = 0, nlights = 0;   // nlights: triangle lights (light indices [0, nlights))
    uint32_t nshapes = 0;               // shape lights in the tree (light indices nlights + i)
    bool stale = false;                 // the shape lights changed since the tree was built
    DevBuf nodes, lightPointers, bitmasks, tris;
    DevBuf nodesX;               // expanded nodes the kernels traverse (decoded direction, SG integral)
    std::vector<rbh_packed_node> hostNodes;
    std::vector<uint32_t> hostPointers;   // nlights + nshapes entries
    std::vector<uint64_t> hostMasks;
};

// Disk and sphere lights (DiskLightData / SphereLightData) and their flux
struct ShapeLights {
    uint32_t n = 0;
    DevBuf recs;
    std::vector<rbh_shape_light> host;
    std::vector<float> flux;
};

struct Geometry {
    uint32_t nverts = 0, ntris = 0;
    int64_t maxLight = -1;   // largest triLight index (-1: no emissive triangles)
    DevBuf verts, indices, attribs, triMaterial, triLight;
};

struct Accel {
    bool built = false;
    uint32_t nnodes = 0, depth = 0;
    DevBuf nodes;     // BvhNode (binary LBVH)
    DevBuf nodes4;    // Bvh4Node (the binary tree collapsed four-wide; traversed)
    DevBuf nodes4q;   // the four-wide nodes quantised to 64 bytes (Bvh4QNode)
    uint32_t nnodes4 = 0;
    DevBuf tris;      // TriRecord in leaf order
    DevBuf scratch;   // build scratch
    size_t keysOffset = 0;   // byte offset of the sorted 64-bit keys in scratch (kept after the build)
    int keyIndexBits = 32;   // low key bits holding the triangle index
    uint32_t ntris = 0;      // triangles the structure was built over
    bool stale = false;      // vertices changed since the last build or refit
};

// HIP-event timing of individual render-loop kernels (RBH_RENDER_TIMING):
// one event pair per launch, recorded on the render stream around the launch.
struct KernelTimer {
    enum { TRACE = 0, SHADE = 1, SHADOW = 2, GUIDE = 3, NCLASS = 4 };
    std::vector<hipEvent_t> pool;   // 2 events per recorded launch
    std::vector<int> cls;           // kernel class of launch i
    KernelTimer() = default;
    KernelTimer(const KernelTimer&) = delete;
    KernelTimer& operator=(const KernelTimer&) = delete;
    ~KernelTimer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    void reset() { cls.clear(); }
    hipError_t begin(hipStream_t s, int k) {
        const size_t i = cls.size();
        while (pool.size() < 2 * i + 2) {
            hipEvent_t e = nullptr;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            pool.push_back(e);
        }
        cls.push_back(k);
        return hipEventRecord(pool[2 * i], s);
    }
    hipError_t end(hipStream_t s) { return hipEventRecord(pool[2 * cls.size() - 1], s); }
    // Sums the durations per class; the stream must have been synchronised.
    hipError_t collect(double ms[NCLASS], uint32_t n[NCLASS]) const {
        for (int k = 0; k < NCLASS; ++k) { ms[k] = 0.0; n[k] = 0; }
        for (size_t i = 0; i < cls.size(); ++i) {
            float t = 0.0f;
            hipError_t r = hipEventElapsedTime(&t, pool[2 * i], pool[2 * i + 1]);
            if (r != hipSuccess) return r;
            ms[cls[i]] += t;
            n[cls[i]] += 1;
        }
        return hipSuccess;
    }
};

struct RenderState {
    uint32_t width = 0, height = 0;
    DevBuf accum;     // W*H*4 doubles
    DevBuf paths;     // SoA path state
    DevBuf queues;    // ray/shadow queues + counters
    DevBuf image;     // RGBA8 resolve output
    DevBuf spill;     // traversal-stack spill columns
    DevBuf gfeat;     // guided renders: network input per hit-queue entry
    DevBuf gmix;      // guided renders: mixture per hit-queue entry
    size_t pathCapacity = 0;
    // path slots one pass may hold (samples per pass = slotBudget / pixels);
    // about 200 B of HBM per slot; 0 = as many as a quarter of the free
    // device memory holds (fewer passes pay the deep-bounce tail fewer times)
    size_t slotBudget = 0;
    // launch shape: the fastest of a sweep of full 1080p 64-spp frames of the
    // 4.1M-triangle spheres scene (persistent grids capped at one lane per
    // path slot); rbh_autotune / rbh_set_tuning replace it per context
    int traceBlocksPerCU = 256;
    int shadeBlocksPerCU = 256;
    int shadeWaves = 4;          // shade-kernel variant (1, 3, 4 waves per SIMD)
    int traceFetch = 0;          // 1: trace/shadow lanes fetch a new ray as soon as theirs completes
    int traceWaves = 7;          // trace/shadow register budget (1, 7, 8 waves per SIMD)
    int quantNodes = 0;          // traverse the 64-byte quantised four-wide nodes
    int sortHits = 1;            // shade conductor hits as their own run of the hit queue
    DevBuf sceneTable;           // batched passes: SceneView per scene and pass
    bool useGraph = false;       // replay each pass's launch sequence as a captured HIP graph
    hipGraphExec_t graphExec = nullptr;
    KernelTimer timer;
    ~RenderState() {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
    }
};

}  // namespace rbh

struct rbh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    rbh::DevBuf spectra;  uint32_t nspectra = 0;
    rbh::DevBuf colorTables;  // 6 x 471: X, Y, Z, basis R, G, B
    bool haveColor = false;
    rbh::DevBuf luts;         // conductor | dielOut | dielIn | reflOut | reflIn | sheen
    bool haveLuts = false;
    rbh::DevBuf textures, texels; uint32_t ntextures = 0;
    rbh::DevBuf srgbLut;         // sRGB decode of the 256 8-bit values, filled with the textures
    rbh::DevBuf materials; uint32_t nmaterials = 0;
    rbh::Geometry geom;
    rbh::Accel accel;
    rbh::LightTreeDev tree;
    rbh::ShapeLights shapes;
    rbh::EnvMap env;
    rbh::RenderState rs;
    rbh_stats stats{};
    void* guide = nullptr;    // guiding state (guide.hip)
};

namespace rbh {

inline int fail(rbh_ctx c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

}  // namespace rbh

#define RBH_HIP(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return rbh::fail((ctx), RBH_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define RBH_ALLOC(ctx, buf, n)                                                                          \
    do {                                                                                                \
        hipError_t e_ = (buf).reserve(n);                                                               \
        if (e_ != hipSuccess) return rbh::fail((ctx), RBH_ERR_NOMEM, "allocation of %zu bytes failed: %s", (size_t)(n), hipGetErrorString(e_)); \
    } while (0)
