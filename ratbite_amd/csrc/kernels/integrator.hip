// Wavefront spectral path tracer replacing the OptiX megakernel
// __raygen__main (src/kernels/pathtrace.cu:1114-1813). The per-path bodies
// live in device/wavefront.hpp; the kernels here run them over persistent
// grids and compact the queues.
//
// One pass renders S consecutive samples of every pixel at once (S*W*H path
// slots, slot = sample-major pixel index) so that the long tail of a few deep
// paths is paid once per pass rather than once per sample; path state lives in
// structure-of-arrays buffers in HBM. Per bounce:
//   k_trace    closest hit for the queued paths           (pathtrace.cu:1162-1177),
//              sorting them into a hit queue (by shading class) and a miss
//              queue; at the first bounce it generates the camera rays
//              (pathtrace.cu:1114-1160)
//   k_miss     sky emission + MIS for the misses; the path ends
//              (pathtrace.cu:1185-1274, sky branch)
//   k_shade    emission + MIS, surface attributes, light-tree NEE, BxDF,
//              Russian roulette, next origin, depth caps   (pathtrace.cu:1179-1785)
//   k_shadow   any-hit test of the deferred NEE rays; adds the NEE term of
//              the bounce when the light is visible        (pathtrace.cu:1704-1720)
// and after the last bounce k_accumulate adds each sample's XYZ to the FP64
// accumulation buffer (pathtrace.cu:1789-1812). Queues are compacted with
// wave ballots; the trace and shade kernels stage finished paths in per-wave
// LDS lists and write them 64 or more at a time (one atomic per run). Queue lengths stay on the device; every
// kernel reads its count and exits early, so the host issues a fixed launch
// sequence per pass.
#include "../context.hpp"
#include "../device/mlp_wave.hpp"
#include "../device/traverse.hpp"
#include "../device/wavefront.hpp"
#include "../integrator.hpp"

namespace rbh {

using namespace dev;
using namespace wf;

namespace {

constexpr int kStackLds = 16;   // traversal stack entries kept in LDS per lane

__device__ __forceinline__ uint32_t wave_push(uint32_t* __restrict__ counter, bool pred) {
    // one atomic per wave: ballot, leader adds the popcount, lanes take
    // consecutive slots by their rank among the active lanes
    const unsigned long long mask = __ballot(pred);
    const uint32_t n = (uint32_t)__popcll(mask);
    if (n == 0) return 0xFFFFFFFFu;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, n);
    base = __shfl(base, leader);
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + rank : 0xFFFFFFFFu;
}

__device__ __forceinline__ void wave_count(unsigned long long* __restrict__ counter, uint32_t v) {
    // sum v over the wave with one atomic
    uint32_t s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0 && s) atomicAdd(counter, (unsigned long long)s);
}

// The queue counters at the start of a pass: every slot traces its camera
// ray (the first closest-hit pass generates it).
__global__ void k_pass_start(Frame1 f) {
    if (threadIdx.x == 0) {
        f.counters[CNT_QUEUE0] = f.NP;
        f.counters[CNT_QUEUE1] = 0u;
        f.counters[CNT_SHADOW] = 0u;
        f.counters[CNT_HIT] = 0u;
        f.counters[CNT_MISS] = 0u;
        f.counters[CNT_TRACE_HEAD] = 0u;
        f.counters[CNT_SHADOW_HEAD] = 0u;
        for (uint32_t c = 1; c < kShadeClasses; ++c) f.counters[hit_counter(c)] = 0u;
    }
}

// The ray a closest-hit pass traces for path slot pid: at the first bounce
// the camera ray, generated here and stored for the shading kernels (with a
// fresh training-record chain when collecting), afterwards the stored ray.
template <bool BATCH>
__device__ __forceinline__ void pass_ray(const Frame1& f, const SceneView& S, uint32_t pid, bool first, V3& o, V3& d) {
    if (first) {
        camera_ray(S, pid, o, d);
        f.p.od[pid] = make_float4(o.x, o.y, o.z, 1.0f);   // bxdfPdf = 1
        f.p.dr[pid] = make_float4(d.x, d.y, d.z, 1.0f);   // refractionScale = 1
        if (f.collect) f.p.prec[pid] = kNoRec;
    } else {
        const float4 od = f.p.od[pid], dr = f.p.dr[pid];
        o = V3{od.x, od.y, od.z};
        d = V3{dr.x, dr.y, dr.z};
    }
}

__global__ void k_reset(uint32_t* __restrict__ counters, uint32_t which) {
    if (threadIdx.x == 0) counters[which] = 0u;
    if (threadIdx.x == 1) counters[CNT_SHADOW] = 0u;
    if (threadIdx.x == 2) counters[CNT_SHADOW_HEAD] = 0u;
}

__global__ void k_reset_hits(uint32_t* __restrict__ counters) {
    if (threadIdx.x < kShadeClasses) counters[hit_counter(threadIdx.x)] = 0u;
    if (threadIdx.x == kShadeClasses) counters[CNT_MISS] = 0u;
    if (threadIdx.x == kShadeClasses + 1) counters[CNT_TRACE_HEAD] = 0u;
}

// Hit-queue entries in shading order: the segments of the shading classes
// one after the other (class c's entries at hq[c * NP ...]).
struct HitQueue {
    uint32_t start[kShadeClasses + 1];   // first shading-order index of each class, then the total
    __device__ __forceinline__ explicit HitQueue(const Frame1& f) {
        start[0] = 0u;
#pragma unroll
        for (uint32_t c = 0; c < kShadeClasses; ++c) start[c + 1] = start[c] + f.counters[hit_counter(c)];
    }
    __device__ __forceinline__ uint32_t count() const { return start[kShadeClasses]; }
    __device__ __forceinline__ uint32_t entry(const Frame1& f, uint32_t i) const {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t k = 1; k < kShadeClasses; ++k) c += i >= start[k] ? 1u : 0u;
        return f.p.hq[(size_t)c * f.NP + (i - start[c])];
    }
};

// Work fetch of the persistent trace and shadow kernels. A wave claims
// chunks of kChunk consecutive queue entries with one atomic on *head and
// hands them to its idle lanes whenever at least kRefill lanes are idle (or
// all of them), so a lane takes a new ray as soon as its query completes
// while the wave's rays stay a contiguous, coherent stretch of the queue.
constexpr int kRefill = 16;
constexpr uint32_t kChunk = 256;
struct WaveChunk {
    uint32_t next = 0, end = 0;   // wave-uniform: unclaimed entries [next, end) of the wave's chunk
    bool drained = false;         // the queue has no further chunk
    // Returns this lane's new entry or ~0. Every lane of the wave calls it.
    __device__ __forceinline__ uint32_t fetch(uint32_t* __restrict__ head, uint32_t count, bool idle, bool anyBusy) {
        const unsigned long long mask = __ballot(idle);
        const uint32_t n = (uint32_t)__popcll(mask);
        if (n == 0 || (anyBusy && n < (uint32_t)kRefill)) return 0xFFFFFFFFu;
        if (next >= end && !drained) {
            const int lane = __lane_id();
            const int leader = __ffsll((long long)mask) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(head, kChunk);
            base = __shfl(base, leader);
            if (base >= count) {
                drained = true;
            } else {
                next = base;
                end = min(base + kChunk, count);
            }
        }
        if (next >= end) return 0xFFFFFFFFu;
        const uint32_t idx = next + (uint32_t)__popcll(mask & ((1ull << __lane_id()) - 1ull));
        next = min(next + n, end);
        return idle && idx < end ? idx : 0xFFFFFFFFu;
    }
    // no lane active, nothing left to hand out
    __device__ __forceinline__ bool done(bool anyActive) const { return !anyActive && drained && next >= end; }
};

// ------------------------------------------------------------- closest hit
__device__ __forceinline__ TravStack<kStackLds> lane_stack(const Frame1& f, int* lds) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    return TravStack<kStackLds>{lds + threadIdx.x, (int)blockDim.x, f.spill + t, f.spillStride};
}

// Wave-private staging of queue pushes: finished rays are appended to an
// LDS list of the wave and written out 64 or more at a time with one atomic,
// instead of one atomic per wave for every traversal step in which some lane
// finished.
struct WaveList {
    uint32_t* buf;   // 128 entries of this wave
    uint32_t n;      // entries held (wave-uniform)
    __device__ __forceinline__ void add(bool pred, uint32_t v) {
        const unsigned long long mask = __ballot(pred);
        const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << __lane_id()) - 1ull));
        if (pred) buf[n + rank] = v;
        n += (uint32_t)__popcll(mask);
    }
    // writes the held entries to queue[*counter ...] when at least `min` are held
    __device__ __forceinline__ void flush(uint32_t* __restrict__ counter, uint32_t* __restrict__ queue, uint32_t min) {
        if (n == 0 || n < min) return;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // the lanes' LDS writes are visible
        const int lane = __lane_id();
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(counter, n);
        base = __shfl(base, 0);
        for (uint32_t k = (uint32_t)lane; k < n; k += 64u) queue[base + k] = buf[k];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // reads done before the list is refilled
        n = 0;
    }
};

// The same staging for the hit queue, with each entry's shading class: a
// flush writes every class present in the list as one run of that class's
// queue segment, with one atomic on the class's counter.
struct ClassList {
    uint32_t* buf;   // 128 path slots of this wave
    uint8_t* cls;    // their shading classes
    uint32_t n;      // entries held (wave-uniform)
    __device__ __forceinline__ void add(bool pred, uint32_t v, uint32_t c) {
        const unsigned long long mask = __ballot(pred);
        const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << __lane_id()) - 1ull));
        if (pred) {
            buf[n + rank] = v;
            cls[n + rank] = (uint8_t)c;
        }
        n += (uint32_t)__popcll(mask);
    }
    // at most 127 entries are held (64 added at a time, flushed from 64 on)
    __device__ __forceinline__ void flush(const Frame1& f, uint32_t min) {
        if (n == 0 || n < min) return;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const uint32_t lane = (uint32_t)__lane_id();
        const bool v0 = lane < n, v1 = lane + 64u < n;
        const uint32_t p0 = v0 ? buf[lane] : 0u, p1 = v1 ? buf[lane + 64u] : 0u;
        const uint32_t c0 = v0 ? cls[lane] : (uint32_t)kShadeClasses, c1 = v1 ? cls[lane + 64u] : (uint32_t)kShadeClasses;
        const unsigned long long lt = (1ull << lane) - 1ull;
        for (uint32_t c = 0; c < kShadeClasses; ++c) {
            const unsigned long long m0 = __ballot(c0 == c), m1 = __ballot(c1 == c);
            const uint32_t k0 = (uint32_t)__popcll(m0), k = k0 + (uint32_t)__popcll(m1);
            if (k == 0u) continue;
            uint32_t base = 0;
            if (lane == 0u) base = atomicAdd(f.counters + hit_counter(c), k);
            base = __shfl(base, 0);
            uint32_t* __restrict__ q = f.p.hq + (size_t)c * f.NP + base;
            if (c0 == c) q[__popcll(m0 & lt)] = p0;
            if (c1 == c) q[k0 + (uint32_t)__popcll(m1 & lt)] = p1;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        n = 0;
    }
};

// The four-wide nodes a kernel traverses: 128-byte or quantised 64-byte
template <bool QN>
__device__ __forceinline__ auto scene_nodes(const SceneView& S) {
    if constexpr (QN) return S.nodes4q;
    else return S.nodes4;
}

// WAVES: minimum resident waves per SIMD the register allocation must allow
// (1 = unconstrained), traded by the launch shape against spills; QN: the
// quantised 64-byte nodes
// CAM0: the first bounce of a pass without depth of field, whose camera
// rays all start at the camera position, the origin of the camera-relative
// scene: the origin is the constant 0 there, which removes the origin
// subtractions of every box and triangle test (x - 0 is x exactly, so the
// walk is the same).
template <bool BATCH, bool FETCH, int WAVES, bool QN, bool CAM0>
__global__ __launch_bounds__(kTraceBlock, WAVES) void k_trace(Frame1 f, uint32_t qsel) {
    __shared__ int stack[kStackLds * kTraceBlock];
    __shared__ uint32_t staged[2 * kTraceBlock * 2];   // per wave: 128 hits, 128 misses
    __shared__ uint8_t stagedCls[2 * kTraceBlock];     // per wave: the hits' shading classes
    const TravStack<kStackLds> st = lane_stack(f, stack);
    const uint32_t count = f.counters[qsel ? CNT_QUEUE1 : CNT_QUEUE0];
    const uint32_t* __restrict__ queue = qsel ? f.p.q1 : f.p.q0;
    const bool identity = f.bounce == 0;
    uint32_t nrays = 0;
    // finished rays are staged per wave and queued 64 or more at a time;
    // hits in their shading class's segment when sorting by material
    const uint32_t wave = threadIdx.x >> 6;
    ClassList hits{staged + wave * 256u, stagedCls + wave * 128u, 0u};
    WaveList misses{staged + wave * 256u + 128u, 0u};
    if (FETCH) {
        // lanes take a new ray as soon as theirs completes
        TravRay ray;
        const Bvh4Node* nodes = nullptr;
        const TriRecord* tris = nullptr;
        uint32_t pid = 0;
        bool active = false;
        WaveChunk chunk;
        while (true) {
            const bool anyBusy = __ballot(active) != 0ull;
            const uint32_t idx = chunk.fetch(f.counters + CNT_TRACE_HEAD, count, !active, anyBusy);
            if (idx != 0xFFFFFFFFu) {
                pid = identity ? idx : queue[idx];
                const SceneView& S = scene_of<BATCH>(f, pid);
                nodes = S.nodes4;
                tris = S.tris;
                V3 o, d;
                pass_ray<BATCH>(f, S, pid, identity, o, d);
                if constexpr (CAM0) o = V3{0.0f, 0.0f, 0.0f};
                trav_init(ray, o, d, 3.402823466e+38f, S.nnodes4);
                // the nearest shape light bounds the walk; a nearer triangle replaces it
                if (S.nshapes) shape_closest(S.shapes, S.nshapes, ray.o, ray.d, 3.402823466e+38f, ray.h.t, ray.h.leaf);
                active = true;
                ++nrays;
            }
            if (chunk.done(__ballot(active) != 0ull)) break;
            bool done = false;
            if (active) done = trav_step<false>(nodes, tris, ray, st);
            const bool hit = done && ray.h.leaf != 0xFFFFFFFFu;
            uint32_t cls = 0;
            if (hit) {
                f.p.hit[pid] = make_float4(ray.h.t, ray.h.b1, ray.h.b2, __uint_as_float(ray.h.leaf));
                if (f.sortHits) cls = (ray.h.leaf & kShapeHit) ? (uint32_t)SC_SHAPE : tris[ray.h.leaf].shadeClass;
            }
            hits.add(hit, pid, cls);
            misses.add(done && !hit, pid);
            hits.flush(f, 64u);
            misses.flush(f.counters + CNT_MISS, f.p.mq, 64u);
            if (done) active = false;
        }
    } else {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < count; i += gridDim.x * blockDim.x) {
            bool hit = false;
            uint32_t cls = 0;
            uint32_t pid = 0;
            if (i < count) {
                pid = identity ? i : queue[i];
                const SceneView& S = scene_of<BATCH>(f, pid);
                V3 o, d;
                pass_ray<BATCH>(f, S, pid, identity, o, d);
                if constexpr (CAM0) o = V3{0.0f, 0.0f, 0.0f};
                // the nearest shape light bounds the walk; a nearer triangle replaces it
                float ts = 3.402823466e+38f;
                uint32_t shape = 0xFFFFFFFFu;
                if (S.nshapes) shape_closest(S.shapes, S.nshapes, o, d, ts, ts, shape);
                Hit h = traverse4<false>(scene_nodes<QN>(S), S.tris, S.nnodes4, o, d, ts, st);
                if (h.leaf == 0xFFFFFFFFu && shape != 0xFFFFFFFFu) h = Hit{ts, 0.0f, 0.0f, shape};
                hit = h.leaf != 0xFFFFFFFFu;
                if (hit) {
                    f.p.hit[pid] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.leaf));
                    // material-sorted shading: the hit joins its shading class's segment
                    if (f.sortHits) cls = (h.leaf & kShapeHit) ? (uint32_t)SC_SHAPE : S.tris[h.leaf].shadeClass;
                }
                ++nrays;
            }
            const bool live = i < count;
            hits.add(live && hit, pid, cls);
            misses.add(live && !hit, pid);
            hits.flush(f, 64u);
            misses.flush(f.counters + CNT_MISS, f.p.mq, 64u);
        }
    }
    hits.flush(f, 1u);
    misses.flush(f.counters + CNT_MISS, f.p.mq, 1u);
    wave_count(f.stats + STAT_CLOSEST, nrays);
}

// Network input of every queued hit, one thread per hit (the gathers run
// fully parallel), kept for the training records of the vertex.
__global__ __launch_bounds__(256) void k_guide_features(Frame1 f) {
    const HitQueue hq(f);
    const uint32_t count = hq.count();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        guide_input(f, f.sc, hq.entry(f, i), f.gfeat + (size_t)i * guide::kRec);
}

// Mixture of every queued hit: wave-resident MLP inference over the
// feature records (BF16 or FP8 E4M3 weights).
template <bool FP8>
__global__ __launch_bounds__(256, 2) void k_guide_mlp(Frame1 f) {
    __shared__ mlp::WaveLds L[4];
    mlp::wave_infer<FP8, true>(L[threadIdx.x >> 6], f.gfeat, HitQueue(f).count(), f.gwf, f.gwq, f.gws, f.gbias, f.gmix);
}

template <bool BATCH>
__global__ __launch_bounds__(256) void k_miss(Frame1 f) {
    const uint32_t count = f.counters[CNT_MISS];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t pid = f.p.mq[i];
        shade_miss(f, scene_of<BATCH>(f, pid), pid);
    }
}

template <bool BATCH, bool FETCH, int WAVES, bool QN>
__global__ __launch_bounds__(kTraceBlock, WAVES) void k_shadow(Frame1 f) {
    __shared__ int stack[kStackLds * kTraceBlock];
    const TravStack<kStackLds> st = lane_stack(f, stack);
    const uint32_t count = f.counters[CNT_SHADOW];
    uint32_t nrays = 0;
    if (FETCH) {
        TravRay ray;
        const Bvh4Node* nodes = nullptr;
        const TriRecord* tris = nullptr;
        uint32_t pid = 0;
        bool active = false;
        WaveChunk chunk;
        while (true) {
            const bool anyBusy = __ballot(active) != 0ull;
            const uint32_t idx = chunk.fetch(f.counters + CNT_SHADOW_HEAD, count, !active, anyBusy);
            if (idx != 0xFFFFFFFFu) {
                pid = f.p.sq[idx];
                const SceneView& S = scene_of<BATCH>(f, pid);
                nodes = S.nodes4;
                tris = S.tris;
                const float4 so = f.p.sro[pid], sd = f.p.srd[pid];
                const V3 o{so.x, so.y, so.z}, d{sd.x, sd.y, sd.z};
                // a shape light in the way settles the ray without a walk
                if (!(S.nshapes && shape_any(S.shapes, S.nshapes, o, d, so.w))) {
                    trav_init(ray, o, d, so.w, S.nnodes4);
                    active = true;
                }
                ++nrays;
            }
            if (chunk.done(__ballot(active) != 0ull)) break;
            if (active && trav_step<true>(nodes, tris, ray, st)) {
                if (ray.h.leaf == 0xFFFFFFFFu) add_nee(f, pid);
                active = false;
            }
        }
    } else {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < count; i += gridDim.x * blockDim.x) {
            if (i < count) {
                const uint32_t pid = f.p.sq[i];
                const SceneView& S = scene_of<BATCH>(f, pid);
                const float4 so = f.p.sro[pid], sd = f.p.srd[pid];
                const V3 o{so.x, so.y, so.z}, d{sd.x, sd.y, sd.z};
                if (!(S.nshapes && shape_any(S.shapes, S.nshapes, o, d, so.w))) {
                    const Hit h = traverse4<true>(scene_nodes<QN>(S), S.tris, S.nnodes4, o, d, so.w, st);
                    if (h.leaf == 0xFFFFFFFFu) add_nee(f, pid);
                }
                ++nrays;
            }
        }
    }
    wave_count(f.stats + STAT_SHADOW, nrays);
}

// WAVES: minimum resident waves per SIMD the register allocation must allow
// (1 = unconstrained); the autotuner picks between fewer spills and more
// latency hiding.
template <bool GUIDED, bool BATCH, int WAVES>
__global__ __launch_bounds__(kShadeBlock, WAVES) void k_shade(Frame1 f, uint32_t qsel) {
    // shades the paths whose ray hit a surface or a shape light (misses were finished by k_trace)
    const HitQueue hq(f);
    const uint32_t count = hq.count();
    uint32_t* __restrict__ qout = qsel ? f.p.q0 : f.p.q1;
    uint32_t* __restrict__ cout = f.counters + (qsel ? CNT_QUEUE0 : CNT_QUEUE1);
    uint32_t nvert = 0;
    // continuing paths and shadow rays are staged per wave and queued 64 or more at a time
    __shared__ uint32_t staged[2 * kShadeBlock * 2];
    const uint32_t wave = threadIdx.x >> 6;
    WaveList next{staged + wave * 256u, 0u}, shadows{staged + wave * 256u + 128u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i - threadIdx.x < count; i += gridDim.x * blockDim.x) {
        bool alive = false, shadow = false;
        bool recordSlot = false;      // guided + collecting: record this vertex's direction
        V3 recDir{0.0f, 0.0f, 0.0f};
        float recPb = 0.0f;           // BSDF sampler's density of the recorded direction
        S4 recT = s4(0.0f);           // throughput before the recorded vertex
        uint32_t pid = 0;
        if (i < count) {
            pid = hq.entry(f, i);
            ++nvert;
            const SceneView& S = scene_of<BATCH>(f, pid);
            shade_hit<GUIDED>(f, S, pid, i, alive, shadow, recordSlot, recDir, recPb, recT);
        }
        next.add(alive, pid);
        shadows.add(shadow, pid);
        next.flush(cout, qout, 64u);
        shadows.flush(f.counters + CNT_SHADOW, f.p.sq, 64u);
        if (GUIDED && f.collect) {
            // a recorded direction needs the path to go on (its radiance is measured later)
            const bool rec = recordSlot && alive;
            const uint32_t r = wave_push(f.grecCount, rec);
            if (rec && r < f.grecCap) {
                float* d = f.grec + (size_t)r * kPendingRec;
                const float* in = f.gfeat + (size_t)i * guide::kRec;
                for (int k = 0; k < 11; ++k) d[k] = in[k];
                d[11] = recDir.x;
                d[12] = recDir.y;
                d[13] = recDir.z;
                d[14] = recPb;
                d[15] = __uint_as_float(f.p.prec[pid]);
                d[16] = recT.v[0]; d[17] = recT.v[1]; d[18] = recT.v[2]; d[19] = recT.v[3];
                f.p.prec[pid] = r | kRecPending;
            }
        }
    }
    next.flush(cout, qout, 1u);
    shadows.flush(f.counters + CNT_SHADOW, f.p.sq, 1u);
    wave_count(f.stats + STAT_VERTICES, nvert);
}

// ----------------------------------------------------------- accumulation
template <bool BATCH>
__global__ __launch_bounds__(256) void k_accumulate(Frame1 f) {
    // per pixel, the S samples of the pass are added one at a time in sample
    // order, so any split of a sample range into passes gives the same sums
    const SceneView& S = BATCH ? f.scenes[blockIdx.y] : f.sc;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;   // scene-local slot of sample 0
    if (i >= S.P) return;
    accumulate_slot(f, S, i);
}

}  // namespace

// Expanded light-tree nodes: the packed nodes decoded once.
__global__ __launch_bounds__(256) void k_expand_tree(const PackedNode* __restrict__ in, TreeNodeX* __restrict__ out,
                                                     uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = expand_node(load_node(in, i));
}

// sRGB decode of the 256 8-bit channel values
__global__ void k_srgb_lut(float* __restrict__ out) {
    const uint32_t k = threadIdx.x;
    if (k < 256u) out[k] = srgb_decode((float)k / 255.0f);
}

// ------------------------------------------------------------------ host
hipError_t srgb_lut_fill(rbh_ctx c) {
    hipLaunchKernelGGL(k_srgb_lut, dim3(1), dim3(256), 0, c->stream, c->srgbLut.as<float>());
    return hipGetLastError();
}

hipError_t light_tree_expand(rbh_ctx c) {
    const uint32_t n = c->tree.nnodes;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand_tree, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->tree.nodes.as<PackedNode>(),
                       c->tree.nodesX.as<TreeNodeX>(), n);
    return hipGetLastError();
}

// Spill entries per lane of the traversal stack beyond the LDS part.
size_t trace_spill_words(size_t lanes) { return lanes * (size_t)(kMaxStack4 - kStackLds); }

template <bool BATCH, bool FETCH, bool QN>
static void launch_trace_q(rbh_ctx c, const Frame1& f, uint32_t qsel, const LaunchShape& ls) {
    const dim3 g(ls.traceGrid), b(kTraceBlock);
    if (f.bounce == 0u && f.camOrigin0) {
        if (ls.traceWaves >= 8) hipLaunchKernelGGL((k_trace<BATCH, FETCH, 8, QN, true>), g, b, 0, c->stream, f, qsel);
        else if (ls.traceWaves == 7) hipLaunchKernelGGL((k_trace<BATCH, FETCH, 7, QN, true>), g, b, 0, c->stream, f, qsel);
        else hipLaunchKernelGGL((k_trace<BATCH, FETCH, 1, QN, true>), g, b, 0, c->stream, f, qsel);
        return;
    }
    if (ls.traceWaves >= 8) hipLaunchKernelGGL((k_trace<BATCH, FETCH, 8, QN, false>), g, b, 0, c->stream, f, qsel);
    else if (ls.traceWaves == 7) hipLaunchKernelGGL((k_trace<BATCH, FETCH, 7, QN, false>), g, b, 0, c->stream, f, qsel);
    else hipLaunchKernelGGL((k_trace<BATCH, FETCH, 1, QN, false>), g, b, 0, c->stream, f, qsel);
}

// the work-fetch variants read the 128-byte nodes
template <bool BATCH>
static void launch_trace(rbh_ctx c, const Frame1& f, uint32_t qsel, const LaunchShape& ls) {
    if (ls.traceFetch) launch_trace_q<BATCH, true, false>(c, f, qsel, ls);
    else if (ls.quantNodes) launch_trace_q<BATCH, false, true>(c, f, qsel, ls);
    else launch_trace_q<BATCH, false, false>(c, f, qsel, ls);
}

template <bool BATCH, bool FETCH, bool QN>
static void launch_shadow_q(rbh_ctx c, const Frame1& f, const LaunchShape& ls) {
    const dim3 g(ls.traceGrid), b(kTraceBlock);
    if (ls.traceWaves >= 8) hipLaunchKernelGGL((k_shadow<BATCH, FETCH, 8, QN>), g, b, 0, c->stream, f);
    else if (ls.traceWaves == 7) hipLaunchKernelGGL((k_shadow<BATCH, FETCH, 7, QN>), g, b, 0, c->stream, f);
    else hipLaunchKernelGGL((k_shadow<BATCH, FETCH, 1, QN>), g, b, 0, c->stream, f);
}

template <bool BATCH>
static void launch_shadow(rbh_ctx c, const Frame1& f, const LaunchShape& ls) {
    if (ls.traceFetch) launch_shadow_q<BATCH, true, false>(c, f, ls);
    else if (ls.quantNodes) launch_shadow_q<BATCH, false, true>(c, f, ls);
    else launch_shadow_q<BATCH, false, false>(c, f, ls);
}

// One wavefront pass: camera rays for every slot, depthMax bounces of
// trace / miss / (guide) / shade / shadow, then accumulation. f.sc (single
// scene) or f.scenes (batched) must describe the pass's samples and slots.
template <bool BATCH>
static hipError_t launch_pass_t(rbh_ctx c, Frame1 f, uint32_t depthMax, const LaunchShape& ls, KernelTimer* kt) {
    // guiding: feature gather over all hit slots, then resident blocks of four
    // MLP waves, as many per CU as the register budget allows (weights and
    // the prefetched next records in registers: 2 waves per SIMD)
    const int nbGuide = (int)std::min<uint32_t>((f.NP + 255) / 256, 65535u * 16u);
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
    const int mlpGrid = 2 * cus;
    const dim3 accGrid((f.maxP + 255) / 256, BATCH ? f.nscenes : 1u);
    hipError_t e = hipSuccess;
#define RBH_TIMED(cls, launch)                                       \
    do {                                                             \
        if (kt && (e = kt->begin(c->stream, cls)) != hipSuccess) return e; \
        launch;                                                      \
        if (kt && (e = kt->end(c->stream)) != hipSuccess) return e;  \
    } while (0)
    f.bounce = 0;
    f.sortHits = ls.sortHits ? 1u : 0u;
    hipLaunchKernelGGL(k_pass_start, dim3(1), dim3(64), 0, c->stream, f);
    for (uint32_t b = 0; b < depthMax; ++b) {
        const uint32_t qsel = b & 1u;
        f.bounce = b;
        if (b > 0) hipLaunchKernelGGL(k_reset_hits, dim3(1), dim3(64), 0, c->stream, f.counters);
        RBH_TIMED(KernelTimer::TRACE, launch_trace<BATCH>(c, f, qsel, ls));
        RBH_TIMED(KernelTimer::SHADE,
                  hipLaunchKernelGGL(k_miss<BATCH>, dim3(ls.shadeGrid), dim3(256), 0, c->stream, f));
        if (!BATCH && f.guided) {
            RBH_TIMED(KernelTimer::GUIDE,
                      hipLaunchKernelGGL(k_guide_features, dim3(nbGuide), dim3(256), 0, c->stream, f));
            RBH_TIMED(KernelTimer::GUIDE,
                      if (f.gwq) hipLaunchKernelGGL((k_guide_mlp<true>), dim3(mlpGrid), dim3(256), 0, c->stream, f);
                      else hipLaunchKernelGGL((k_guide_mlp<false>), dim3(mlpGrid), dim3(256), 0, c->stream, f));
        }
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, c->stream, f.counters, qsel ? CNT_QUEUE0 : CNT_QUEUE1);
        if (!BATCH && f.guided && ls.shadeWaves >= 3)
            RBH_TIMED(KernelTimer::SHADE, hipLaunchKernelGGL((k_shade<true, false, 3>), dim3(ls.shadeGrid),
                                                             dim3(kShadeBlock), 0, c->stream, f, qsel));
        else if (!BATCH && f.guided)
            RBH_TIMED(KernelTimer::SHADE, hipLaunchKernelGGL((k_shade<true, false, 1>), dim3(ls.shadeGrid),
                                                             dim3(kShadeBlock), 0, c->stream, f, qsel));
        else if (ls.shadeWaves >= 4)
            RBH_TIMED(KernelTimer::SHADE, hipLaunchKernelGGL((k_shade<false, BATCH, 4>), dim3(ls.shadeGrid),
                                                             dim3(kShadeBlock), 0, c->stream, f, qsel));
        else if (ls.shadeWaves == 3)
            RBH_TIMED(KernelTimer::SHADE, hipLaunchKernelGGL((k_shade<false, BATCH, 3>), dim3(ls.shadeGrid),
                                                             dim3(kShadeBlock), 0, c->stream, f, qsel));
        else
            RBH_TIMED(KernelTimer::SHADE, hipLaunchKernelGGL((k_shade<false, BATCH, 1>), dim3(ls.shadeGrid),
                                                             dim3(kShadeBlock), 0, c->stream, f, qsel));
        RBH_TIMED(KernelTimer::SHADOW, launch_shadow<BATCH>(c, f, ls));
    }
#undef RBH_TIMED
    hipLaunchKernelGGL(k_accumulate<BATCH>, accGrid, dim3(256), 0, c->stream, f);
    return hipGetLastError();
}

hipError_t launch_pass(rbh_ctx c, const Frame1& f, uint32_t depthMax, const LaunchShape& ls, KernelTimer* kt) {
    return f.scenes ? launch_pass_t<true>(c, f, depthMax, ls, kt) : launch_pass_t<false>(c, f, depthMax, ls, kt);
}

}  // namespace rbh
