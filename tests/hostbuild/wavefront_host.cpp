// TEST INFRASTRUCTURE. A host build of the product's wavefront integrator:
// the per-path bodies of ratbite_amd/csrc/device/wavefront.hpp (camera,
// sky hit, surface shading, NEE term, FP64 accumulation) and the four-wide
// traversal of device/traverse.hpp, compiled as plain C++ against the shim
// in shim/hip/hip_runtime.h and driven by the same per-pass launch sequence
// as kernels/integrator.hip (single scene, plain and no-drop renders). Every
// buffer is its own heap allocation of the exact size, so a build with
// -fsanitize=address,undefined reports any out-of-range access or invalid
// conversion of the product code with its source line. The four-wide BVH is
// built here on the CPU (Morton keys with the triangle index in the low bits,
// Karras hierarchy, refit, top-down collapse that opens the largest inner
// child first), the same construction as kernels/lbvh.hip.
//
// Besides the sanitizers, the driver checks the state the kernels hand to
// each other (hit records, ray directions and origins of live paths, queue
// counts) and counts non-finite values, so that the paths that carry them can
// be found.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <omp.h>

#include "../../include/ratbite_hip.h"
#include "../../ratbite_amd/csrc/device/traverse.hpp"
#include "../../ratbite_amd/csrc/device/wavefront.hpp"

using namespace rbh;
using namespace rbh::dev;
using namespace rbh::wf;

namespace {

constexpr int kLdsDepth = 16;
constexpr int kSpill = kMaxStack4 - kLdsDepth;

template <class T>
T* copy_of(const void* src, size_t n) {
    if (n == 0) return nullptr;
    T* p = new T[n];
    std::memcpy(p, src, n * sizeof(T));
    return p;
}

uint32_t push(uint32_t* counter) { return __atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED); }

// ---------------------------------------------------------------- LBVH (CPU)
struct Bvh4 {
    std::vector<Bvh4Node> nodes;
    std::vector<TriRecord> tris;
    uint32_t nnodes4 = 0;
    uint32_t depth = 0;
};

uint64_t spread21(uint32_t x) {
    uint64_t y = x & 0x1FFFFFu;
    y = (y | (y << 32)) & 0x001F00000000FFFFull;
    y = (y | (y << 16)) & 0x001F0000FF0000FFull;
    y = (y | (y << 8)) & 0x100F00F00F00F00Full;
    y = (y | (y << 4)) & 0x10C30C30C30C30C3ull;
    y = (y | (y << 2)) & 0x1249249249249249ull;
    return y;
}

int prefix(const std::vector<uint64_t>& k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __builtin_clzll(k[i] ^ k[j]);
}

float half_area(const float* b) {
    const float x = b[3] - b[0], y = b[4] - b[1], z = b[5] - b[2];
    return x * y + y * z + z * x;
}

int build_bvh(const float* v, const uint32_t* idx, uint32_t n, const uint32_t* triMat, const uint32_t* triLight, Bvh4& out) {
    if (n == 0) return -1;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> cen(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const float a = v[3 * idx[3 * i] + k], b = v[3 * idx[3 * i + 1] + k], c = v[3 * idx[3 * i + 2] + k];
            const float ce = 0.5f * (std::fmin(std::fmin(a, b), c) + std::fmax(std::fmax(a, b), c));
            cen[3 * (size_t)i + k] = ce;
            lo[k] = std::fmin(lo[k], ce);
            hi[k] = std::fmax(hi[k], ce);
        }
    int idxBits = 1;
    while (idxBits < 32 && ((uint64_t)1 << idxBits) < (uint64_t)n) ++idxBits;
    const int axisBits = std::min(21, (64 - idxBits) / 3);
    const float cells = (float)(1u << axisBits);
    std::vector<uint64_t> keys(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            const float ext = hi[k] - lo[k];
            const float u = ext > 0.0f ? (cen[3 * (size_t)i + k] - lo[k]) / ext : 0.5f;
            q[k] = (uint32_t)std::fmin(std::fmax(u * cells, 0.0f), cells - 1.0f);
        }
        keys[i] = (((spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2])) << idxBits) | i;
    }
    std::sort(keys.begin(), keys.end());
    out.tris.resize(n);
    std::vector<float> lbox(6 * (size_t)n);
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t t = (uint32_t)(keys[j] & ((1ull << idxBits) - 1ull));
        const uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
        TriRecord r;
        r.v0x = v[3 * a]; r.v0y = v[3 * a + 1]; r.v0z = v[3 * a + 2]; r.prim = t;
        r.v1x = v[3 * b]; r.v1y = v[3 * b + 1]; r.v1z = v[3 * b + 2]; r.material = triMat[t];
        r.v2x = v[3 * c]; r.v2y = v[3 * c + 1]; r.v2z = v[3 * c + 2]; r.light = triLight[t];
        r.i0 = a; r.i1 = b; r.i2 = c; r.shadeClass = 0u;
        out.tris[j] = r;
        float* lb = &lbox[6 * (size_t)j];
        lb[0] = std::fmin(std::fmin(r.v0x, r.v1x), r.v2x); lb[3] = std::fmax(std::fmax(r.v0x, r.v1x), r.v2x);
        lb[1] = std::fmin(std::fmin(r.v0y, r.v1y), r.v2y); lb[4] = std::fmax(std::fmax(r.v0y, r.v1y), r.v2y);
        lb[2] = std::fmin(std::fmin(r.v0z, r.v1z), r.v2z); lb[5] = std::fmax(std::fmax(r.v0z, r.v1z), r.v2z);
    }
    if (n == 1) {
        out.nnodes4 = 0;
        out.nodes.assign(1, Bvh4Node{});
        return 0;
    }
    // Karras 2012 hierarchy over the sorted keys
    const int ni = (int)n - 1;
    std::vector<int2> child(ni);
    for (int i = 0; i < ni; ++i) {
        const int nn = (int)n;
        const int d = (prefix(keys, nn, i, i + 1) - prefix(keys, nn, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = prefix(keys, nn, i, i - d);
        int lmax = 2;
        while (prefix(keys, nn, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (prefix(keys, nn, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = prefix(keys, nn, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (prefix(keys, nn, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int split = i + s * d + std::min(d, 0);
        const int first = std::min(i, j), last = std::max(i, j);
        child[i] = int2{first == split ? ~split : split, last == split + 1 ? ~(split + 1) : split + 1};
    }
    // refit in post order; depth of the deepest leaf
    std::vector<float> ibox(6 * (size_t)ni);
    std::vector<std::pair<int, int>> st;   // (node, state)
    std::vector<int> ndepth(ni, 0);
    st.push_back({0, 0});
    uint32_t maxDepth = 0;
    while (!st.empty()) {
        auto& top = st.back();
        const int nd = top.first;
        const int2 ch = child[nd];
        if (top.second == 0) {
            top.second = 1;
            for (int c : {ch.x, ch.y}) {
                if (c >= 0) {
                    ndepth[c] = ndepth[nd] + 1;
                    st.push_back({c, 0});
                } else {
                    maxDepth = std::max<uint32_t>(maxDepth, (uint32_t)ndepth[nd] + 1);
                }
            }
            continue;
        }
        const float* A = ch.x >= 0 ? &ibox[6 * (size_t)ch.x] : &lbox[6 * (size_t)(~ch.x)];
        const float* B = ch.y >= 0 ? &ibox[6 * (size_t)ch.y] : &lbox[6 * (size_t)(~ch.y)];
        float* o = &ibox[6 * (size_t)nd];
        for (int k = 0; k < 3; ++k) {
            o[k] = std::fmin(A[k], B[k]);
            o[3 + k] = std::fmax(A[3 + k], B[3 + k]);
        }
        st.pop_back();
    }
    out.depth = maxDepth;
    if (maxDepth > 64) return -2;
    // collapse top down
    out.nodes.assign(ni, Bvh4Node{});
    std::vector<int2> level{int2{0, 0}}, next;
    uint32_t count = 1;
    while (!level.empty()) {
        next.clear();
        for (const int2 w : level) {
            int slot[4];
            int ns = 2;
            slot[0] = child[w.x].x;
            slot[1] = child[w.x].y;
            while (ns < 4) {
                int best = -1;
                float area = -1.0f;
                for (int k = 0; k < ns; ++k)
                    if (slot[k] >= 0) {
                        const float a = half_area(&ibox[6 * (size_t)slot[k]]);
                        if (a > area) { area = a; best = k; }
                    }
                if (best < 0) break;
                const int2 ch = child[slot[best]];
                slot[best] = ch.x;
                slot[ns++] = ch.y;
            }
            Bvh4Node nd{};
            for (int k = 0; k < 4; ++k) {
                if (k >= ns) {
                    nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
                    nd.hix[k] = nd.hiy[k] = nd.hiz[k] = -INFINITY;
                    nd.child[k] = kBvh4Empty;
                    continue;
                }
                const int c = slot[k];
                const float* b = c >= 0 ? &ibox[6 * (size_t)c] : &lbox[6 * (size_t)(~c)];
                nd.lox[k] = b[0]; nd.loy[k] = b[1]; nd.loz[k] = b[2];
                nd.hix[k] = b[3]; nd.hiy[k] = b[4]; nd.hiz[k] = b[5];
                if (c >= 0) {
                    const int id = (int)count++;
                    nd.child[k] = id;
                    next.push_back(int2{c, id});
                } else {
                    nd.child[k] = c;
                }
            }
            out.nodes[w.y] = nd;
        }
        std::swap(level, next);
    }
    out.nnodes4 = count;
    out.nodes.resize(count);
    return 0;
}

// ------------------------------------------------------------- anomaly log
struct Anomalies {
    std::atomic<uint64_t> hitNonFinite{0}, rayNonFinite{0}, shadowNonFinite{0}, queueOverflow{0};
    std::atomic<int> printed{0};
    int maxPrint = 20;
    void note(const char* what, const SceneView& S, uint32_t pid, uint32_t bounce, float4 a, float4 b) {
        if (printed.fetch_add(1) >= maxPrint) return;
        const uint32_t sk = pid / S.P, pix = pid - sk * S.P;
        uint32_t x, y;
        slot_pixel(S, pix, x, y);
        std::fprintf(stderr, "[hostbuild] %s: sample %u pixel (%u,%u) bounce %u: (%g %g %g %g) (%g %g %g %g)\n", what,
                     S.sample + sk, x, y, bounce, a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
    }
};

bool finite4(float4 v, bool w = true) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z) && (!w || std::isfinite(v.w));
}

}  // namespace

extern "C" {

// Scene arrays as the rbh_upload_* entry points take them, plus the built
// light tree (packed nodes, pointers, bitmasks) and env-map tables.
struct HostScene {
    const float* verts; uint32_t nverts;
    const uint32_t* indices; uint32_t ntris;
    const uint8_t* attribs; uint64_t attribBytes;
    const uint32_t* triMaterial;
    const uint32_t* triLight;
    const void* materials; uint32_t nmaterials;
    const void* textures; uint32_t ntextures;
    const uint8_t* texels; uint64_t texelBytes;
    const float* spectra; uint32_t nspectra;
    const float* colors;       // 6 x 471: X, Y, Z, basis R, G, B
    const float* luts;         // kLutTotal floats in the context's LUT layout
    const void* lights; uint32_t nlights;
    const void* treeNodes; uint32_t ntreeNodes;
    const uint32_t* treePointers;
    const uint64_t* treeMasks;
    uint32_t envOn, envW, envH, envC;
    const float* envTex;
    const uint16_t* envCond;
    const uint16_t* envMarg;
    float envIntegral;
};

struct HostStats {
    uint64_t closestRays, shadowRays, pathVertices, nonFiniteSamples;
    uint64_t hitNonFinite, rayNonFinite, shadowNonFinite, queueOverflow;
    uint32_t bvhDepth, nnodes4;
};

// Renders lp->sampleCount samples from lp->sampleOffset into accum (W*H*4
// doubles, read when the launch continues an accumulation) in passes of at
// most passSamples samples per pixel, like rbh_render. Returns 0, or a
// negative code for invalid input.
int host_render(const HostScene* hs, const rbh_launch_params* lp, uint32_t passSamples, int threads, int sortHits,
                double* accum, HostStats* stats) {
    if (!hs || !lp || !accum || !stats || lp->width == 0 || lp->height == 0 || lp->maxPathDepth == 0) return -1;
    if (lp->flags & ~(uint32_t)(RBH_RENDER_NO_DROP | RBH_RENDER_TIMING)) return -2;
    if (threads > 0) omp_set_num_threads(threads);
    Bvh4 bvh;
    if (build_bvh(hs->verts, hs->indices, hs->ntris, hs->triMaterial, hs->triLight, bvh) != 0) return -3;
    // exact-size copies: the sanitizer's bounds are the scene's
    Bvh4Node* nodes4 = copy_of<Bvh4Node>(bvh.nodes.data(), bvh.nodes.size());
    TriRecord* tris = copy_of<TriRecord>(bvh.tris.data(), bvh.tris.size());
    uint8_t* attribs = copy_of<uint8_t>(hs->attribs, hs->attribBytes ? hs->attribBytes : 0);
    Material* materials = copy_of<Material>(hs->materials, hs->nmaterials);
    TextureDesc* texDescs = copy_of<TextureDesc>(hs->textures, hs->ntextures);
    uint8_t* texels = copy_of<uint8_t>(hs->texels, hs->texelBytes);
    float* spectra = copy_of<float>(hs->spectra, (size_t)hs->nspectra * kDense);
    float* colors = copy_of<float>(hs->colors, 6 * (size_t)kDense);
    float* luts = copy_of<float>(hs->luts, kLutTotal);
    EmissiveTri* lights = copy_of<EmissiveTri>(hs->lights, hs->nlights);
    TreeNodeX* treeX = nullptr;
    uint32_t* treePtr = nullptr;
    uint64_t* treeMask = nullptr;
    if (hs->nlights) {
        treeX = new TreeNodeX[hs->ntreeNodes];
        const PackedNode* pn = static_cast<const PackedNode*>(hs->treeNodes);
        for (uint32_t i = 0; i < hs->ntreeNodes; ++i) treeX[i] = expand_node(load_node(pn, i));
        treePtr = copy_of<uint32_t>(hs->treePointers, hs->nlights);
        treeMask = copy_of<uint64_t>(hs->treeMasks, hs->nlights);
    }
    float* srgb = new float[256];
    for (int k = 0; k < 256; ++k) srgb[k] = srgb_decode((float)k / 255.0f);
    float* envTex = nullptr;
    uint16_t *envCond = nullptr, *envMarg = nullptr;
    if (hs->envOn) {
        envTex = copy_of<float>(hs->envTex, (size_t)hs->envW * hs->envH * hs->envC);
        envCond = copy_of<uint16_t>(hs->envCond, (size_t)hs->envW * hs->envH);
        envMarg = copy_of<uint16_t>(hs->envMarg, hs->envH);
    }
    const size_t P = (size_t)lp->width * lp->height;
    double4* acc = new double4[P];
    std::memcpy(acc, accum, P * sizeof(double4));

    SceneView v{};
    v.P = (uint32_t)P;
    v.W = lp->width;
    v.H = lp->height;
    v.maxDepth = std::max(1u, lp->maxPathDepth);
    v.maxRefl = std::min(v.maxDepth, lp->maxReflectedPathDepth);
    v.maxTrans = std::min(v.maxDepth, lp->maxTransmittedPathDepth);
    v.invW = 1.0f / (float)lp->width;
    v.invH = 1.0f / (float)lp->height;
    v.psW = lp->perspectiveScaleW;
    v.psH = lp->perspectiveScaleH;
    v.camU = V3{lp->camU[0], lp->camU[1], lp->camU[2]};
    v.camV = V3{lp->camV[0], lp->camV[1], lp->camV[2]};
    v.camW = V3{lp->camW[0], lp->camW[1], lp->camW[2]};
    v.dof = lp->depthOfField;
    v.aperture = lp->aperture;
    v.focus = lp->focusDistance;
    v.nodes4 = nodes4;
    v.nnodes4 = bvh.nnodes4;
    v.tris = tris;
    v.attribs = attribs;
    v.materials = materials;
    v.texDescs = texDescs;
    v.texels = texels;
    v.srgbLut = srgb;
    v.spectra = spectra;
    v.colors = colors;
    v.luts = luts;
    v.treeNodes = treeX;
    v.treePointers = treePtr;
    v.treeMasks = treeMask;
    v.lights = lights;
    v.envOn = hs->envOn;
    v.envW = hs->envW;
    v.envH = hs->envH;
    v.envC = hs->envC;
    v.envTex = envTex;
    v.envCond = envCond;
    v.envMarg = envMarg;
    v.envIntegral = hs->envIntegral;
    v.accum = acc;

    const uint32_t spp = lp->sampleCount;
    const uint32_t S = std::max(1u, std::min(passSamples ? passSamples : spp, std::max(1u, spp)));
    const size_t NPmax = P * S;
    Frame1 f{};
    f.sc = v;
    PathBuffers& pb = f.p;
    pb.od = new float4[NPmax]; pb.dr = new float4[NPmax]; pb.L = new float4[NPmax]; pb.T = new float4[NPmax];
    pb.nf = new float4[NPmax]; pb.hit = new float4[NPmax]; pb.Lnee = new float4[NPmax];
    pb.sro = new float4[NPmax]; pb.srd = new float4[NPmax];
    pb.depth = new uint32_t[NPmax];
    pb.q0 = new uint32_t[NPmax]; pb.q1 = new uint32_t[NPmax]; pb.hq = new uint32_t[NPmax * kShadeClasses];
    pb.mq = new uint32_t[NPmax]; pb.sq = new uint32_t[NPmax]; pb.prec = new uint32_t[NPmax];
    uint32_t* counters = new uint32_t[CNT_COUNT]();
    unsigned long long* st = new unsigned long long[STAT_COUNT]();
    f.counters = counters;
    f.stats = st;
    f.noDrop = (lp->flags & RBH_RENDER_NO_DROP) ? 1u : 0u;
    f.sortHits = sortHits ? 1u : 0u;
    Anomalies an;
    const char* mp = std::getenv("HOSTBUILD_MAX_PRINT");
    if (mp) an.maxPrint = std::atoi(mp);
    const bool progress = std::getenv("HOSTBUILD_PROGRESS") != nullptr;

    for (uint32_t done = 0; done < spp;) {
        f.sc.sample = lp->sampleOffset + done;
        f.sc.S = std::min(S, spp - done);
        f.sc.slot0 = 0;
        f.sc.first = (lp->sampleOffset == 0 && done == 0) ? 1u : 0u;
        f.NP = (uint32_t)(P * f.sc.S);
        f.maxP = (uint32_t)P;
        const uint32_t NP = f.NP;
        const SceneView& sv = f.sc;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)NP; ++i) camera_slot(f, sv, (uint32_t)i);
        std::fill(counters, counters + CNT_COUNT, 0u);
        counters[CNT_QUEUE0] = NP;
        for (uint32_t b = 0; b < sv.maxDepth; ++b) {
            const uint32_t qsel = b & 1u;
            f.bounce = b;
            if (b > 0) {
                counters[CNT_MISS] = counters[CNT_TRACE_HEAD] = 0u;
                for (uint32_t c = 0; c < kShadeClasses; ++c) counters[hit_counter(c)] = 0u;
            }
            // closest hit (k_trace, wave-synchronous form)
            {
                const uint32_t count = counters[qsel ? CNT_QUEUE1 : CNT_QUEUE0];
                const uint32_t* queue = qsel ? pb.q1 : pb.q0;
                uint64_t nrays = 0;
#pragma omp parallel reduction(+ : nrays)
                {
                    int lds[kLdsDepth];
                    int* spill = new int[kSpill];
                    const TravStack<kLdsDepth> stk{lds, 1, spill, 1u};
#pragma omp for schedule(dynamic, 4096)
                    for (int64_t ii = 0; ii < (int64_t)count; ++ii) {
                        const uint32_t pid = b == 0 ? (uint32_t)ii : queue[ii];
                        const float4 od = pb.od[pid], dr = pb.dr[pid];
                        if (!finite4(od, false) || !finite4(dr, false)) {
                            ++an.rayNonFinite;
                            an.note("non-finite ray traced", sv, pid, b, od, dr);
                        }
                        const Hit h = traverse4<false>(nodes4, tris, sv.nnodes4, V3{od.x, od.y, od.z}, V3{dr.x, dr.y, dr.z},
                                                       3.402823466e+38f, stk);
                        ++nrays;
                        if (h.leaf != 0xFFFFFFFFu) {
                            pb.hit[pid] = make_float4(h.t, h.b1, h.b2, __uint_as_float(h.leaf));
                            if (!(std::isfinite(h.t) && std::isfinite(h.b1) && std::isfinite(h.b2))) {
                                ++an.hitNonFinite;
                                an.note("non-finite hit", sv, pid, b, pb.hit[pid], dr);
                            }
                            // material-sorted shading: one queue segment per shading class
                            const uint32_t cls = f.sortHits ? shade_class(materials[tris[h.leaf].material]) : 0u;
                            pb.hq[(size_t)cls * NP + push(counters + hit_counter(cls))] = pid;
                        } else {
                            pb.mq[push(counters + CNT_MISS)] = pid;
                        }
                    }
                    delete[] spill;
                }
                st[STAT_CLOSEST] += nrays;
                uint64_t hits = 0;
                for (uint32_t c = 0; c < kShadeClasses; ++c) hits += counters[hit_counter(c)];
                if (hits > NP) { ++an.queueOverflow; break; }
            }
            // sky hits (k_miss)
            {
                const uint32_t count = counters[CNT_MISS];
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < (int64_t)count; ++i) shade_miss(f, sv, pb.mq[i]);
            }
            counters[qsel ? CNT_QUEUE0 : CNT_QUEUE1] = 0u;
            counters[CNT_SHADOW] = 0u;
            counters[CNT_SHADOW_HEAD] = 0u;
            // surface shading (k_shade)
            {
                uint32_t start[kShadeClasses + 1] = {0u};
                for (uint32_t c = 0; c < kShadeClasses; ++c) start[c + 1] = start[c] + counters[hit_counter(c)];
                const uint32_t count = start[kShadeClasses];
                uint32_t* qout = qsel ? pb.q0 : pb.q1;
                uint32_t* cout = counters + (qsel ? CNT_QUEUE0 : CNT_QUEUE1);
#pragma omp parallel for schedule(dynamic, 4096)
                for (int64_t ii = 0; ii < (int64_t)count; ++ii) {
                    const uint32_t i = (uint32_t)ii;
                    uint32_t cls = 0;
                    while (cls + 1 < kShadeClasses && i >= start[cls + 1]) ++cls;
                    const uint32_t pid = pb.hq[(size_t)cls * NP + (i - start[cls])];
                    bool alive = false, shadow = false, rec = false;
                    V3 recDir{0.0f, 0.0f, 0.0f};
                    float recPb = 0.0f;
                    S4 recT = s4(0.0f);
                    shade_hit<false>(f, sv, pid, i, alive, shadow, rec, recDir, recPb, recT);
                    if (alive) {
                        qout[push(cout)] = pid;
                        const float4 od = pb.od[pid], dr = pb.dr[pid];
                        if (!finite4(od, false) || !finite4(dr, false)) {
                            ++an.rayNonFinite;
                            an.note("live path with a non-finite ray", sv, pid, b, od, dr);
                        }
                    }
                    if (shadow) {
                        pb.sq[push(counters + CNT_SHADOW)] = pid;
                        if (!finite4(pb.sro[pid]) || !finite4(pb.srd[pid], false)) {
                            ++an.shadowNonFinite;
                            an.note("non-finite shadow ray", sv, pid, b, pb.sro[pid], pb.srd[pid]);
                        }
                    }
                }
                st[STAT_VERTICES] += count;
            }
            // shadow rays (k_shadow)
            {
                const uint32_t count = counters[CNT_SHADOW];
#pragma omp parallel
                {
                    int lds[kLdsDepth];
                    int* spill = new int[kSpill];
                    const TravStack<kLdsDepth> stk{lds, 1, spill, 1u};
#pragma omp for schedule(dynamic, 4096)
                    for (int64_t i = 0; i < (int64_t)count; ++i) {
                        const uint32_t pid = pb.sq[i];
                        const float4 so = pb.sro[pid], sd = pb.srd[pid];
                        const Hit h = traverse4<true>(nodes4, tris, sv.nnodes4, V3{so.x, so.y, so.z}, V3{sd.x, sd.y, sd.z},
                                                      so.w, stk);
                        if (h.leaf == 0xFFFFFFFFu) add_nee(f, pid);
                    }
                    delete[] spill;
                }
                st[STAT_SHADOW] += count;
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)P; ++i) accumulate_slot(f, sv, (uint32_t)i);
        done += f.sc.S;
        if (progress)
            std::fprintf(stderr, "[hostbuild] samples %u..%u done (%llu closest, %llu shadow rays)\n", sv.sample,
                         sv.sample + sv.S - 1, (unsigned long long)st[STAT_CLOSEST], (unsigned long long)st[STAT_SHADOW]);
    }
    std::memcpy(accum, acc, P * sizeof(double4));
    stats->closestRays = st[STAT_CLOSEST];
    stats->shadowRays = st[STAT_SHADOW];
    stats->pathVertices = st[STAT_VERTICES];
    stats->nonFiniteSamples = st[STAT_NONFINITE];
    stats->hitNonFinite = an.hitNonFinite;
    stats->rayNonFinite = an.rayNonFinite;
    stats->shadowNonFinite = an.shadowNonFinite;
    stats->queueOverflow = an.queueOverflow;
    stats->bvhDepth = bvh.depth;
    stats->nnodes4 = bvh.nnodes4;

    for (void* q : {(void*)pb.od, (void*)pb.dr, (void*)pb.L, (void*)pb.T, (void*)pb.nf, (void*)pb.hit, (void*)pb.Lnee,
                    (void*)pb.sro, (void*)pb.srd})
        delete[] static_cast<float4*>(q);
    for (uint32_t* q : {pb.depth, pb.q0, pb.q1, pb.hq, pb.mq, pb.sq, pb.prec, counters}) delete[] q;
    delete[] st;
    delete[] acc;
    delete[] nodes4; delete[] tris; delete[] attribs; delete[] materials; delete[] texDescs; delete[] texels;
    delete[] spectra; delete[] colors; delete[] luts; delete[] lights; delete[] treeX; delete[] treePtr;
    delete[] treeMask; delete[] srgb; delete[] envTex; delete[] envCond; delete[] envMarg;
    return 0;
}

}  // extern "C"
