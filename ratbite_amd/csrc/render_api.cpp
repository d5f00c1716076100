// C ABI entry points of the acceleration structure, light tree and render
// loop (include/ratbite_hip.h).
#include <algorithm>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "guide.hpp"
#include "integrator.hpp"

namespace rbh {
int lbvh_build(rbh_ctx c);
hipError_t launch_pass(rbh_ctx c, Frame1 f, uint32_t depthMax, bool firstPass, int traceGrid, int shadeGrid,
                       KernelTimer* kt);
hipError_t resolve_launch(rbh_ctx c, const float* colorspace, float expo, void* out);
size_t trace_spill_words(size_t lanes);
namespace host {
int build_light_tree(const EmissiveTri* tris, const float* flux, uint32_t n, std::vector<PackedNode>& nodes,
                     std::vector<uint32_t>& pointers, std::vector<uint64_t>& masks);
}
}  // namespace rbh

using namespace rbh;

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Path-state buffers for P slots; returns the byte size and fills pointers
// when base != nullptr.
size_t layout_paths(char* base, size_t P, PathBuffers* pb, uint32_t** counters, unsigned long long** stats) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off = align256(off + bytes);
        return p;
    };
    float4** f4[] = {&pb->od, &pb->dr, &pb->lam, &pb->lpdf, &pb->L, &pb->T, &pb->nf, &pb->hit, &pb->Lnee, &pb->sro, &pb->srd};
    for (float4** q : f4) *q = reinterpret_cast<float4*>(take(P * sizeof(float4)));
    pb->depth = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->q0 = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->q1 = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->hq = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->mq = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->prec = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->sq = reinterpret_cast<uint32_t*>(take(P * 4));
    *counters = reinterpret_cast<uint32_t*>(take(CNT_COUNT * 4));
    *stats = reinterpret_cast<unsigned long long*>(take(STAT_COUNT * 8));
    return off;
}

__global__ void k_zero_stats(unsigned long long* s) {
    if (threadIdx.x < STAT_COUNT) s[threadIdx.x] = 0ull;
}

int cu_count(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    return n;
}

}  // namespace

extern "C" {

int rbh_build_accel(rbh_ctx c) {
    if (!c) return RBH_ERR_INVALID;
    return lbvh_build(c);
}

int rbh_build_light_tree(rbh_ctx c, const rbh_emissive_tri* tris, const float* flux, uint32_t n) {
    if (!c) return RBH_ERR_INVALID;
    if (n && (!tris || !flux)) return fail(c, RBH_ERR_INVALID, "light arrays are null");
    for (uint32_t i = 0; i < n; ++i) {
        if (!(flux[i] > 0.0f) || !std::isfinite(flux[i])) return fail(c, RBH_ERR_INVALID, "light %u: flux must be positive and finite", i);
        if (tris[i].materialIndex >= c->nmaterials) return fail(c, RBH_ERR_INVALID, "light %u: material %u of %u", i, tris[i].materialIndex, c->nmaterials);
    }
    LightTreeDev& t = c->tree;
    const EmissiveTri* et = reinterpret_cast<const EmissiveTri*>(tris);
    std::vector<PackedNode> nodes;
    const int r = host::build_light_tree(et, flux, n, nodes, t.hostPointers, t.hostMasks);
    if (r == -1) return fail(c, RBH_ERR_RANGE, "light count %u exceeds the tree's 27-bit leaf offsets", n);
    if (r < 0) return fail(c, RBH_ERR_INVALID, "light tree produced an invalid bitmask");
    t.hostNodes.resize(nodes.size());
    std::memcpy(t.hostNodes.data(), nodes.data(), nodes.size() * sizeof(PackedNode));
    t.nnodes = (uint32_t)nodes.size();
    t.nlights = n;
    if (n == 0) return RBH_OK;
    RBH_ALLOC(c, t.nodes, nodes.size() * sizeof(PackedNode));
    RBH_ALLOC(c, t.lightPointers, (size_t)n * 4);
    RBH_ALLOC(c, t.bitmasks, (size_t)n * 8);
    RBH_ALLOC(c, t.tris, (size_t)n * sizeof(EmissiveTri));
    RBH_HIP(c, hipMemcpyAsync(t.nodes.p, nodes.data(), nodes.size() * sizeof(PackedNode), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(t.lightPointers.p, t.hostPointers.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(t.bitmasks.p, t.hostMasks.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(t.tris.p, tris, (size_t)n * sizeof(EmissiveTri), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

int rbh_light_tree_build_host(const rbh_emissive_tri* tris, const float* flux, uint32_t n, rbh_packed_node* nodes,
                              uint32_t* lightPointers, uint64_t* bitmasks) {
    if (n && (!tris || !flux || !nodes || !lightPointers || !bitmasks)) return RBH_ERR_INVALID;
    std::vector<PackedNode> nv;
    std::vector<uint32_t> lp;
    std::vector<uint64_t> bm;
    const int r = host::build_light_tree(reinterpret_cast<const EmissiveTri*>(tris), flux, n, nv, lp, bm);
    if (r == -1) return RBH_ERR_RANGE;
    if (r < 0) return RBH_ERR_INVALID;
    if (r) {
        std::memcpy(nodes, nv.data(), nv.size() * sizeof(PackedNode));
        std::memcpy(lightPointers, lp.data(), lp.size() * 4);
        std::memcpy(bitmasks, bm.data(), bm.size() * 8);
    }
    return r;
}

int rbh_get_light_tree(rbh_ctx c, rbh_packed_node* nodes, uint32_t* nnodes, uint32_t* lightPointers, uint64_t* bitmasks) {
    if (!c) return RBH_ERR_INVALID;
    const LightTreeDev& t = c->tree;
    if (nnodes) *nnodes = t.nnodes;
    if (nodes && t.nnodes) std::memcpy(nodes, t.hostNodes.data(), (size_t)t.nnodes * sizeof(PackedNode));
    if (lightPointers && t.nlights) std::memcpy(lightPointers, t.hostPointers.data(), (size_t)t.nlights * 4);
    if (bitmasks && t.nlights) std::memcpy(bitmasks, t.hostMasks.data(), (size_t)t.nlights * 8);
    return RBH_OK;
}

int rbh_render(rbh_ctx c, const rbh_launch_params* lp) {
    if (!c || !lp) return RBH_ERR_INVALID;
    if (!c->accel.built) return fail(c, RBH_ERR_STATE, "rbh_build_accel has not been run on the current geometry");
    if (!c->haveColor || !c->haveLuts) return fail(c, RBH_ERR_STATE, "colour tables and LUTs must be uploaded first");
    if (lp->width == 0 || lp->height == 0) return fail(c, RBH_ERR_INVALID, "empty film");
    if ((uint64_t)lp->width * lp->height > (1ull << 28)) return fail(c, RBH_ERR_RANGE, "film above 2^28 pixels");
    if (lp->maxPathDepth == 0 || lp->maxPathDepth > 1024) return fail(c, RBH_ERR_INVALID, "maxPathDepth must be in [1, 1024]");
    const uint32_t known = RBH_RENDER_GUIDED | RBH_RENDER_COLLECT | RBH_RENDER_TIMING | RBH_RENDER_NO_DROP;
    if (lp->flags & ~known)
        return fail(c, RBH_ERR_INVALID, "render flags 0x%x are not supported by this entry point", lp->flags);
    const bool guided = (lp->flags & RBH_RENDER_GUIDED) != 0;
    const bool collect = (lp->flags & RBH_RENDER_COLLECT) != 0;
    if (collect && !guided) return fail(c, RBH_ERR_INVALID, "RBH_RENDER_COLLECT needs RBH_RENDER_GUIDED");
    GuideState* gst = static_cast<GuideState*>(c->guide);
    if (guided && !(gst && gst->configured)) return fail(c, RBH_ERR_STATE, "guided render before rbh_guide_configure");
    const bool treeOn = c->tree.nlights > 0;
    if (c->geom.maxLight >= (int64_t)c->tree.nlights)
        return fail(c, RBH_ERR_STATE, "triangle light index %lld has no light record (%u lights built)",
                    (long long)c->geom.maxLight, c->tree.nlights);
    RenderState& rs = c->rs;
    const size_t P = (size_t)lp->width * lp->height;
    // samples per pass: as many as fit the path-slot budget (at least one)
    const uint32_t spp = lp->sampleCount;
    const uint32_t S = (uint32_t)std::max<size_t>(1, std::min<size_t>(spp ? spp : 1, rs.slotBudget / P));
    const size_t NP = P * S;
    PathBuffers pb{};
    uint32_t* counters = nullptr;
    unsigned long long* stats = nullptr;
    const size_t bytes = layout_paths(nullptr, NP, &pb, &counters, &stats);
    RBH_ALLOC(c, rs.paths, bytes);
    layout_paths(rs.paths.as<char>(), NP, &pb, &counters, &stats);
    const bool sizeChanged = rs.width != lp->width || rs.height != lp->height;
    RBH_ALLOC(c, rs.accum, P * sizeof(double) * 4);
    if (sizeChanged && lp->sampleOffset != 0) return fail(c, RBH_ERR_STATE, "sampleOffset != 0 but no accumulation of this size exists");
    rs.width = lp->width;
    rs.height = lp->height;

    Frame1 f{};
    f.P = (uint32_t)P;
    f.W = lp->width;
    f.H = lp->height;
    f.maxDepth = std::max(1u, lp->maxPathDepth);
    f.maxRefl = std::min(f.maxDepth, lp->maxReflectedPathDepth);
    f.maxTrans = std::min(f.maxDepth, lp->maxTransmittedPathDepth);
    f.invW = 1.0f / (float)lp->width;
    f.invH = 1.0f / (float)lp->height;
    f.psW = lp->perspectiveScaleW;
    f.psH = lp->perspectiveScaleH;
    f.camU = dev::V3{lp->camU[0], lp->camU[1], lp->camU[2]};
    f.camV = dev::V3{lp->camV[0], lp->camV[1], lp->camV[2]};
    f.camW = dev::V3{lp->camW[0], lp->camW[1], lp->camW[2]};
    f.dof = lp->depthOfField;
    f.aperture = lp->aperture;
    f.focus = lp->focusDistance;
    f.nodes = c->accel.nodes.as<BvhNode>();
    f.tris = c->accel.tris.as<TriRecord>();
    f.nnodes = c->accel.nnodes;
    f.verts = c->geom.verts.as<float>();
    f.indices = c->geom.indices.as<uint32_t>();
    f.attribs = c->geom.attribs.as<uint8_t>();
    f.triMaterial = c->geom.triMaterial.as<uint32_t>();
    f.triLight = c->geom.triLight.as<uint32_t>();
    f.materials = c->materials.as<Material>();
    f.texDescs = c->textures.as<TextureDesc>();
    f.texels = c->texels.as<uint8_t>();
    f.spectra = c->spectra.as<float>();
    f.colors = c->colorTables.as<float>();
    f.luts = c->luts.as<float>();
    f.treeNodes = treeOn ? c->tree.nodes.as<PackedNode>() : nullptr;
    f.treePointers = treeOn ? c->tree.lightPointers.as<uint32_t>() : nullptr;
    f.treeMasks = treeOn ? c->tree.bitmasks.as<uint64_t>() : nullptr;
    f.lights = treeOn ? c->tree.tris.as<EmissiveTri>() : nullptr;
    f.envOn = c->env.enabled ? 1u : 0u;
    if (c->env.enabled) {
        f.envW = (uint32_t)c->env.width;
        f.envH = (uint32_t)c->env.height;
        f.envC = (uint32_t)c->env.channels;
        f.envTex = c->env.texels.as<float>();
        f.envCond = c->env.cond.as<uint16_t>();
        f.envMarg = c->env.marg.as<uint16_t>();
        f.envIntegral = c->env.integral;
    }
    f.p = pb;
    f.noDrop = (guided || (lp->flags & RBH_RENDER_NO_DROP)) ? 1u : 0u;
    if (guided) {
        RBH_ALLOC(c, rs.gfeat, NP * 16 * sizeof(float));
        RBH_ALLOC(c, rs.gmix, NP * 40 * sizeof(float));
        f.guided = 1u;
        f.gwf = static_cast<const __bf16*>(gst->wf);
        f.gbias = static_cast<const float*>(gst->bias);
        f.gfeat = rs.gfeat.as<float>();
        f.gmix = rs.gmix.as<float>();
        for (int a = 0; a < 3; ++a) {
            f.gmin[a] = gst->sceneMin[a];
            f.ginv[a] = gst->sceneInvExtent[a];
        }
        if (collect) {
            f.collect = 1u;
            f.grec = static_cast<float*>(gst->records);
            f.gtrain = static_cast<float*>(gst->trainRecs);
            f.grecCount = static_cast<uint32_t*>(gst->recCount);
            f.grecCap = gst->recCapacity;
            // keep about as many vertices as the buffer holds
            const double frac = std::min(1.0, (double)gst->recCapacity / ((double)P * std::max(1u, lp->sampleCount)));
            f.recThreshold = frac >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(frac * 4294967295.0);
            RBH_HIP(c, hipMemsetAsync(gst->recCount, 0, 4, c->stream));
        }
    }
    f.counters = counters;
    f.stats = stats;
    f.accum = rs.accum.as<double4>();

    const int cus = cu_count(c->device);
    const int traceGrid = cus * 16;
    const int shadeGrid = cus * 16;
    const size_t lanes = (size_t)traceGrid * kTraceBlock;
    RBH_ALLOC(c, rs.spill, trace_spill_words(lanes) * sizeof(int));
    f.spill = rs.spill.as<int>();
    f.spillStride = (uint32_t)lanes;
    KernelTimer* kt = (lp->flags & RBH_RENDER_TIMING) ? &rs.timer : nullptr;
    if (kt) kt->reset();
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, c->stream, stats);
    uint32_t passes = 0;
    for (uint32_t done = 0; done < spp;) {
        f.sample = lp->sampleOffset + done;
        f.S = std::min(S, spp - done);
        f.NP = (uint32_t)(P * f.S);
        const bool first = (lp->sampleOffset == 0 && done == 0);
        RBH_HIP(c, launch_pass(c, f, f.maxDepth, first, traceGrid, shadeGrid, kt));
        done += f.S;
        ++passes;
    }
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    RBH_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    unsigned long long st[STAT_COUNT] = {};
    RBH_HIP(c, hipMemcpy(st, stats, sizeof(st), hipMemcpyDeviceToHost));
    c->stats.renderMs = ms;
    c->stats.closestRays = st[STAT_CLOSEST];
    c->stats.shadowRays = st[STAT_SHADOW];
    c->stats.pathVertices = st[STAT_VERTICES];
    c->stats.samples = (uint64_t)P * lp->sampleCount;
    c->stats.traceLaunches = passes * f.maxDepth;
    c->stats.traceMs = c->stats.shadeMs = c->stats.shadowMs = c->stats.guideMs = 0.0;
    if (kt) {
        double kms[KernelTimer::NCLASS];
        uint32_t kn[KernelTimer::NCLASS];
        RBH_HIP(c, kt->collect(kms, kn));
        c->stats.traceMs = kms[KernelTimer::TRACE];
        c->stats.shadeMs = kms[KernelTimer::SHADE];
        c->stats.shadowMs = kms[KernelTimer::SHADOW];
        c->stats.guideMs = kms[KernelTimer::GUIDE];
    }
    return RBH_OK;
}

int rbh_set_path_slots(rbh_ctx c, uint64_t slots) {
    if (!c) return RBH_ERR_INVALID;
    if (slots == 0 || slots > (1ull << 31)) return fail(c, RBH_ERR_RANGE, "path slots must be in [1, 2^31]");
    c->rs.slotBudget = (size_t)slots;
    return RBH_OK;
}

int rbh_read_accum(rbh_ctx c, double* rgbw) {
    if (!c || !rgbw) return RBH_ERR_INVALID;
    if (!c->rs.accum.p) return fail(c, RBH_ERR_STATE, "nothing rendered");
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    RBH_HIP(c, hipMemcpy(rgbw, c->rs.accum.p, (size_t)c->rs.width * c->rs.height * 32, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_write_accum(rbh_ctx c, const double* rgbw, uint32_t w, uint32_t h) {
    if (!c || !rgbw || w == 0 || h == 0) return RBH_ERR_INVALID;
    RBH_ALLOC(c, c->rs.accum, (size_t)w * h * 32);
    RBH_HIP(c, hipMemcpyAsync(c->rs.accum.p, rgbw, (size_t)w * h * 32, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->rs.width = w;
    c->rs.height = h;
    return RBH_OK;
}

int rbh_resolve(rbh_ctx c, const float* colorspace, float exposureScale, uint8_t* rgba8) {
    if (!c || !colorspace || !rgba8) return RBH_ERR_INVALID;
    if (!c->rs.accum.p) return fail(c, RBH_ERR_STATE, "nothing rendered");
    const size_t n = (size_t)c->rs.width * c->rs.height;
    RBH_ALLOC(c, c->rs.image, n * 4);
    RBH_HIP(c, resolve_launch(c, colorspace, exposureScale, c->rs.image.p));
    RBH_HIP(c, hipMemcpyAsync(rgba8, c->rs.image.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

}  // extern "C"
