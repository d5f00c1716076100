// C ABI entry points of the acceleration structure, light tree and render
// loop (include/ratbite_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "guide.hpp"
#include "integrator.hpp"

namespace rbh {
int lbvh_build(rbh_ctx c);
int lbvh_refit(rbh_ctx c);
hipError_t translate_points(rbh_ctx c, float* p, size_t npoints, uint32_t strideFloats, const float* delta);
hipError_t launch_pass(rbh_ctx c, const Frame1& f, uint32_t depthMax, const LaunchShape& ls, KernelTimer* kt);
hipError_t resolve_launch(rbh_ctx c, const float* colorspace, float expo, void* out);
size_t trace_spill_words(size_t lanes);
hipError_t light_tree_expand(rbh_ctx c);
hipError_t refresh_tri_classes(rbh_ctx c);
namespace host {
int build_light_tree(const EmissiveTri* tris, const float* flux, uint32_t n, const ShapeLight* shapes,
                     const float* shapeFlux, uint32_t m, std::vector<PackedNode>& nodes,
                     std::vector<uint32_t>& pointers, std::vector<uint64_t>& masks, bool fp64);
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
    float4** f4[] = {&pb->od, &pb->dr, &pb->L, &pb->T, &pb->nf, &pb->hit, &pb->Lnee, &pb->sro, &pb->srd};
    for (float4** q : f4) *q = reinterpret_cast<float4*>(take(P * sizeof(float4)));
    pb->depth = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->q0 = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->q1 = reinterpret_cast<uint32_t*>(take(P * 4));
    pb->hq = reinterpret_cast<uint32_t*>(take(P * 4 * kShadeClasses));   // one segment per shading class
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
    return rbh_build_light_tree_ex(c, tris, flux, n, 0u);
}

int rbh_refit_accel(rbh_ctx c) {
    if (!c) return RBH_ERR_INVALID;
    return lbvh_refit(c);
}

int rbh_update_vertices(rbh_ctx c, uint32_t first, uint32_t count, const float* verts) {
    if (!c || (count && !verts)) return RBH_ERR_INVALID;
    Geometry& g = c->geom;
    if ((uint64_t)first + count > g.nverts)
        return fail(c, RBH_ERR_RANGE, "vertices %u..%u of %u", first, first + count, g.nverts);
    if (!count) return RBH_OK;
    RBH_HIP(c, hipMemcpyAsync(g.verts.as<char>() + (size_t)first * 12, verts, (size_t)count * 12,
                              hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->accel.stale = true;
    return RBH_OK;
}

int rbh_translate_scene(rbh_ctx c, const float* delta) {
    if (!c || !delta) return RBH_ERR_INVALID;
    if (!c->accel.built) return fail(c, RBH_ERR_STATE, "rbh_build_accel first");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(delta[a])) return fail(c, RBH_ERR_INVALID, "translation must be finite");
    Geometry& g = c->geom;
    RBH_HIP(c, translate_points(c, g.verts.as<float>(), g.nverts, 3, delta));
    LightTreeDev& t = c->tree;
    ShapeLights& sh = c->shapes;
    if (sh.n) {
        for (auto& r : sh.host)
            for (int a = 0; a < 3; ++a) r.position[a] -= delta[a];
        RBH_HIP(c, hipMemcpyAsync(sh.recs.p, sh.host.data(), (size_t)sh.n * sizeof(rbh_shape_light),
                                  hipMemcpyHostToDevice, c->stream));
    }
    if (t.nlights) {
        // emitter corners (9 floats at the start of each 68-byte record)
        RBH_HIP(c, translate_points(c, t.tris.as<float>(), (size_t)t.nlights * 3, 0, delta));
    }
    if (t.nnodes) {
        for (auto& nd : t.hostNodes)
            for (int a = 0; a < 3; ++a) nd.spatialMean[a] -= delta[a];
        RBH_HIP(c, hipMemcpyAsync(t.nodes.p, t.hostNodes.data(), t.hostNodes.size() * sizeof(rbh_packed_node),
                                  hipMemcpyHostToDevice, c->stream));
        RBH_HIP(c, light_tree_expand(c));
    }
    c->accel.stale = true;
    return lbvh_refit(c);
}

int rbh_update_materials(rbh_ctx c, uint32_t first, uint32_t count, const rbh_material* m) {
    if (!c || (count && !m)) return RBH_ERR_INVALID;
    if ((uint64_t)first + count > c->nmaterials)
        return fail(c, RBH_ERR_RANGE, "materials %u..%u of %u", first, first + count, c->nmaterials);
    for (uint32_t i = 0; i < count; ++i) {
        const rbh_material& x = m[i];
        if (x.bxdf > 2) return fail(c, RBH_ERR_INVALID, "material %u: bxdf index %u", first + i, x.bxdf);
        for (int k = 0; k < 7; ++k)
            if (((x.textures >> k) & 1u) && (x.tex[k] < 0 || (uint32_t)x.tex[k] >= c->ntextures))
                return fail(c, RBH_ERR_INVALID, "material %u: texture slot %d references %d of %u", first + i, k,
                            x.tex[k], c->ntextures);
        if (x.bxdf == 0 && (x.iorSpectrum >= c->nspectra || x.acSpectrum >= c->nspectra))
            return fail(c, RBH_ERR_INVALID, "material %u: conductor spectra out of range", first + i);
        if (x.bxdf == 1 && x.iorSpectrum >= c->nspectra)
            return fail(c, RBH_ERR_INVALID, "material %u: dielectric ior spectrum out of range", first + i);
        if ((x.flags & 4u) && x.emissionSpectrum >= c->nspectra)
            return fail(c, RBH_ERR_INVALID, "material %u: emission spectrum out of range", first + i);
    }
    if (!count) return RBH_OK;
    // the vertex-attribute layout must stay inside the attribute buffer
    const uint64_t attribBytes = c->geom.attribs.bytes;
    for (uint32_t i = 0; i < count && c->geom.nverts; ++i) {
        const rbh_material& x = m[i];
        struct { uint32_t bit, off, size; } at[] = {{1u, x.normalOffset, 8u}, {2u, x.frameOffset, 16u},
                                                    {8u, x.texCoord1Offset, 8u}, {16u, x.texCoord2Offset, 8u}};
        for (auto& a : at)
            if ((x.attributes & a.bit) &&
                ((uint64_t)(c->geom.nverts - 1) * x.attributeStride + a.off + a.size > attribBytes ||
                 (a.off % 4) || (x.attributeStride % 4)))
                return fail(c, RBH_ERR_INVALID, "material %u: attribute (bit %u) outside the attribute buffer",
                            first + i, a.bit);
    }
    RBH_HIP(c, hipMemcpyAsync(c->materials.as<char>() + (size_t)first * sizeof(Material), m,
                              (size_t)count * sizeof(Material), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, refresh_tri_classes(c));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

}  // extern "C"

namespace {

// Checks a packed light tree before the kernels walk it: nodes in pre-order
// (the left child of inner node i is i + 1, its right child lies after it,
// so every walk ends), leaves holding 1..15 triangle lights inside the
// pointer array, every light's bitmask path from the root ending in the leaf
// that holds it (light_tree_types.h:119-206, light_tree.cpp:71-73).
int check_tree(rbh_ctx c, const PackedNode* nodes, uint32_t nnodes, const uint32_t* ptrs, const uint64_t* masks,
               uint32_t ntri, uint32_t nshape) {
    const uint32_t n = ntri + nshape;
    if (nnodes == 0 || nnodes > 2 * (uint64_t)n) return fail(c, RBH_ERR_INVALID, "%u nodes for %u lights", nnodes, n);
    for (uint32_t i = 0; i < nnodes; ++i) {
        const uint32_t core = nodes[i].core;
        if (core >> 31) {
            const uint32_t cnt = (core >> 27) & 15u, off = core & ((1u << 27) - 1u);
            if (cnt == 0 || (uint64_t)off + cnt > n)
                return fail(c, RBH_ERR_INVALID, "leaf %u holds lights [%u, %u) of %u", i, off, off + cnt, n);
        } else if (i + 1 >= nnodes || core <= i + 1 || core >= nnodes) {
            return fail(c, RBH_ERR_INVALID, "inner node %u: children %u and %u are not in pre-order below %u nodes", i,
                        i + 1, core, nnodes);
        }
    }
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t type = ptrs[k] >> 30, idx = ptrs[k] & ((1u << 30) - 1u);
        const bool okTri = type == 0u && idx < ntri, okShape = type == 3u && idx >= ntri && idx < n;
        if (!okTri && !okShape)
            return fail(c, RBH_ERR_INVALID,
                        "light pointer %u (0x%08x) is neither a triangle light below %u nor a shape light in [%u, %u)",
                        k, ptrs[k], ntri, ntri, n);
        if (seen[idx]++) return fail(c, RBH_ERR_INVALID, "light %u appears twice in the pointer array", idx);
    }
    for (uint32_t l = 0; l < n; ++l) {
        uint32_t cur = 0;
        for (int d = 0; d < 64 && !(nodes[cur].core >> 31); ++d) cur = ((masks[l] >> d) & 1ull) ? nodes[cur].core : cur + 1;
        const uint32_t core = nodes[cur].core;
        if (!(core >> 31)) return fail(c, RBH_ERR_INVALID, "light %u: its bitmask does not reach a leaf", l);
        const uint32_t cnt = (core >> 27) & 15u, off = core & ((1u << 27) - 1u);
        bool found = false;
        for (uint32_t k = off; k < off + cnt; ++k) found = found || (ptrs[k] & ((1u << 30) - 1u)) == l;
        if (!found) return fail(c, RBH_ERR_INVALID, "light %u: its bitmask ends in leaf %u, which does not hold it", l, cur);
    }
    return RBH_OK;
}

// Uploads a checked tree with its triangle records and expands the nodes.
// n triangle records; pointers and masks of the n triangle and the context's
// shape lights.
int upload_tree(rbh_ctx c, const PackedNode* nodes, uint32_t nnodes, const uint32_t* ptrs, const uint64_t* masks,
                const rbh_emissive_tri* tris, uint32_t n) {
    LightTreeDev& t = c->tree;
    const uint32_t m = nnodes ? c->shapes.n : 0u, total = n + m;
    t.hostNodes.resize(nnodes);
    if (nnodes) std::memcpy(t.hostNodes.data(), nodes, (size_t)nnodes * sizeof(PackedNode));
    t.hostPointers.assign(ptrs, ptrs + total);
    t.hostMasks.assign(masks, masks + total);
    t.nnodes = nnodes;
    t.nlights = n;
    t.nshapes = m;
    t.stale = false;
    if (total == 0) return RBH_OK;
    RBH_ALLOC(c, t.nodes, (size_t)nnodes * sizeof(PackedNode));
    RBH_ALLOC(c, t.lightPointers, (size_t)total * 4);
    RBH_ALLOC(c, t.bitmasks, (size_t)total * 8);
    RBH_HIP(c, hipMemcpyAsync(t.nodes.p, nodes, (size_t)nnodes * sizeof(PackedNode), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(t.lightPointers.p, ptrs, (size_t)total * 4, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(t.bitmasks.p, masks, (size_t)total * 8, hipMemcpyHostToDevice, c->stream));
    if (n) {
        RBH_ALLOC(c, t.tris, (size_t)n * sizeof(EmissiveTri));
        RBH_HIP(c, hipMemcpyAsync(t.tris.p, tris, (size_t)n * sizeof(EmissiveTri), hipMemcpyHostToDevice, c->stream));
    }
    RBH_ALLOC(c, t.nodesX, (size_t)nnodes * 48);
    RBH_HIP(c, light_tree_expand(c));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

constexpr uint32_t kMaxShapeLights = 4096;

// Shape-light records and flux; c may be null (host-only builds: no message,
// no material check when nmaterials is ~0).
int check_shapes(rbh_ctx c, const rbh_shape_light* s, const float* flux, uint32_t m, uint32_t nmaterials) {
    auto bad = [&](const char* what, uint32_t i) {
        return c ? fail(c, RBH_ERR_INVALID, "shape light %u: %s", i, what) : RBH_ERR_INVALID;
    };
    for (uint32_t i = 0; i < m; ++i) {
        const rbh_shape_light& x = s[i];
        if (x.type > 1u) return bad("type must be 0 (disk) or 1 (sphere)", i);
        if (!(x.radius > 0.0f) || !std::isfinite(x.radius)) return bad("radius must be positive and finite", i);
        if (!(flux[i] > 0.0f) || !std::isfinite(flux[i])) return bad("flux must be positive and finite", i);
        if (!std::isfinite(x.powerScale)) return bad("power scale must be finite", i);
        double q2 = 0.0;
        for (int k = 0; k < 4; ++k) {
            if (!std::isfinite(x.frame[k])) return bad("frame must be finite", i);
            q2 += (double)x.frame[k] * x.frame[k];
        }
        if (std::fabs(std::sqrt(q2) - 1.0) > 1e-3) return bad("frame must be a unit quaternion", i);
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(x.position[k])) return bad("position must be finite", i);
        if (nmaterials != ~0u && x.materialIndex >= nmaterials) return bad("material index out of range", i);
    }
    return RBH_OK;
}

int check_light_records(rbh_ctx c, const rbh_emissive_tri* tris, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i)
        if (tris[i].materialIndex >= c->nmaterials)
            return fail(c, RBH_ERR_INVALID, "light %u: material %u of %u", i, tris[i].materialIndex, c->nmaterials);
    return RBH_OK;
}

}  // namespace

extern "C" {

int rbh_build_light_tree_ex(rbh_ctx c, const rbh_emissive_tri* tris, const float* flux, uint32_t n, uint32_t flags) {
    if (!c) return RBH_ERR_INVALID;
    if (flags & ~(uint32_t)RBH_LIGHT_TREE_FP64) return fail(c, RBH_ERR_INVALID, "unknown light-tree flags 0x%x", flags);
    if (n && (!tris || !flux)) return fail(c, RBH_ERR_INVALID, "light arrays are null");
    for (uint32_t i = 0; i < n; ++i)
        if (!(flux[i] > 0.0f) || !std::isfinite(flux[i])) return fail(c, RBH_ERR_INVALID, "light %u: flux must be positive and finite", i);
    int r = check_light_records(c, tris, n);
    if (r) return r;
    const EmissiveTri* et = reinterpret_cast<const EmissiveTri*>(tris);
    std::vector<PackedNode> nodes;
    std::vector<uint32_t> ptrs;
    std::vector<uint64_t> masks;
    const ShapeLights& sh = c->shapes;
    r = host::build_light_tree(et, flux, n, reinterpret_cast<const ShapeLight*>(sh.host.data()), sh.flux.data(), sh.n,
                               nodes, ptrs, masks, (flags & RBH_LIGHT_TREE_FP64) != 0);
    if (r == -1) return fail(c, RBH_ERR_RANGE, "light count %u exceeds the tree's 27-bit leaf offsets", n + sh.n);
    if (r < 0) return fail(c, RBH_ERR_INVALID, "light tree produced an invalid bitmask");
    return upload_tree(c, nodes.data(), (uint32_t)nodes.size(), ptrs.data(), masks.data(), tris, n);
}

int rbh_set_light_tree(rbh_ctx c, const rbh_packed_node* nodes, uint32_t nnodes, const uint32_t* lightPointers,
                       const uint64_t* bitmasks, const rbh_emissive_tri* tris, uint32_t n) {
    if (!c) return RBH_ERR_INVALID;
    const uint32_t m = c->shapes.n;
    if (n == 0 && m == 0) return upload_tree(c, nullptr, 0, nullptr, nullptr, nullptr, 0);
    if (!nodes || !lightPointers || !bitmasks || (n && !tris))
        return fail(c, RBH_ERR_INVALID, "light-tree arrays are null");
    if (n + (uint64_t)m > (1u << 27)) return fail(c, RBH_ERR_RANGE, "light count %u exceeds the tree's 27-bit leaf offsets", n + m);
    const PackedNode* pn = reinterpret_cast<const PackedNode*>(nodes);
    int r = check_light_records(c, tris, n);
    if (r) return r;
    r = check_tree(c, pn, nnodes, lightPointers, bitmasks, n, m);
    if (r) return r;
    return upload_tree(c, pn, nnodes, lightPointers, bitmasks, tris, n);
}

int rbh_light_tree_build_host(const rbh_emissive_tri* tris, const float* flux, uint32_t n, rbh_packed_node* nodes,
                              uint32_t* lightPointers, uint64_t* bitmasks) {
    return rbh_light_tree_build_host_ex(tris, flux, n, nodes, lightPointers, bitmasks, 0u);
}

int rbh_light_tree_build_host_ex(const rbh_emissive_tri* tris, const float* flux, uint32_t n, rbh_packed_node* nodes,
                                 uint32_t* lightPointers, uint64_t* bitmasks, uint32_t flags) {
    if (n && (!tris || !flux || !nodes || !lightPointers || !bitmasks)) return RBH_ERR_INVALID;
    if (flags & ~(uint32_t)RBH_LIGHT_TREE_FP64) return RBH_ERR_INVALID;
    std::vector<PackedNode> nv;
    std::vector<uint32_t> lp;
    std::vector<uint64_t> bm;
    const int r = host::build_light_tree(reinterpret_cast<const EmissiveTri*>(tris), flux, n, nullptr, nullptr, 0u, nv,
                                         lp, bm, (flags & RBH_LIGHT_TREE_FP64) != 0);
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
    const size_t total = t.hostPointers.size();
    if (nnodes) *nnodes = t.nnodes;
    if (nodes && t.nnodes) std::memcpy(nodes, t.hostNodes.data(), (size_t)t.nnodes * sizeof(PackedNode));
    if (lightPointers && total) std::memcpy(lightPointers, t.hostPointers.data(), total * 4);
    if (bitmasks && total) std::memcpy(bitmasks, t.hostMasks.data(), total * 8);
    return RBH_OK;
}

int rbh_light_tree_build_host_shapes(const rbh_emissive_tri* tris, const float* flux, uint32_t n,
                                     const rbh_shape_light* shapes, const float* shapeFlux, uint32_t m,
                                     rbh_packed_node* nodes, uint32_t* lightPointers, uint64_t* bitmasks,
                                     uint32_t flags) {
    if ((n && (!tris || !flux)) || (m && (!shapes || !shapeFlux)) || ((n || m) && (!nodes || !lightPointers || !bitmasks)))
        return RBH_ERR_INVALID;
    if (flags & ~(uint32_t)RBH_LIGHT_TREE_FP64) return RBH_ERR_INVALID;
    if (m > kMaxShapeLights) return RBH_ERR_RANGE;
    if (check_shapes(nullptr, shapes, shapeFlux, m, ~0u) != RBH_OK) return RBH_ERR_INVALID;
    std::vector<PackedNode> nv;
    std::vector<uint32_t> lp;
    std::vector<uint64_t> bm;
    const int r = host::build_light_tree(reinterpret_cast<const EmissiveTri*>(tris), flux, n,
                                         reinterpret_cast<const ShapeLight*>(shapes), shapeFlux, m, nv, lp, bm,
                                         (flags & RBH_LIGHT_TREE_FP64) != 0);
    if (r == -1) return RBH_ERR_RANGE;
    if (r < 0) return RBH_ERR_INVALID;
    if (r) {
        std::memcpy(nodes, nv.data(), nv.size() * sizeof(PackedNode));
        std::memcpy(lightPointers, lp.data(), lp.size() * 4);
        std::memcpy(bitmasks, bm.data(), bm.size() * 8);
    }
    return r;
}

int rbh_upload_shape_lights(rbh_ctx c, const rbh_shape_light* lights, const float* flux, uint32_t m) {
    if (!c) return RBH_ERR_INVALID;
    if (m && (!lights || !flux)) return fail(c, RBH_ERR_INVALID, "shape-light arrays are null");
    if (m > kMaxShapeLights) return fail(c, RBH_ERR_RANGE, "%u shape lights (at most %u)", m, kMaxShapeLights);
    const int r = check_shapes(c, lights, flux, m, c->nmaterials);
    if (r) return r;
    ShapeLights& sh = c->shapes;
    sh.host.assign(lights, lights + m);
    sh.flux.assign(flux, flux + m);
    sh.n = m;
    if (m) {
        RBH_ALLOC(c, sh.recs, (size_t)m * sizeof(rbh_shape_light));
        RBH_HIP(c, hipMemcpyAsync(sh.recs.p, lights, (size_t)m * sizeof(rbh_shape_light), hipMemcpyHostToDevice,
                                  c->stream));
        RBH_HIP(c, hipStreamSynchronize(c->stream));
    }
    // the tree holds the previous set until it is rebuilt
    c->tree.stale = c->tree.nshapes != m || m != 0;
    return RBH_OK;
}

}  // extern "C"

namespace {

// Checks that a context holds a renderable scene for these launch parameters.
int check_scene(rbh_ctx c, const rbh_launch_params* lp) {
    if (!c->accel.built) return fail(c, RBH_ERR_STATE, "rbh_build_accel has not been run on the current geometry");
    if (c->accel.stale)
        return fail(c, RBH_ERR_STATE, "vertices changed since the last build: rbh_refit_accel or rbh_build_accel first");
    if (!c->haveColor || !c->haveLuts) return fail(c, RBH_ERR_STATE, "colour tables and LUTs must be uploaded first");
    if (lp->width == 0 || lp->height == 0) return fail(c, RBH_ERR_INVALID, "empty film");
    if ((uint64_t)lp->width * lp->height > (1ull << 28)) return fail(c, RBH_ERR_RANGE, "film above 2^28 pixels");
    if (lp->maxPathDepth == 0 || lp->maxPathDepth > 1024) return fail(c, RBH_ERR_INVALID, "maxPathDepth must be in [1, 1024]");
    if (c->tree.stale || c->tree.nshapes != c->shapes.n)
        return fail(c, RBH_ERR_STATE, "shape lights changed since the light tree was built: rebuild the light tree");
    if (c->geom.maxLight >= (int64_t)c->tree.nlights)
        return fail(c, RBH_ERR_STATE, "triangle light index %lld has no light record (%u lights built)",
                    (long long)c->geom.maxLight, c->tree.nlights);
    return RBH_OK;
}

// Makes sure the accumulation buffer fits the film; a continued launch
// (sampleOffset != 0) needs an existing accumulation of the same size.
int prepare_accum(rbh_ctx c, const rbh_launch_params* lp) {
    RenderState& rs = c->rs;
    const size_t P = (size_t)lp->width * lp->height;
    const bool sizeChanged = rs.width != lp->width || rs.height != lp->height;
    if (sizeChanged && lp->sampleOffset != 0)
        return fail(c, RBH_ERR_STATE, "sampleOffset != 0 but no accumulation of this size exists");
    RBH_ALLOC(c, rs.accum, P * sizeof(double) * 4);
    rs.width = lp->width;
    rs.height = lp->height;
    return RBH_OK;
}

// Scene-level part of a pass view (the pass fields sample, S, slot0, first
// are set per pass).
void fill_view(rbh_ctx c, const rbh_launch_params* lp, SceneView& v) {
    const bool treeOn = c->tree.nnodes > 0;
    v = SceneView{};
    v.P = lp->width * lp->height;
    v.W = lp->width;
    v.H = lp->height;
    v.maxDepth = std::max(1u, lp->maxPathDepth);
    v.maxRefl = std::min(v.maxDepth, lp->maxReflectedPathDepth);
    v.maxTrans = std::min(v.maxDepth, lp->maxTransmittedPathDepth);
    v.invW = 1.0f / (float)lp->width;
    v.invH = 1.0f / (float)lp->height;
    v.psW = lp->perspectiveScaleW;
    v.psH = lp->perspectiveScaleH;
    v.camU = dev::V3{lp->camU[0], lp->camU[1], lp->camU[2]};
    v.camV = dev::V3{lp->camV[0], lp->camV[1], lp->camV[2]};
    v.camW = dev::V3{lp->camW[0], lp->camW[1], lp->camW[2]};
    v.dof = lp->depthOfField;
    v.aperture = lp->aperture;
    v.focus = lp->focusDistance;
    v.nodes = c->accel.nodes.as<BvhNode>();
    v.nodes4 = c->accel.nodes4.as<Bvh4Node>();
    v.nodes4q = c->accel.nodes4q.as<Bvh4QNode>();
    v.nnodes4 = c->accel.nnodes4;
    v.tris = c->accel.tris.as<TriRecord>();
    v.nnodes = c->accel.nnodes;
    v.verts = c->geom.verts.as<float>();
    v.indices = c->geom.indices.as<uint32_t>();
    v.attribs = c->geom.attribs.as<uint8_t>();
    v.triMaterial = c->geom.triMaterial.as<uint32_t>();
    v.triLight = c->geom.triLight.as<uint32_t>();
    v.materials = c->materials.as<Material>();
    v.texDescs = c->textures.as<TextureDesc>();
    v.texels = c->texels.as<uint8_t>();
    v.srgbLut = c->srgbLut.as<float>();
    v.spectra = c->spectra.as<float>();
    v.colors = c->colorTables.as<float>();
    v.luts = c->luts.as<float>();
    v.treeNodes = treeOn ? c->tree.nodesX.as<dev::TreeNodeX>() : nullptr;
    v.treePointers = treeOn ? c->tree.lightPointers.as<uint32_t>() : nullptr;
    v.treeMasks = treeOn ? c->tree.bitmasks.as<uint64_t>() : nullptr;
    v.lights = treeOn && c->tree.nlights ? c->tree.tris.as<EmissiveTri>() : nullptr;
    v.shapes = c->shapes.n ? c->shapes.recs.as<ShapeLight>() : nullptr;
    v.nshapes = c->shapes.n;
    v.shapeBase = c->tree.nlights;
    v.envOn = c->env.enabled ? 1u : 0u;
    if (c->env.enabled) {
        v.envW = (uint32_t)c->env.width;
        v.envH = (uint32_t)c->env.height;
        v.envC = (uint32_t)c->env.channels;
        v.envTex = c->env.texels.as<float>();
        v.envCond = c->env.cond.as<uint16_t>();
        v.envMarg = c->env.marg.as<uint16_t>();
        v.envIntegral = (float)c->env.integral;
    }
    v.accum = c->rs.accum.as<double4>();
}

// Path buffers, counters, spill columns and launch shape of the owner context
// for NP slots.
int prepare_pass_state(rbh_ctx c, size_t NP, Frame1& f, LaunchShape& ls) {
    RenderState& rs = c->rs;
    PathBuffers pb{};
    uint32_t* counters = nullptr;
    unsigned long long* stats = nullptr;
    const size_t bytes = layout_paths(nullptr, NP, &pb, &counters, &stats);
    RBH_ALLOC(c, rs.paths, bytes);
    layout_paths(rs.paths.as<char>(), NP, &pb, &counters, &stats);
    f.p = pb;
    f.counters = counters;
    f.stats = stats;
    const int cus = cu_count(c->device);
    // persistent grids, never more lanes than the pass has path slots
    ls.traceGrid = (int)std::max<size_t>(1, std::min<size_t>((size_t)cus * rs.traceBlocksPerCU, (NP + kTraceBlock - 1) / kTraceBlock));
    ls.shadeGrid = (int)std::max<size_t>(1, std::min<size_t>((size_t)cus * rs.shadeBlocksPerCU, (NP + kShadeBlock - 1) / kShadeBlock));
    ls.shadeWaves = rs.shadeWaves;
    ls.traceFetch = rs.traceFetch;
    ls.traceWaves = rs.traceWaves;
    ls.quantNodes = rs.quantNodes;
    ls.sortHits = rs.sortHits;
    const size_t lanes = (size_t)ls.traceGrid * kTraceBlock;
    RBH_ALLOC(c, rs.spill, trace_spill_words(lanes) * sizeof(int));
    f.spill = rs.spill.as<int>();
    f.spillStride = (uint32_t)lanes;
    return RBH_OK;
}

// One pass: launched directly, or (useGraph, no per-kernel timing) captured
// into a graph whose executable is updated in place for each pass (same
// topology, new kernel arguments) and replayed. The pass has no memset nodes:
// its counters are cleared by kernels.
hipError_t run_pass(rbh_ctx c, const Frame1& f, uint32_t depth, const LaunchShape& ls, KernelTimer* kt) {
    RenderState& rs = c->rs;
    if (!rs.useGraph || kt) return launch_pass(c, f, depth, ls, kt);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t le = launch_pass(c, f, depth, ls, nullptr);
    e = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess) { if (g) (void)hipGraphDestroy(g); return le; }
    if (e != hipSuccess) return e;
    bool fresh = rs.graphExec == nullptr;
    if (!fresh) {
        hipGraphNode_t errNode = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(rs.graphExec, g, &errNode, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(rs.graphExec);
            rs.graphExec = nullptr;
            fresh = true;
        }
    }
    if (fresh) {
        e = hipGraphInstantiate(&rs.graphExec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { rs.graphExec = nullptr; (void)hipGraphDestroy(g); return e; }
    }
    (void)hipGraphDestroy(g);
    return hipGraphLaunch(rs.graphExec, c->stream);
}

// Path slots of a pass: the context's budget, or (budget 0) all the samples
// of the launch if a quarter of the free device memory holds them.
size_t slot_budget(rbh_ctx c, size_t want, size_t bytesPerSlot) {
    if (c->rs.slotBudget) return c->rs.slotBudget;
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) freeB = (size_t)16 << 30;
    // memory already held by this context's path buffers is reusable
    freeB += c->rs.paths.bytes;
    const size_t cap = std::max<size_t>(size_t(1) << 20, freeB / 4 / bytesPerSlot);
    return std::min<size_t>(std::min(want, cap), size_t(1) << 30);
}

int finish_stats(rbh_ctx c, const Frame1& f, KernelTimer* kt, uint64_t samples, uint32_t launches) {
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    RBH_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    unsigned long long st[STAT_COUNT] = {};
    RBH_HIP(c, hipMemcpy(st, f.stats, sizeof(st), hipMemcpyDeviceToHost));
    c->stats.renderMs = ms;
    c->stats.closestRays = st[STAT_CLOSEST];
    c->stats.shadowRays = st[STAT_SHADOW];
    c->stats.pathVertices = st[STAT_VERTICES];
    c->stats.nonFiniteSamples = st[STAT_NONFINITE];
    c->stats.samples = samples;
    c->stats.traceLaunches = launches;
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

}  // namespace

extern "C" {

int rbh_render(rbh_ctx c, const rbh_launch_params* lp) {
    if (!c || !lp) return RBH_ERR_INVALID;
    int rc = check_scene(c, lp);
    if (rc) return rc;
    const uint32_t known = RBH_RENDER_GUIDED | RBH_RENDER_COLLECT | RBH_RENDER_GUIDE_FP8 | RBH_RENDER_TIMING |
                           RBH_RENDER_NO_DROP;
    if (lp->flags & ~known)
        return fail(c, RBH_ERR_INVALID, "render flags 0x%x are not supported by this entry point", lp->flags);
    const bool guided = (lp->flags & RBH_RENDER_GUIDED) != 0;
    const bool collect = (lp->flags & RBH_RENDER_COLLECT) != 0;
    if (collect && !guided) return fail(c, RBH_ERR_INVALID, "RBH_RENDER_COLLECT needs RBH_RENDER_GUIDED");
    if ((lp->flags & RBH_RENDER_GUIDE_FP8) && !guided)
        return fail(c, RBH_ERR_INVALID, "RBH_RENDER_GUIDE_FP8 needs RBH_RENDER_GUIDED");
    GuideState* gst = static_cast<GuideState*>(c->guide);
    if (guided && !(gst && gst->configured)) return fail(c, RBH_ERR_STATE, "guided render before rbh_guide_configure");
    RenderState& rs = c->rs;
    const size_t P = (size_t)lp->width * lp->height;
    // samples per pass: as many as fit the path-slot budget (at least one)
    const uint32_t spp = lp->sampleCount;
    const size_t budget = slot_budget(c, P * std::max(1u, spp), guided ? 480 : 200);
    const uint32_t S = (uint32_t)std::max<size_t>(1, std::min<size_t>(spp ? spp : 1, budget / P));
    const size_t NP = P * S;
    rc = prepare_accum(c, lp);
    if (rc) return rc;
    Frame1 f{};
    fill_view(c, lp, f.sc);
    LaunchShape ls;
    rc = prepare_pass_state(c, NP, f, ls);
    if (rc) return rc;
    f.maxP = (uint32_t)P;
    f.noDrop = (guided || (lp->flags & RBH_RENDER_NO_DROP)) ? 1u : 0u;
    f.camOrigin0 = lp->depthOfField ? 0u : 1u;
    if (guided) {
        RBH_ALLOC(c, rs.gfeat, NP * 16 * sizeof(float));
        RBH_ALLOC(c, rs.gmix, NP * (guide::kOut) * sizeof(float));   // lobes + selection per hit
        f.guided = 1u;
        f.gwf = static_cast<const __bf16*>(gst->wf);
        f.gbias = static_cast<const float*>(gst->bias);
        if (lp->flags & RBH_RENDER_GUIDE_FP8) {
            f.gwq = static_cast<const uint8_t*>(gst->fp8w);
            f.gws = static_cast<const float*>(gst->fp8s);
        }
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
    KernelTimer* kt = (lp->flags & RBH_RENDER_TIMING) ? &rs.timer : nullptr;
    if (kt) kt->reset();
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, c->stream, f.stats);
    uint32_t passes = 0;
    for (uint32_t done = 0; done < spp;) {
        f.sc.sample = lp->sampleOffset + done;
        f.sc.S = std::min(S, spp - done);
        f.sc.slot0 = 0;
        f.sc.first = (lp->sampleOffset == 0 && done == 0) ? 1u : 0u;
        f.NP = (uint32_t)(P * f.sc.S);
        RBH_HIP(c, run_pass(c, f, f.sc.maxDepth, ls, kt));
        done += f.sc.S;
        ++passes;
    }
    return finish_stats(c, f, kt, (uint64_t)P * spp, passes * f.sc.maxDepth);
}

// Many scenes (one per context, all on the owner's device) in one wavefront:
// each pass holds S samples of every unfinished scene's pixels, slots in
// scene order. Per-scene completion times are the end of the pass in which
// the scene's last sample finished.
int rbh_render_batch(rbh_ctx owner, const rbh_ctx* scenes, const rbh_launch_params* lps, uint32_t n,
                     double* completionMs) {
    if (!owner || !scenes || !lps || n == 0) return RBH_ERR_INVALID;
    if (n > 4096) return fail(owner, RBH_ERR_RANGE, "at most 4096 scenes per batch");
    uint64_t totalP = 0, totalSamples = 0;
    uint32_t depth = 1;
    for (uint32_t s = 0; s < n; ++s) {
        rbh_ctx c = scenes[s];
        if (!c) return fail(owner, RBH_ERR_INVALID, "scene %u: null context", s);
        if (c->device != owner->device) return fail(owner, RBH_ERR_INVALID, "scene %u is on another device", s);
        if (lps[s].flags & ~(uint32_t)(RBH_RENDER_TIMING | RBH_RENDER_NO_DROP))
            return fail(owner, RBH_ERR_INVALID, "scene %u: guided flags are not supported in batches", s);
        const int rc = check_scene(c, &lps[s]);
        if (rc) return fail(owner, rc, "scene %u: %s", s, c->err.c_str());
        const int ra = prepare_accum(c, &lps[s]);
        if (ra) return fail(owner, ra, "scene %u: %s", s, c->err.c_str());
        totalP += (uint64_t)lps[s].width * lps[s].height;
        totalSamples += (uint64_t)lps[s].width * lps[s].height * lps[s].sampleCount;
        depth = std::max(depth, std::max(1u, lps[s].maxPathDepth));
    }
    RenderState& rs = owner->rs;
    // Pass schedule: each pass packs the scenes' remaining samples shortest
    // remaining work first (pixels x samples left), whole samples per scene,
    // up to the pass's slot count: small scenes finish in the first passes
    // and large ones stream through the rest (shortest-job-first completion
    // times at full passes). A pass holds at least the largest film and at
    // most the slot budget; without an explicit budget it is sized for about
    // kBatchPasses passes.
    constexpr uint64_t kBatchPasses = 8;
    uint64_t maxFilm = 0;
    for (uint32_t s = 0; s < n; ++s) maxFilm = std::max<uint64_t>(maxFilm, (uint64_t)lps[s].width * lps[s].height);
    const size_t budget = slot_budget(owner, (size_t)totalSamples, 200);
    uint64_t passSlots = rs.slotBudget ? (uint64_t)budget
                                       : std::min<uint64_t>(budget, (totalSamples + kBatchPasses - 1) / kBatchPasses);
    passSlots = std::max<uint64_t>(passSlots, maxFilm);
    if (passSlots > (1ull << 30)) return fail(owner, RBH_ERR_RANGE, "a film of %llu pixels exceeds a pass",
                                              (unsigned long long)maxFilm);
    std::vector<SceneView> base(n);
    for (uint32_t s = 0; s < n; ++s) {
        fill_view(scenes[s], &lps[s], base[s]);
        // a scene may need fewer bounces than the pass runs; its paths end at its own caps
    }
    // the schedule, planned on the host: views per pass (sample, S, slot0, first)
    std::vector<std::vector<SceneView>> tables;
    std::vector<int> finishedAt(n, -1);
    {
        std::vector<uint32_t> done(n, 0u);
        uint32_t left = n;
        for (uint32_t s = 0; s < n; ++s)
            if (lps[s].sampleCount == 0) { --left; finishedAt[s] = -1; }
        std::vector<uint32_t> order(n);
        while (left > 0) {
            for (uint32_t s = 0; s < n; ++s) order[s] = s;
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                const uint64_t wa = (uint64_t)base[a].P * (lps[a].sampleCount - done[a]);
                const uint64_t wb = (uint64_t)base[b].P * (lps[b].sampleCount - done[b]);
                return wa < wb;
            });
            std::vector<SceneView> t;
            uint64_t room = passSlots;
            uint32_t slot = 0;
            for (uint32_t s : order) {
                if (done[s] >= lps[s].sampleCount) continue;
                const uint64_t fit = room / base[s].P;
                if (fit == 0) continue;
                SceneView v = base[s];
                v.sample = lps[s].sampleOffset + done[s];
                v.S = (uint32_t)std::min<uint64_t>(fit, lps[s].sampleCount - done[s]);
                v.slot0 = slot;
                v.first = (lps[s].sampleOffset == 0 && done[s] == 0) ? 1u : 0u;
                slot += v.P * v.S;
                room -= (uint64_t)v.P * v.S;
                done[s] += v.S;
                if (done[s] >= lps[s].sampleCount) {
                    finishedAt[s] = (int)tables.size();
                    --left;
                }
                t.push_back(v);
            }
            // slots in ascending order, as the kernels' scene lookup needs
            tables.push_back(std::move(t));
        }
    }
    Frame1 f{};
    LaunchShape ls;
    int rc = prepare_pass_state(owner, (size_t)passSlots, f, ls);
    if (rc) return rc;
    f.noDrop = 0u;
    f.camOrigin0 = 1u;
    for (uint32_t s = 0; s < n; ++s) {
        if (lps[s].flags & RBH_RENDER_NO_DROP) f.noDrop = 1u;
        if (lps[s].depthOfField) f.camOrigin0 = 0u;
    }
    RBH_ALLOC(owner, rs.sceneTable, std::max<size_t>(1, tables.size()) * n * sizeof(SceneView));
    KernelTimer* kt = (lps[0].flags & RBH_RENDER_TIMING) ? &rs.timer : nullptr;
    if (kt) kt->reset();
    std::vector<hipEvent_t> passEnd;
    RBH_HIP(owner, hipEventRecord(owner->ev0, owner->stream));
    hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, owner->stream, f.stats);
    uint32_t passes = 0;
    int err = RBH_OK;
    for (const std::vector<SceneView>& tb : tables) {
        uint32_t slot = 0, maxP = 0;
        for (const SceneView& v : tb) {
            slot = v.slot0 + v.P * v.S;
            maxP = std::max(maxP, v.P);
        }
        // each pass reads its own table slice, so uploads never overwrite a table in use
        SceneView* dtab = rs.sceneTable.as<SceneView>() + (size_t)passes * n;
        if (hipMemcpyAsync(dtab, tb.data(), tb.size() * sizeof(SceneView), hipMemcpyHostToDevice, owner->stream) !=
            hipSuccess) { err = fail(owner, RBH_ERR_HIP, "scene table upload failed"); break; }
        f.scenes = dtab;
        f.nscenes = (uint32_t)tb.size();
        f.NP = slot;
        f.maxP = maxP;
        const hipError_t e = run_pass(owner, f, depth, ls, kt);
        if (e != hipSuccess) { err = fail(owner, RBH_ERR_HIP, "launch failed: %s", hipGetErrorString(e)); break; }
        hipEvent_t ev = nullptr;
        if (completionMs && hipEventCreate(&ev) == hipSuccess) {
            (void)hipEventRecord(ev, owner->stream);
            passEnd.push_back(ev);
        }
        ++passes;
    }
    if (err == RBH_OK) err = finish_stats(owner, f, kt, totalSamples, passes * depth);
    if (completionMs) {
        for (uint32_t s = 0; s < n; ++s) {
            completionMs[s] = -1.0;
            const int pi = finishedAt[s];
            if (err == RBH_OK && pi >= 0 && pi < (int)passEnd.size()) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, owner->ev0, passEnd[pi]) == hipSuccess) completionMs[s] = ms;
            }
        }
    }
    for (hipEvent_t ev : passEnd) (void)hipEventDestroy(ev);
    return err;
}

int rbh_autotune(rbh_ctx c, const rbh_launch_params* lp, uint32_t trialSpp, rbh_tuning* best) {
    if (!c || !lp || !best || trialSpp == 0) return RBH_ERR_INVALID;
    const int rc0 = check_scene(c, lp);
    if (rc0) return rc0;
    RenderState& rs = c->rs;
    const size_t P = (size_t)lp->width * lp->height;
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) freeB = (size_t)8 << 30;
    const size_t keep = rs.slotBudget;
    const int keepT = rs.traceBlocksPerCU, keepS = rs.shadeBlocksPerCU;
    rbh_launch_params t = *lp;
    t.sampleOffset = 0;
    t.sampleCount = trialSpp;
    t.flags &= ~(uint32_t)(RBH_RENDER_COLLECT | RBH_RENDER_TIMING);
    // the slot budget stays automatic (all samples of a launch in one pass
    // when memory allows); the grids of the persistent kernels are searched
    const int gridList[] = {8, 16, 32, 64, 128, 256};
    double bestRate = -1.0;
    rbh_tuning b{};
    (void)freeB;
    const int keepF = rs.traceFetch, keepW = rs.shadeWaves, keepTW = rs.traceWaves, keepQ = rs.quantNodes;
    const int keepH = rs.sortHits;
    rs.quantNodes = 0;
    b.quantNodes = 0;
    rs.sortHits = 0;
    b.sortHits = 0;
    rs.traceFetch = 0;
    rs.shadeWaves = 1;
    rs.traceWaves = 1;
    b.traceFetch = 0;
    b.traceWaves = 1;
    for (int tg : gridList) {
        for (int sg : gridList) {
            rs.slotBudget = keep;
            rs.traceBlocksPerCU = tg;
            rs.shadeBlocksPerCU = sg;
            const int rc = rbh_render(c, &t);
            if (rc) {
                rs.traceBlocksPerCU = keepT;
                rs.shadeBlocksPerCU = keepS;
                rs.traceFetch = keepF;
                rs.shadeWaves = keepW;
                rs.traceWaves = keepTW;
                rs.quantNodes = keepQ;
                rs.sortHits = keepH;
                return rc;
            }
            const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
            if (rate > bestRate) {
                bestRate = rate;
                b.pathSlots = keep;
                b.traceBlocksPerCU = tg;
                b.shadeBlocksPerCU = sg;
                b.samplesPerSecond = rate;
            }
        }
    }
    if (bestRate >= 0.0) {
        // shade-kernel register budgets on the best grid
        for (int w : {3, 4}) {
            rs.traceBlocksPerCU = b.traceBlocksPerCU;
            rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
            rs.shadeWaves = w;
            const int rc = rbh_render(c, &t);
            rs.shadeWaves = 1;
            if (rc) return rc;
            const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
            if (rate > bestRate) {
                bestRate = rate;
                b.samplesPerSecond = rate;
                b.shadeWaves = w;
            }
        }
    }
    if (bestRate >= 0.0) {
        // trace/shadow register budgets on the best grid
        for (int w : {7, 8}) {
            rs.traceBlocksPerCU = b.traceBlocksPerCU;
            rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
            rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
            rs.traceWaves = w;
            const int rc = rbh_render(c, &t);
            rs.traceWaves = 1;
            if (rc) return rc;
            const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
            if (rate > bestRate) {
                bestRate = rate;
                b.samplesPerSecond = rate;
                b.traceWaves = w;
            }
        }
    }
    if (bestRate >= 0.0) {
        // the 64-byte quantised nodes
        rs.traceBlocksPerCU = b.traceBlocksPerCU;
        rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
        rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
        rs.traceWaves = b.traceWaves;
        rs.quantNodes = 1;
        const int rc = rbh_render(c, &t);
        rs.quantNodes = 0;
        if (rc) return rc;
        const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
        if (rate > bestRate) {
            bestRate = rate;
            b.samplesPerSecond = rate;
            b.quantNodes = 1;
        }
    }
    if (bestRate >= 0.0) {
        // material-sorted shading
        rs.traceBlocksPerCU = b.traceBlocksPerCU;
        rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
        rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
        rs.traceWaves = b.traceWaves;
        rs.quantNodes = b.quantNodes;
        rs.sortHits = 1;
        const int rc = rbh_render(c, &t);
        rs.sortHits = 0;
        rs.quantNodes = 0;
        if (rc) return rc;
        const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
        if (rate > bestRate) {
            bestRate = rate;
            b.samplesPerSecond = rate;
            b.sortHits = 1;
        }
    }
    if (bestRate >= 0.0) {
        // per-lane work fetch instead of wave-synchronous ray batches
        rs.traceBlocksPerCU = b.traceBlocksPerCU;
        rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
        rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
        rs.traceWaves = b.traceWaves;
        rs.sortHits = b.sortHits;
        rs.traceFetch = 1;
        const int rc = rbh_render(c, &t);
        rs.traceFetch = 0;
        rs.sortHits = 0;
        if (rc) return rc;
        const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
        if (rate > bestRate) {
            bestRate = rate;
            b.samplesPerSecond = rate;
            b.traceFetch = 1;
            b.quantNodes = 0;    // the work-fetch kernels read the 128-byte nodes
        }
    }
    if (bestRate >= 0.0) {
        // the best shape once more as a replayed graph
        rs.traceFetch = b.traceFetch;
        rs.traceWaves = b.traceWaves;
        rs.quantNodes = b.quantNodes;
        rs.sortHits = b.sortHits;
        rs.slotBudget = b.pathSlots;
        rs.traceBlocksPerCU = b.traceBlocksPerCU;
        rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
        rs.useGraph = true;
        rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
        const int rc = rbh_render(c, &t);
        rs.useGraph = false;
        if (rc) return rc;
        const double rate = (double)P * trialSpp / (c->stats.renderMs * 1e-3);
        if (rate > bestRate) {
            bestRate = rate;
            b.samplesPerSecond = rate;
            b.useGraph = 1;
        }
    }
    if (bestRate < 0.0) {
        rs.slotBudget = keep;
        rs.traceBlocksPerCU = keepT;
        rs.shadeBlocksPerCU = keepS;
        rs.traceFetch = keepF;
        rs.shadeWaves = keepW;
        rs.traceWaves = keepTW;
        rs.quantNodes = keepQ;
        rs.sortHits = keepH;
        return fail(c, RBH_ERR_NOMEM, "no launch shape fits the free device memory");
    }
    rs.slotBudget = b.pathSlots;
    rs.traceBlocksPerCU = b.traceBlocksPerCU;
    rs.shadeBlocksPerCU = b.shadeBlocksPerCU;
    rs.useGraph = b.useGraph != 0;
    rs.shadeWaves = b.shadeWaves ? b.shadeWaves : 1;
    rs.traceFetch = b.traceFetch;
    rs.traceWaves = b.traceWaves;
    rs.quantNodes = b.quantNodes;
    rs.sortHits = b.sortHits;
    if (!b.shadeWaves) b.shadeWaves = 1;
    *best = b;
    return RBH_OK;
}

int rbh_set_tuning(rbh_ctx c, const rbh_tuning* t) {
    if (!c || !t) return RBH_ERR_INVALID;
    if (t->pathSlots > (1ull << 30) || t->traceBlocksPerCU < 1 || t->traceBlocksPerCU > 256 ||
        t->shadeBlocksPerCU < 1 || t->shadeBlocksPerCU > 256)
        return fail(c, RBH_ERR_RANGE, "tuning out of range");
    c->rs.slotBudget = (size_t)t->pathSlots;
    c->rs.traceBlocksPerCU = t->traceBlocksPerCU;
    c->rs.shadeBlocksPerCU = t->shadeBlocksPerCU;
    c->rs.useGraph = t->useGraph != 0;
    c->rs.shadeWaves = (t->shadeWaves == 3 || t->shadeWaves == 4) ? t->shadeWaves : 1;
    c->rs.traceFetch = t->traceFetch != 0;
    c->rs.traceWaves = (t->traceWaves == 7 || t->traceWaves == 8) ? t->traceWaves : 1;
    c->rs.quantNodes = t->quantNodes != 0;
    c->rs.sortHits = t->sortHits != 0;
    return RBH_OK;
}

int rbh_get_tuning(rbh_ctx c, rbh_tuning* t) {
    if (!c || !t) return RBH_ERR_INVALID;
    *t = rbh_tuning{};
    t->pathSlots = c->rs.slotBudget;
    t->traceBlocksPerCU = c->rs.traceBlocksPerCU;
    t->shadeBlocksPerCU = c->rs.shadeBlocksPerCU;
    t->useGraph = c->rs.useGraph ? 1 : 0;
    t->shadeWaves = c->rs.shadeWaves;
    t->traceFetch = c->rs.traceFetch;
    t->traceWaves = c->rs.traceWaves;
    t->quantNodes = c->rs.quantNodes;
    t->sortHits = c->rs.sortHits;
    return RBH_OK;
}

int rbh_set_path_slots(rbh_ctx c, uint64_t slots) {
    if (!c) return RBH_ERR_INVALID;
    if (slots > (1ull << 30)) return fail(c, RBH_ERR_RANGE, "path slots must be at most 2^30 (0: automatic)");
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
