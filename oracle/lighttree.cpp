// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// Light-tree builder restating src/core/light_tree.cpp:14-418 (binned SAOH,
// 32 bins x 3 axes, at most 8 lights per leaf, depth-first node layout with
// the left child at index+1) and the per-light sort data of
// src/core/rendering_interface.cpp:1116-1150.
//
// std::nth_element leaves ties in an implementation-defined order; both this
// oracle and the product builder replace it with a total order on
// (centroid[axis], light index), which is one valid nth_element outcome, so
// topologies are comparable bit for bit.
#include "rb_math.hpp"
#include "rb_oracle.h"
#include <vector>
#include <limits>

namespace orc {

struct BBox { float mn[3], mx[3]; };
static BBox bbox_default() { return {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}; }
static BBox bbox_union(const BBox& a, const BBox& b) {
    BBox r;
    for (int i = 0; i < 3; ++i) { r.mn[i] = std::min(a.mn[i], b.mn[i]); r.mx[i] = std::max(a.mx[i], b.mx[i]); }
    return r;
}
static float bbox_center(const BBox& b, int d) { return b.mn[d] + (b.mx[d] - b.mn[d]) * 0.5f; }
static bool bbox_valid(const BBox& b) { return b.mn[0] <= b.mx[0] && b.mn[1] <= b.mx[1] && b.mn[2] <= b.mx[2]; }
static float bbox_area(const BBox& b) {
    float x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
    return x * y * 2.0f + x * z * 2.0f + y * z * 2.0f;
}

struct SortData {
    BBox bounds;
    float cone[3];
    float cosCone;
    float flux;
    uint32_t light;
};

struct TreeOut {
    std::vector<PackedNode> nodes;
    std::vector<uint32_t> lightPointers;
    std::vector<uint64_t> bitmasks;
};

// src/core/light_tree_types.h:65-100
static NodeAttr attr_add(const NodeAttr& A, const NodeAttr& B) {
    NodeAttr r{};
    float wa = A.flux / (A.flux + B.flux), wb = B.flux / (A.flux + B.flux);
    float la = vmf_sharpness_to_axis_length(A.sharpness), lb = vmf_sharpness_to_axis_length(B.sharpness);
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = wa * la * A.dir[i] + wb * lb * B.dir[i];
    float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; ++i) d[i] = dl > 0.0001f ? d[i] / dl : 0.0f;
    float md[3] = {B.mean[0] - A.mean[0], B.mean[1] - A.mean[1], B.mean[2] - A.mean[2]};
    float md2 = md[0] * md[0] + md[1] * md[1] + md[2] * md[2];
    for (int i = 0; i < 3; ++i) r.mean[i] = wa * A.mean[i] + wb * B.mean[i];
    r.variance = wa * A.variance + wb * B.variance + wa * wb * md2;
    for (int i = 0; i < 3; ++i) r.dir[i] = d[i];
    r.flux = A.flux + B.flux;
    r.sharpness = vmf_axis_length_to_sharpness(dl);
    return r;
}

// src/core/light_tree_types.h:144-161
static void pack_attr(PackedNode& p, const NodeAttr& a) {
    p.spatialMean[0] = a.mean[0]; p.spatialMean[1] = a.mean[1]; p.spatialMean[2] = a.mean[2];
    p.spatialVariance = a.variance;
    p.packedAverageDirection = oct_encode_u32({a.dir[0], a.dir[1], a.dir[2]});
    p.sharpness = a.sharpness;
    p.flux = a.flux;
}

// Unit normal of a disk light: the third column of its quaternion (x, y, z, w).
static V3 shape_normal(const orc_shape_light& s) {
    const float x = s.frame[0], y = s.frame[1], z = s.frame[2], w = s.frame[3];
    return normalize(V3{2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)});
}

struct Builder {
    const float* verts;  // per light: 9 floats (camera-relative world space)
    const orc_shape_light* shapes = nullptr;   // light index ntri + i
    uint32_t ntri = 0;
    uint32_t maxPerLeaf = 8, binCount = 32;
    TreeOut* out;

    // Shape lights (no reference builder: scene.h:319-492 declares them):
    // a disk emits like a triangle, a one-sided cosine lobe (sharpness of
    // axis length 0.5) with the variance r^2/2 of a uniform disk; a sphere
    // emits in every direction (sharpness 0) with variance r^2.
    NodeAttr shape_attr(const SortData& sd) const {
        const orc_shape_light& s = shapes[sd.light - ntri];
        NodeAttr a{};
        a.mean[0] = s.position[0]; a.mean[1] = s.position[1]; a.mean[2] = s.position[2];
        if (s.type == 0u) {
            V3 n = shape_normal(s);
            a.dir[0] = n.x; a.dir[1] = n.y; a.dir[2] = n.z;
            a.variance = s.radius * s.radius / 2.0f;
            a.sharpness = vmf_axis_length_to_sharpness(0.5f);
        } else {
            a.dir[0] = 0.0f; a.dir[1] = 0.0f; a.dir[2] = 1.0f;
            a.variance = s.radius * s.radius;
            a.sharpness = 0.0f;
        }
        a.flux = sd.flux;
        return a;
    }

    // src/core/light_tree.cpp:14-43
    NodeAttr light_attr(const SortData& s) const {
        if (s.light >= ntri) return shape_attr(s);
        const float* v = verts + (size_t)s.light * 9;
        V3 v0{v[0], v[1], v[2]}, v1{v[3], v[4], v[5]}, v2{v[6], v[7], v[8]};
        V3 e1 = v1 - v0, e2 = v2 - v0;
        V3 n = normalize(cross(e1, e2));
        V3 m = (v0 + v1 + v2) / 3.0f;
        NodeAttr a{};
        a.mean[0] = m.x; a.mean[1] = m.y; a.mean[2] = m.z;
        a.variance = (dot(e1, e1) + dot(e2, e2) - dot(e1, e2)) / 18.0f;
        a.dir[0] = n.x; a.dir[1] = n.y; a.dir[2] = n.z;
        a.sharpness = vmf_axis_length_to_sharpness(0.5f);
        a.flux = s.flux;
        return a;
    }

    struct Split { uint32_t axis = UINT32_MAX, index = UINT32_MAX; bool valid() const { return axis != UINT32_MAX && index != UINT32_MAX; } };

    // src/core/light_tree.cpp:167-185
    static float merge_cone(const float* da, float ca, const float* db, float cb) {
        float r = -1.0f;
        if (ca != -1.0f && cb != -1.0f) {
            float cd = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
            float sd = std::sqrt(1.0f - cd * cd);
            float sb = std::sqrt(1.0f - cb * cb);
            float ct = cb * cd - sb * sd;
            float st = sb * cd + cb * sd;
            if (st > 0.0f) r = std::min(cb, ct);
        }
        return r;
    }
    // src/core/light_tree.cpp:187-199
    static float saoh(const BBox& b, float flux, float cosT) {
        const float pi = PI;
        float ac = bbox_valid(b) ? bbox_area(b) : 0.0f;
        float thO = cosT != -1.0f ? std::acos(std::max(std::min(cosT, 1.0f), -1.0f)) : pi;
        float thW = std::min(thO + pi / 2.0f, pi);
        float sO = std::sin(thO), cO = std::cos(thO);
        float oc = 2.0f * pi * (1.0f - cO) + pi / 2.0f * (2.0f * thW * sO - std::cos(thO - 2.0f * thW) - 2.0f * thO * sO + cO);
        return flux * ac * oc;
    }

    // A default-constructed bin has a zero (not an empty) box, as the
    // reference's value-initialised AABB::BBox (light_tree.cpp:211), so bin
    // and sweep bounds always contain the origin (the camera position).
    struct Bin {
        BBox b = BBox{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
        uint32_t n = 0;
        float flux = 0.0f;
        float cone[3] = {0, 0, 0};
        float cosCone = 1.0f;
        void add(const BBox& bb, uint32_t cnt, float f, const float* c) {
            b = bbox_union(b, bb); n += cnt; flux += f;
            cone[0] += c[0]; cone[1] += c[1]; cone[2] += c[2];
        }
    };

    // src/core/light_tree.cpp:162-394
    Split split_fn(std::vector<SortData>& sd, uint32_t begin, uint32_t end, const BBox& bounds) {
        float splitCost = std::numeric_limits<float>::max();
        Split split;
        float dims[3] = {bounds.mx[0] - bounds.mn[0], bounds.mx[1] - bounds.mn[1], bounds.mx[2] - bounds.mn[2]};
        uint32_t largest = dims[0] > dims[1] && dims[0] > dims[2] ? 0u : (dims[1] > dims[2] ? 1u : 2u);
        const uint32_t nc = binCount - 1;
        std::vector<Bin> bins(binCount);
        std::vector<float> costs(nc);
        for (uint32_t dim = 0; dim < 3; ++dim) {
            auto binId = [&](const SortData& s) {
                float bmin = bounds.mn[dim], bmax = bounds.mx[dim];
                float w = bmax - bmin;
                float scale = w > std::numeric_limits<float>::min() ? (float)binCount / w : 0.0f;
                float p = bbox_center(s.bounds, dim);
                return std::min((uint32_t)((p - bmin) * scale), binCount - 1);
            };
            for (auto& b : bins) b = Bin{};
            for (uint32_t i = begin; i < end; ++i) bins[binId(sd[i])].add(sd[i].bounds, 1, sd[i].flux, sd[i].cone);
            for (auto& b : bins) {
                float l = std::sqrt(b.cone[0] * b.cone[0] + b.cone[1] * b.cone[1] + b.cone[2] * b.cone[2]);
                if (l < std::numeric_limits<float>::min()) {
                    b.cosCone = -1.0f;
                } else {
                    b.cosCone = 1.0f;
                    b.cone[0] /= l; b.cone[1] /= l; b.cone[2] /= l;
                }
            }
            for (uint32_t i = begin; i < end; ++i) {
                Bin& b = bins[binId(sd[i])];
                b.cosCone = merge_cone(b.cone, b.cosCone, sd[i].cone, sd[i].cosCone);
            }
            Bin tot;
            for (uint32_t i = 0; i < nc; ++i) {
                tot.add(bins[i].b, bins[i].n, bins[i].flux, bins[i].cone);
                float cosT = -1.0f;
                float l = std::sqrt(tot.cone[0] * tot.cone[0] + tot.cone[1] * tot.cone[1] + tot.cone[2] * tot.cone[2]);
                if (l > std::numeric_limits<float>::min()) {
                    cosT = 1.0f;
                    float cd[3] = {tot.cone[0] / l, tot.cone[1] / l, tot.cone[2] / l};
                    for (uint32_t j = 0; j <= i; ++j) cosT = merge_cone(cd, cosT, bins[j].cone, bins[j].cosCone);
                }
                costs[i] = saoh(tot.b, tot.flux, cosT);
            }
            tot = Bin{};
            for (int i = (int)nc; i > 0; --i) {
                tot.add(bins[i].b, bins[i].n, bins[i].flux, bins[i].cone);
                float cosT = -1.0f;
                float l = std::sqrt(tot.cone[0] * tot.cone[0] + tot.cone[1] * tot.cone[1] + tot.cone[2] * tot.cone[2]);
                if (l > std::numeric_limits<float>::min()) {
                    cosT = 1.0f;
                    float cd[3] = {tot.cone[0] / l, tot.cone[1] / l, tot.cone[2] / l};
                    for (uint32_t j = (uint32_t)i; j <= nc; ++j) cosT = merge_cone(cd, cosT, bins[j].cone, bins[j].cosCone);
                }
                costs[i - 1] += saoh(tot.b, tot.flux, cosT);
            }
            float axisCost = std::numeric_limits<float>::max();
            Split axisSplit{dim, 0};
            uint32_t lIdx = begin;
            for (uint32_t i = 0; i < nc; ++i) {
                lIdx += bins[i].n;
                if (costs[i] < axisCost) { axisCost = costs[i]; axisSplit = {dim, lIdx}; }
            }
            axisCost *= dims[largest] / dims[dim];
            if (axisSplit.index == begin || axisSplit.index == end) continue;
            if (axisCost < splitCost) { splitCost = axisCost; split = axisSplit; }
        }
        if (!split.valid()) {
            if (end - begin <= maxPerLeaf) return Split{};
            return Split{largest, (begin + end) / 2};
        }
        return split;
    }

    // src/core/light_tree.cpp:80-161
    uint32_t build_node(std::vector<SortData>& sd, uint64_t mask, uint32_t depth, uint32_t begin, uint32_t end) {
        BBox b = bbox_default();
        for (uint32_t i = begin; i < end; ++i) b = bbox_union(b, sd[i].bounds);
        Split sp = (end - begin) > maxPerLeaf ? split_fn(sd, begin, end, b) : Split{};
        if (sp.valid()) {
            const uint32_t ax = sp.axis;
            std::sort(sd.begin() + begin, sd.begin() + end, [ax](const SortData& p, const SortData& q) {
                float cp = bbox_center(p.bounds, ax), cq = bbox_center(q.bounds, ax);
                if (cp != cq) return cp < cq;
                return p.light < q.light;
            });
            uint32_t idx = (uint32_t)out->nodes.size();
            out->nodes.push_back(PackedNode{});
            build_node(sd, mask | (0ull << depth), depth + 1, begin, sp.index);
            uint32_t right = build_node(sd, mask | (1ull << depth), depth + 1, sp.index, end);
            PackedNode br{};
            br.coreData = ((1u << 31) - 1u) & right;
            out->nodes[idx] = br;
            return idx;
        }
        uint32_t idx = (uint32_t)out->nodes.size();
        out->nodes.push_back(PackedNode{});
        NodeAttr a = light_attr(sd[begin]);
        for (uint32_t i = begin + 1; i < end; ++i) a = attr_add(a, light_attr(sd[i]));
        PackedNode lf{};
        pack_attr(lf, a);
        uint32_t cnt = end - begin;
        lf.coreData = (1u << 31) | ((cnt & ((1u << KLightCountBits) - 1u)) << KLightOffsetBits) | (begin & ((1u << KLightOffsetBits) - 1u));
        out->nodes[idx] = lf;
        for (uint32_t i = begin; i < end; ++i) {
            out->lightPointers[i] = lptr_pack(sd[i].light < ntri ? LT_TRIANGLE : LT_SHAPE, sd[i].light);
            out->bitmasks[sd[i].light] = mask;
        }
        return idx;
    }

    // src/core/light_tree.cpp:396-418 (leaf attributes are re-read from the
    // packed node, i.e. with the decoded direction)
    NodeAttr fill_branch(uint32_t i) {
        UnpackedNode u = unpack_node(out->nodes[i]);
        if (u.leaf) return u.a;
        NodeAttr l = fill_branch(i + 1);
        NodeAttr r = fill_branch(u.right);
        NodeAttr a = attr_add(l, r);
        pack_attr(out->nodes[i], a);
        return a;
    }
};

// Sort data of src/core/rendering_interface.cpp:1116-1150.
static SortData make_sort_data(const float* v, float flux, uint32_t light) {
    SortData s{};
    V3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
    V3 n = normalize(cross(b - a, c - a));
    s.bounds.mn[0] = std::min(a.x, std::min(b.x, c.x));
    s.bounds.mn[1] = std::min(a.y, std::min(b.y, c.y));
    s.bounds.mn[2] = std::min(a.z, std::min(b.z, c.z));
    s.bounds.mx[0] = std::max(a.x, std::max(b.x, c.x));
    s.bounds.mx[1] = std::max(a.y, std::max(b.y, c.y));
    s.bounds.mx[2] = std::max(a.z, std::max(b.z, c.z));
    for (int i = 0; i < 3; ++i)
        if (s.bounds.mx[i] <= s.bounds.mn[i]) {
            s.bounds.mn[i] = s.bounds.mx[i] - std::numeric_limits<float>::epsilon();
            s.bounds.mx[i] = s.bounds.mx[i] + std::numeric_limits<float>::epsilon();
        }
    s.cone[0] = n.x; s.cone[1] = n.y; s.cone[2] = n.z;
    s.cosCone = 1.0f;
    s.flux = flux;
    s.light = light;
    return s;
}

// Sort data of a shape light: a disk's exact box (half-extent
// r sqrt(1 - n_k^2) on axis k, one ulp either side where that is 0, as
// scene.h:355-369 bounds a disk) and its normal cone; a sphere's cube and the
// whole sphere of directions.
static SortData make_shape_sort_data(const orc_shape_light& s, float flux, uint32_t light) {
    SortData d{};
    V3 n = shape_normal(s);
    const float nk[3] = {n.x, n.y, n.z};
    for (int k = 0; k < 3; ++k) {
        if (s.type == 0u) {
            float span = std::max(0.0f, std::sqrt(1.0f - nk[k] * nk[k]));
            d.bounds.mn[k] = span == 0.0f ? std::nextafter(s.position[k], -FLT_MAX) : s.position[k] - span * s.radius;
            d.bounds.mx[k] = span == 0.0f ? std::nextafter(s.position[k], FLT_MAX) : s.position[k] + span * s.radius;
            d.cone[k] = nk[k];
        } else {
            d.bounds.mn[k] = s.position[k] - s.radius;
            d.bounds.mx[k] = s.position[k] + s.radius;
            d.cone[k] = 0.0f;
        }
    }
    d.cosCone = s.type == 0u ? 1.0f : -1.0f;
    d.flux = flux;
    d.light = light;
    return d;
}

// Returns node count (>0) or a negative error. n triangle lights, m shape
// lights (light indices n + i).
int light_tree_build(const float* tri_verts, const float* flux, uint32_t n, const orc_shape_light* shapes,
                     const float* shape_flux, uint32_t m, std::vector<PackedNode>& nodes,
                     std::vector<uint32_t>& lptrs, std::vector<uint64_t>& masks) {
    const uint32_t total = n + m;
    if (total == 0) { nodes.clear(); lptrs.clear(); masks.clear(); return 0; }
    if (total > KMaxLeafLightOffset + 8) return -1;
    TreeOut out;
    out.lightPointers.assign(total, 0);
    out.bitmasks.assign(total, ~0ull);
    std::vector<SortData> sd(total);
    for (uint32_t i = 0; i < n; ++i) sd[i] = make_sort_data(tri_verts + (size_t)i * 9, flux[i], i);
    for (uint32_t i = 0; i < m; ++i) sd[n + i] = make_shape_sort_data(shapes[i], shape_flux[i], n + i);
    Builder b;
    b.verts = tri_verts;
    b.shapes = shapes;
    b.ntri = n;
    b.out = &out;
    b.build_node(sd, 0ull, 0u, 0u, total);
    for (uint32_t i = 0; i < total; ++i)
        if (out.bitmasks[i] == ~0ull) return -2;
    b.fill_branch(0);
    nodes = std::move(out.nodes);
    lptrs = std::move(out.lightPointers);
    masks = std::move(out.bitmasks);
    return (int)nodes.size();
}

}  // namespace orc

namespace orc {
int light_tree_build(const float* tri_verts, const float* flux, uint32_t n, std::vector<PackedNode>& nodes,
                     std::vector<uint32_t>& lptrs, std::vector<uint64_t>& masks) {
    return light_tree_build(tri_verts, flux, n, nullptr, nullptr, 0u, nodes, lptrs, masks);
}
}  // namespace orc
