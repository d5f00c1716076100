// Host light-tree builder (the reference's LightTree::Builder,
// src/core/light_tree.cpp:14-418, fed as RenderingInterface::buildLightTree
// does at src/core/rendering_interface.cpp:1065-1161).
//
// Binned SAOH split search over 32 bins on each axis, at most 8 lights per
// leaf, nodes in depth-first order with the left child at index+1, one
// 64-bit path mask per light, branch attributes filled bottom-up from the
// packed (octahedral-quantised) child attributes. The implementation-defined
// std::nth_element partition is replaced by a stable total order on
// (centroid[axis], light index) so the topology is deterministic.
//
// Compiled without floating-point contraction so its FP32 results match a
// straightforward scalar evaluation bit for bit.
//
// The builder is a template over the type that carries the light attributes
// (flux sums, flux-weighted means, variances, directions) and the split-cost
// sums. float is the reference's arithmetic. double is the FP64 power mode:
// the sums and the SAOH costs run in double, and branch attributes merge the
// children's unrounded double attributes instead of the packed ones, so each
// node's flux is the float rounding of a double sum of its lights' fluxes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../device/types.hpp"

namespace rbh {
namespace host {

namespace {

constexpr uint32_t kMaxPerLeaf = 8, kBins = 32;

struct Aabb {
    float lo[3], hi[3];
};
inline Aabb aabb_zero() { return Aabb{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}; }
inline Aabb aabb_empty() {
    const float inf = std::numeric_limits<float>::infinity();
    return Aabb{{inf, inf, inf}, {-inf, -inf, -inf}};
}
inline void grow(Aabb& a, const Aabb& b) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = std::min(a.lo[k], b.lo[k]);
        a.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
}
inline float center(const Aabb& a, int k) { return a.lo[k] + (a.hi[k] - a.lo[k]) * 0.5f; }
inline bool valid(const Aabb& a) { return a.lo[0] <= a.hi[0] && a.lo[1] <= a.hi[1] && a.lo[2] <= a.hi[2]; }

// vMF axis length <-> sharpness (math_util.h:562-587)
template <typename T>
inline T axis_to_sharpness(T l) { return l * (T(3) - l * l) / (T(1) - l * l); }
template <typename T>
inline T sharpness_to_axis(T s) {
    const T a = s / T(3);
    const T b = a * a * a;
    const T c = std::sqrt(T(1) + T(3) * (a * a) * (T(1) + a * a));
    const T th = std::atan2(c, b) / T(3);
    const T d = std::sin(th) * T(1.7320508075688772935274463415059L) - std::cos(th);
    return (s > T(0x1.0p25)) ? T(1) : std::sqrt(T(1) + a * a) * d + a;
}

template <typename T>
struct Attr {
    T mean[3], var, flux, dir[3], sharp;
};

// light_tree_types.h:65-100
template <typename T>
Attr<T> merge(const Attr<T>& A, const Attr<T>& B) {
    Attr<T> r{};
    const T wa = A.flux / (A.flux + B.flux), wb = B.flux / (A.flux + B.flux);
    const T la = sharpness_to_axis(A.sharp), lb = sharpness_to_axis(B.sharp);
    T d[3];
    for (int k = 0; k < 3; ++k) d[k] = wa * la * A.dir[k] + wb * lb * B.dir[k];
    const T dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) r.dir[k] = dl > T(0.0001f) ? d[k] / dl : T(0);
    const T m[3] = {B.mean[0] - A.mean[0], B.mean[1] - A.mean[1], B.mean[2] - A.mean[2]};
    const T m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    for (int k = 0; k < 3; ++k) r.mean[k] = wa * A.mean[k] + wb * B.mean[k];
    r.var = wa * A.var + wb * B.var + wa * wb * m2;
    r.flux = A.flux + B.flux;
    r.sharp = axis_to_sharpness(dl);
    return r;
}

// Octahedral 16+16 encode (math_util.h:147-207) and the [0,1]-scaled decode
// used by unpackNode (light_tree_types.h:174-175).
uint32_t oct_pack(const float* v) {
    const float inv = 1.0f / (std::fabs(v[0]) + std::fabs(v[1]) + std::fabs(v[2]));
    float px = v[0] * inv, py = v[1] * inv;
    if (v[2] <= 0.0f) {
        const float qx = (1.0f - std::fabs(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float qy = (1.0f - std::fabs(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx;
        py = qy;
    }
    auto q = [](float x) { return (uint32_t)(std::min(std::max(x * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.99f); };
    return (q(py) << 16) | q(px);
}
void oct_unpack_unit(uint32_t e, float* out) {
    float x = (float)(e & 0xFFFFu) / 65535.0f, y = (float)(e >> 16) / 65535.0f;
    float z = 1.0f - std::fabs(x) - std::fabs(y);
    const float t = std::max(0.0f, -z);
    x += x >= 0.0f ? -t : t;
    y += y >= 0.0f ? -t : t;
    const float n = 1.0f / std::sqrt(x * x + y * y + z * z);
    out[0] = x * n; out[1] = y * n; out[2] = z * n;
}

struct Prim {
    Aabb box;
    float cone[3];
    float cosCone;
    float flux;
    uint32_t id;
};

template <typename T>
struct Bin {
    Aabb box = aabb_zero();
    uint32_t n = 0;
    T flux = T(0);
    T cone[3] = {T(0), T(0), T(0)};
    T cosCone = T(1);
    template <typename U>
    void add(const Aabb& b, uint32_t c, T f, const U* d) {
        grow(box, b);
        n += c;
        flux += f;
        cone[0] += d[0]; cone[1] += d[1]; cone[2] += d[2];
    }
};

// light_tree.cpp:167-185
template <typename T, typename U>
T cone_union(const T* da, T ca, const U* db, T cb) {
    if (ca == T(-1) || cb == T(-1)) return T(-1);
    const T cd = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
    const T sd = std::sqrt(T(1) - cd * cd);
    const T sb = std::sqrt(T(1) - cb * cb);
    const T ct = cb * cd - sb * sd;
    const T st = sb * cd + cb * sd;
    return st > T(0) ? std::min(cb, ct) : T(-1);
}
// light_tree.cpp:187-199 (surface area x orientation measure x flux)
template <typename T>
T saoh(const Aabb& b, T flux, T cosT) {
    const T pi = T(3.14159265358979323846L);
    T ac = T(0);
    if (valid(b)) {
        const T x = T(b.hi[0]) - T(b.lo[0]), y = T(b.hi[1]) - T(b.lo[1]), z = T(b.hi[2]) - T(b.lo[2]);
        ac = x * y * T(2) + x * z * T(2) + y * z * T(2);
    }
    const T tO = cosT != T(-1) ? std::acos(std::max(std::min(cosT, T(1)), T(-1))) : pi;
    const T tW = std::min(tO + pi / T(2), pi);
    const T sO = std::sin(tO), cO = std::cos(tO);
    const T oc = T(2) * pi * (T(1) - cO) + pi / T(2) * (T(2) * tW * sO - std::cos(tO - T(2) * tW) - T(2) * tO * sO + cO);
    return flux * ac * oc;
}

// Third column of a shape light's unit quaternion (x, y, z, w): the disk's
// emitting normal.
template <typename T>
void shape_normal(const ShapeLight& s, T* n) {
    const T x = s.q[0], y = s.q[1], z = s.q[2], w = s.q[3];
    n[0] = T(2) * (x * z + w * y);
    n[1] = T(2) * (y * z - w * x);
    n[2] = T(1) - T(2) * (x * x + y * y);
    const T inv = T(1) / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] *= inv;
}

template <typename T>
struct Builder {
    static constexpr bool kExact = sizeof(T) > sizeof(float);
    const EmissiveTri* tris;
    const ShapeLight* shapes;   // light index ntri + i
    uint32_t ntri;
    std::vector<Prim> prims;
    std::vector<PackedNode> nodes;
    std::vector<Attr<T>> leafAttrs;  // FP64 mode: unrounded leaf attributes by node index
    std::vector<uint32_t> pointers;
    std::vector<uint64_t> masks;

    // Shape lights: a disk is a one-sided cosine emitter like a triangle
    // (variance r^2/2 of a uniform disk); a sphere emits in every direction
    // (sharpness 0; variance r^2 of a uniform sphere surface).
    Attr<T> shape_attr(const ShapeLight& s) const {
        Attr<T> a{};
        for (int k = 0; k < 3; ++k) a.mean[k] = T(s.pos[k]);
        const T r = T(s.radius);
        if (s.type == 0u) {
            shape_normal<T>(s, a.dir);
            a.var = r * r / T(2);
            a.sharp = axis_to_sharpness(T(0.5));
        } else {
            a.dir[0] = T(0); a.dir[1] = T(0); a.dir[2] = T(1);
            a.var = r * r;
            a.sharp = T(0);
        }
        a.flux = T(0);
        return a;
    }

    Attr<T> leaf_attr(uint32_t id) const {
        if (id >= ntri) return shape_attr(shapes[id - ntri]);
        const float* v = tris[id].v;
        const T e1[3] = {T(v[3]) - T(v[0]), T(v[4]) - T(v[1]), T(v[5]) - T(v[2])};
        const T e2[3] = {T(v[6]) - T(v[0]), T(v[7]) - T(v[1]), T(v[8]) - T(v[2])};
        T n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const T inv = T(1) / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        Attr<T> a{};
        for (int k = 0; k < 3; ++k) {
            a.mean[k] = (T(v[k]) + T(v[3 + k]) + T(v[6 + k])) / T(3);
            a.dir[k] = n[k] * inv;
        }
        const T d11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
        const T d22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
        const T d12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
        a.var = (d11 + d22 - d12) / T(18);
        a.sharp = axis_to_sharpness(T(0.5));
        a.flux = T(0);
        return a;
    }

    static void pack(PackedNode& p, const Attr<T>& a) {
        float dir[3];
        for (int k = 0; k < 3; ++k) {
            p.mean[k] = (float)a.mean[k];
            dir[k] = (float)a.dir[k];
        }
        p.variance = (float)a.var;
        p.dir = oct_pack(dir);
        p.sharpness = (float)a.sharp;
        p.flux = (float)a.flux;
    }

    // light_tree.cpp:162-394; returns axis/index, or index = UINT32_MAX for none
    void split(uint32_t begin, uint32_t end, const Aabb& bounds, uint32_t& outAxis, uint32_t& outIndex) {
        T best = std::numeric_limits<T>::max();
        outAxis = outIndex = UINT32_MAX;
        const float dims[3] = {bounds.hi[0] - bounds.lo[0], bounds.hi[1] - bounds.lo[1], bounds.hi[2] - bounds.lo[2]};
        const uint32_t largest = dims[0] > dims[1] && dims[0] > dims[2] ? 0u : (dims[1] > dims[2] ? 1u : 2u);
        Bin<T> bins[kBins];
        T costs[kBins - 1];
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const float w = bounds.hi[ax] - bounds.lo[ax];
            const float scale = w > std::numeric_limits<float>::min() ? (float)kBins / w : 0.0f;
            auto binOf = [&](const Prim& p) {
                return std::min((uint32_t)((center(p.box, (int)ax) - bounds.lo[ax]) * scale), kBins - 1);
            };
            for (auto& b : bins) b = Bin<T>{};
            for (uint32_t i = begin; i < end; ++i)
                bins[binOf(prims[i])].add(prims[i].box, 1, T(prims[i].flux), prims[i].cone);
            for (auto& b : bins) {
                const T l = std::sqrt(b.cone[0] * b.cone[0] + b.cone[1] * b.cone[1] + b.cone[2] * b.cone[2]);
                if (l < std::numeric_limits<T>::min()) {
                    b.cosCone = T(-1);
                } else {
                    b.cosCone = T(1);
                    b.cone[0] /= l; b.cone[1] /= l; b.cone[2] /= l;
                }
            }
            for (uint32_t i = begin; i < end; ++i) {
                Bin<T>& b = bins[binOf(prims[i])];
                b.cosCone = cone_union(b.cone, b.cosCone, prims[i].cone, T(prims[i].cosCone));
            }
            auto sweep_cost = [&](const Bin<T>& tot, uint32_t j0, uint32_t j1) {
                T cosT = T(-1);
                const T l = std::sqrt(tot.cone[0] * tot.cone[0] + tot.cone[1] * tot.cone[1] + tot.cone[2] * tot.cone[2]);
                if (l > std::numeric_limits<T>::min()) {
                    cosT = T(1);
                    const T cd[3] = {tot.cone[0] / l, tot.cone[1] / l, tot.cone[2] / l};
                    for (uint32_t j = j0; j <= j1; ++j) cosT = cone_union(cd, cosT, bins[j].cone, bins[j].cosCone);
                }
                return saoh(tot.box, tot.flux, cosT);
            };
            Bin<T> tot;
            for (uint32_t i = 0; i < kBins - 1; ++i) {
                tot.add(bins[i].box, bins[i].n, bins[i].flux, bins[i].cone);
                costs[i] = sweep_cost(tot, 0, i);
            }
            tot = Bin<T>{};
            for (uint32_t i = kBins - 1; i > 0; --i) {
                tot.add(bins[i].box, bins[i].n, bins[i].flux, bins[i].cone);
                costs[i - 1] += sweep_cost(tot, i, kBins - 1);
            }
            T axCost = std::numeric_limits<T>::max();
            uint32_t axIndex = 0, run = begin;
            for (uint32_t i = 0; i < kBins - 1; ++i) {
                run += bins[i].n;
                if (costs[i] < axCost) { axCost = costs[i]; axIndex = run; }
            }
            axCost *= T(dims[largest] / dims[ax]);
            if (axIndex == begin || axIndex == end) continue;
            if (axCost < best) { best = axCost; outAxis = ax; outIndex = axIndex; }
        }
        if (outIndex == UINT32_MAX && end - begin > kMaxPerLeaf) {
            outAxis = largest;
            outIndex = (begin + end) / 2;
        }
    }

    uint32_t node(uint64_t mask, uint32_t depth, uint32_t begin, uint32_t end) {
        Aabb b = aabb_empty();
        for (uint32_t i = begin; i < end; ++i) grow(b, prims[i].box);
        uint32_t ax = UINT32_MAX, at = UINT32_MAX;
        if (end - begin > kMaxPerLeaf) split(begin, end, b, ax, at);
        const uint32_t me = (uint32_t)nodes.size();
        nodes.push_back(PackedNode{});
        if (at != UINT32_MAX) {
            std::sort(prims.begin() + begin, prims.begin() + end, [ax](const Prim& p, const Prim& q) {
                const float cp = center(p.box, (int)ax), cq = center(q.box, (int)ax);
                return cp != cq ? cp < cq : p.id < q.id;
            });
            node(mask, depth + 1, begin, at);
            const uint32_t right = node(mask | (1ull << depth), depth + 1, at, end);
            nodes[me].core = ((1u << 31) - 1u) & right;
            return me;
        }
        Attr<T> a = leaf_attr(prims[begin].id);
        a.flux = T(prims[begin].flux);
        for (uint32_t i = begin + 1; i < end; ++i) {
            Attr<T> o = leaf_attr(prims[i].id);
            o.flux = T(prims[i].flux);
            a = merge(a, o);
        }
        pack(nodes[me], a);
        if (kExact) {
            if (leafAttrs.size() <= me) leafAttrs.resize(me + 1);
            leafAttrs[me] = a;
        }
        const uint32_t cnt = end - begin;
        nodes[me].core = (1u << 31) | ((cnt & 15u) << 27) | (begin & ((1u << 27) - 1u));
        for (uint32_t i = begin; i < end; ++i) {
            pointers[i] = ((prims[i].id < ntri ? 0u : 3u) << 30) | (prims[i].id & ((1u << 30) - 1u));
            masks[prims[i].id] = mask;
        }
        return me;
    }

    Attr<T> unpacked(uint32_t i) const {
        const PackedNode& p = nodes[i];
        Attr<T> a{};
        for (int k = 0; k < 3; ++k) a.mean[k] = p.mean[k];
        a.var = p.variance;
        a.flux = p.flux;
        a.sharp = p.sharpness;
        float dir[3];
        oct_unpack_unit(p.dir, dir);
        for (int k = 0; k < 3; ++k) a.dir[k] = dir[k];
        return a;
    }
    // light_tree.cpp:396-418 (float: merges the packed children as the
    // reference does; double: merges the unrounded children)
    Attr<T> fill(uint32_t i) {
        if ((nodes[i].core >> 31) != 0u) return kExact ? leafAttrs[i] : unpacked(i);
        const Attr<T> l = fill(i + 1);
        const Attr<T> r = fill(nodes[i].core);
        const Attr<T> a = merge(l, r);
        pack(nodes[i], a);
        return a;
    }
};

// Sort data of a shape light: the disk's box as scene.h:355-369 gives it
// (extent r sqrt(1 - n_k^2) along axis k, one ulp either side where that is
// 0), the sphere's cube; the disk's normal cone, the whole sphere for a sphere.
inline void shape_prim(const ShapeLight& s, float flux, uint32_t id, Prim& p) {
    float n[3];
    shape_normal<float>(s, n);
    for (int k = 0; k < 3; ++k) {
        if (s.type == 0u) {
            const float span = std::max(0.0f, std::sqrt(1.0f - n[k] * n[k]));
            p.box.lo[k] = span == 0.0f ? std::nextafter(s.pos[k], -std::numeric_limits<float>::max())
                                       : s.pos[k] - span * s.radius;
            p.box.hi[k] = span == 0.0f ? std::nextafter(s.pos[k], std::numeric_limits<float>::max())
                                       : s.pos[k] + span * s.radius;
            p.cone[k] = n[k];
        } else {
            p.box.lo[k] = s.pos[k] - s.radius;
            p.box.hi[k] = s.pos[k] + s.radius;
            p.cone[k] = 0.0f;
        }
    }
    p.cosCone = s.type == 0u ? 1.0f : -1.0f;
    p.flux = flux;
    p.id = id;
}

template <typename T>
int build(const EmissiveTri* tris, const float* flux, uint32_t n, const ShapeLight* shapes, const float* shapeFlux,
          uint32_t m, std::vector<PackedNode>& nodes, std::vector<uint32_t>& pointers, std::vector<uint64_t>& masks) {
    Builder<T> b;
    b.tris = tris;
    b.shapes = shapes;
    b.ntri = n;
    b.prims.resize(n + m);
    for (uint32_t i = 0; i < m; ++i) shape_prim(shapes[i], shapeFlux[i], n + i, b.prims[n + i]);
    for (uint32_t i = 0; i < n; ++i) {
        const float* v = tris[i].v;
        Prim& p = b.prims[i];
        const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float inv = 1.0f / std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        for (int k = 0; k < 3; ++k) {
            p.box.lo[k] = std::min(v[k], std::min(v[3 + k], v[6 + k]));
            p.box.hi[k] = std::max(v[k], std::max(v[3 + k], v[6 + k]));
            if (p.box.hi[k] <= p.box.lo[k]) {
                p.box.lo[k] = p.box.hi[k] - std::numeric_limits<float>::epsilon();
                p.box.hi[k] = p.box.hi[k] + std::numeric_limits<float>::epsilon();
            }
            p.cone[k] = nn[k] * inv;
        }
        p.cosCone = 1.0f;
        p.flux = flux[i];
        p.id = i;
    }
    const uint32_t total = n + m;
    b.pointers.assign(total, 0u);
    b.masks.assign(total, ~0ull);
    b.nodes.reserve(2 * (size_t)total);
    b.node(0ull, 0u, 0u, total);
    for (uint32_t i = 0; i < total; ++i)
        if (b.masks[i] == ~0ull) return -2;
    b.fill(0);
    nodes = std::move(b.nodes);
    pointers = std::move(b.pointers);
    masks = std::move(b.masks);
    return (int)nodes.size();
}

}  // namespace

// Returns the node count, 0 for no lights, or a negative error. n triangle
// lights (light indices [0, n)) and m shape lights (light indices n + i);
// fp64 selects the FP64 power mode.
int build_light_tree(const EmissiveTri* tris, const float* flux, uint32_t n, const ShapeLight* shapes,
                     const float* shapeFlux, uint32_t m, std::vector<PackedNode>& nodes,
                     std::vector<uint32_t>& pointers, std::vector<uint64_t>& masks, bool fp64) {
    nodes.clear();
    pointers.clear();
    masks.clear();
    if (n + (uint64_t)m == 0) return 0;
    if (n + (uint64_t)m > (1u << 27) - 1u + kMaxPerLeaf) return -1;
    return fp64 ? build<double>(tris, flux, n, shapes, shapeFlux, m, nodes, pointers, masks)
                : build<float>(tris, flux, n, shapes, shapeFlux, m, nodes, pointers, masks);
}

}  // namespace host
}  // namespace rbh
