// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// Ray queries for the CPU path tracer. OptiX's acceleration structure and
// intersection (src/kernels/pathtrace.cu:1162-1177, 1704-1720) are closed, so
// the oracle uses its own median-split BVH and the watertight ray/triangle
// test of Woop, Benthin and Wald (JCGT 2013). The brute-force query below is
// the checker for the product's LBVH traversal.
#include "rb_math.hpp"
#include "rb_bvh.hpp"
#include <vector>
#include <numeric>

namespace orc {

// Woop, Benthin, Wald 2013, "Watertight Ray/Triangle Intersection".
bool ray_triangle(const V3& o, const V3& d, const V3& a, const V3& b, const V3& c, float tmin, float tmax, float& t,
                  float& b1, float& b2) {
    float ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    int kz = ad[0] > ad[1] ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    const float dd[3] = {d.x, d.y, d.z};
    if (dd[kz] < 0.0f) std::swap(kx, ky);
    float Sx = dd[kx] / dd[kz], Sy = dd[ky] / dd[kz], Sz = 1.0f / dd[kz];
    const float A[3] = {a.x - o.x, a.y - o.y, a.z - o.z};
    const float B[3] = {b.x - o.x, b.y - o.y, b.z - o.z};
    const float C[3] = {c.x - o.x, c.y - o.y, c.z - o.z};
    float Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    float Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    float Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * By - (double)Cy * Bx);
        V = (float)((double)Ax * Cy - (double)Ay * Cx);
        W = (float)((double)Bx * Ay - (double)By * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f) { T = -T; det = -det; U = -U; V = -V; W = -W; }
    if (T < tmin * det || T > tmax * det) return false;
    float r = 1.0f / det;
    t = T * r;
    b1 = V * r;
    b2 = W * r;
    return true;
}

namespace {
struct Build {
    const float* v; const uint32_t* idx;
    std::vector<float> cen;  // 3 per tri
    std::vector<float> lo, hi;
    std::vector<uint32_t> order;
    std::vector<CpuBvh::Node>* nodes;
    void tri_box(uint32_t t, float* mn, float* mx) const {
        for (int k = 0; k < 3; ++k) { mn[k] = INFINITY; mx[k] = -INFINITY; }
        for (int j = 0; j < 3; ++j) {
            const float* p = v + 3 * (size_t)idx[3 * (size_t)t + j];
            for (int k = 0; k < 3; ++k) { mn[k] = std::min(mn[k], p[k]); mx[k] = std::max(mx[k], p[k]); }
        }
    }
    uint32_t rec(uint32_t b, uint32_t e) {
        CpuBvh::Node n{};
        for (int k = 0; k < 3; ++k) { n.mn[k] = INFINITY; n.mx[k] = -INFINITY; }
        float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = b; i < e; ++i) {
            uint32_t t = order[i];
            for (int k = 0; k < 3; ++k) {
                n.mn[k] = std::min(n.mn[k], lo[3 * (size_t)t + k]);
                n.mx[k] = std::max(n.mx[k], hi[3 * (size_t)t + k]);
                cmn[k] = std::min(cmn[k], cen[3 * (size_t)t + k]);
                cmx[k] = std::max(cmx[k], cen[3 * (size_t)t + k]);
            }
        }
        uint32_t id = (uint32_t)nodes->size();
        nodes->push_back(n);
        if (e - b <= 4) {
            (*nodes)[id].first = b; (*nodes)[id].count = e - b;
            return id;
        }
        int ax = 0;
        float ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
        if (ext[1] > ext[ax]) ax = 1;
        if (ext[2] > ext[ax]) ax = 2;
        uint32_t m = (b + e) / 2;
        std::nth_element(order.begin() + b, order.begin() + m, order.begin() + e, [&](uint32_t p, uint32_t q) {
            float cp = cen[3 * (size_t)p + ax], cq = cen[3 * (size_t)q + ax];
            return cp != cq ? cp < cq : p < q;
        });
        uint32_t l = rec(b, m);
        (void)l;
        uint32_t r = rec(m, e);
        (*nodes)[id].first = r;  // left child is id+1
        (*nodes)[id].count = 0;
        return id;
    }
};
}  // namespace

void CpuBvh::build(const float* v, const uint32_t* idx, uint32_t nt) {
    verts = v; indices = idx; ntris = nt;
    nodes.clear();
    if (nt == 0) return;
    Build bd{v, idx};
    bd.cen.resize(3 * (size_t)nt); bd.lo.resize(3 * (size_t)nt); bd.hi.resize(3 * (size_t)nt);
    bd.order.resize(nt);
    std::iota(bd.order.begin(), bd.order.end(), 0u);
    for (uint32_t t = 0; t < nt; ++t) {
        bd.tri_box(t, &bd.lo[3 * (size_t)t], &bd.hi[3 * (size_t)t]);
        for (int k = 0; k < 3; ++k) bd.cen[3 * (size_t)t + k] = 0.5f * (bd.lo[3 * (size_t)t + k] + bd.hi[3 * (size_t)t + k]);
    }
    nodes.reserve(2 * (size_t)nt / 2 + 16);
    bd.nodes = &nodes;
    bd.rec(0, nt);
    prims = std::move(bd.order);
}

static inline bool slab(const CpuBvh::Node& n, const V3& o, const V3& inv, float tmax, float& tnear) {
    float t0x = (n.mn[0] - o.x) * inv.x, t1x = (n.mx[0] - o.x) * inv.x;
    float t0y = (n.mn[1] - o.y) * inv.y, t1y = (n.mx[1] - o.y) * inv.y;
    float t0z = (n.mn[2] - o.z) * inv.z, t1z = (n.mx[2] - o.z) * inv.z;
    float tn = std::max(std::max(std::min(t0x, t1x), std::min(t0y, t1y)), std::max(std::min(t0z, t1z), 0.0f));
    float tf = std::min(std::min(std::max(t0x, t1x), std::max(t0y, t1y)), std::min(std::max(t0z, t1z), tmax));
    // a little slack so that boxes grazed by the ray are not culled by rounding
    tf *= 1.0f + 4.0f * FLT_EPSILON;
    tnear = tn;
    return tn <= tf;
}

bool CpuBvh::trace(const V3& o, const V3& d, float tmax, bool anyhit, Hit& h) const {
    if (nodes.empty()) return false;
    V3 inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    bool found = false;
    float best = tmax;
    while (sp) {
        uint32_t ni = stack[--sp];
        const Node& n = nodes[ni];
        float tn;
        if (!slab(n, o, inv, best, tn)) continue;
        if (n.count) {
            for (uint32_t i = n.first; i < n.first + n.count; ++i) {
                uint32_t t = prims[i];
                const float* a = verts + 3 * (size_t)indices[3 * (size_t)t];
                const float* b = verts + 3 * (size_t)indices[3 * (size_t)t + 1];
                const float* c = verts + 3 * (size_t)indices[3 * (size_t)t + 2];
                float tt, b1, b2;
                if (ray_triangle(o, d, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, 0.0f, best, tt, b1, b2)) {
                    found = true;
                    best = tt;
                    h.t = tt; h.prim = t; h.b1 = b1; h.b2 = b2;
                    if (anyhit) return true;
                }
            }
        } else {
            uint32_t l = ni + 1, r = n.first;
            float tl, tr;
            bool hl = slab(nodes[l], o, inv, best, tl), hr = slab(nodes[r], o, inv, best, tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = r; stack[sp++] = l; } else { stack[sp++] = l; stack[sp++] = r; }
            } else if (hl) {
                stack[sp++] = l;
            } else if (hr) {
                stack[sp++] = r;
            }
        }
    }
    return found;
}

}  // namespace orc
