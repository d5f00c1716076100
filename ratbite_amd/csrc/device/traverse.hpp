// LBVH traversal: closest-hit and any-hit queries replacing optixTrace /
// optixTraverse(TERMINATE_ON_FIRST_HIT) (src/kernels/pathtrace.cu:1162-1177,
// 1704-1720). Stack in LDS (one column per lane, so consecutive lanes hit
// consecutive banks); ray/box slabs with a conservative far bound; the
// watertight ray/triangle test of Woop, Benthin and Wald (JCGT 2013) with
// per-ray shear constants.
#pragma once
#include "vec.hpp"
#include "types.hpp"

namespace rbh {
namespace dev {

struct RayShear {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
__device__ __forceinline__ float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ RayShear make_shear(V3 d) {
    RayShear s;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    s.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    s.kx = s.kz == 2 ? 0 : s.kz + 1;
    s.ky = s.kx == 2 ? 0 : s.kx + 1;
    const float dz = comp(d, s.kz);
    if (dz < 0.0f) { const int t = s.kx; s.kx = s.ky; s.ky = t; }
    s.Sx = comp(d, s.kx) / dz;
    s.Sy = comp(d, s.ky) / dz;
    s.Sz = 1.0f / dz;
    return s;
}

struct Hit {
    float t, b1, b2;
    uint32_t leaf;   // leaf-order triangle record of the hit; 0xFFFFFFFF: miss
};

__device__ __forceinline__ bool tri_hit(const RayShear& s, V3 o, const TriRecord& tr, float tmax, float& t, float& b1,
                                        float& b2) {
    // No FMA contraction: the 2D edge functions of a shared edge must round
    // identically from both triangles for the test to stay watertight.
#pragma clang fp contract(off)
    const V3 A{tr.v0x - o.x, tr.v0y - o.y, tr.v0z - o.z};
    const V3 B{tr.v1x - o.x, tr.v1y - o.y, tr.v1z - o.z};
    const V3 C{tr.v2x - o.x, tr.v2y - o.y, tr.v2z - o.z};
    const float Akz = comp(A, s.kz), Bkz = comp(B, s.kz), Ckz = comp(C, s.kz);
    const float Ax = comp(A, s.kx) - s.Sx * Akz, Ay = comp(A, s.ky) - s.Sy * Akz;
    const float Bx = comp(B, s.kx) - s.Sx * Bkz, By = comp(B, s.ky) - s.Sy * Bkz;
    const float Cx = comp(C, s.kx) - s.Sx * Ckz, Cy = comp(C, s.ky) - s.Sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * By - (double)Cy * Bx);
        V = (float)((double)Ax * Cy - (double)Ay * Cx);
        W = (float)((double)Bx * Ay - (double)By * Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (s.Sz * Akz) + V * (s.Sz * Bkz) + W * (s.Sz * Ckz);
    if (det < 0.0f) { T = -T; det = -det; V = -V; W = -W; }
    if (T < 0.0f || T > tmax * det) return false;
    const float r = 1.0f / det;
    t = T * r;
    b1 = V * r;
    b2 = W * r;
    return true;
}

// Traversal stack of at most kMaxStack entries (the LBVH builder rejects
// trees deeper than that). The LDS_DEPTH entries nearest the bottom live in
// LDS, one column per lane (stride = lanes per block), so consecutive lanes
// hit consecutive banks; deeper entries go to a per-lane spill column in
// global memory (entry k of lane t at spill[k * spillStride + t]), which only
// rays with many deferred subtrees touch. A small LDS part keeps more waves
// resident per CU than a full 64-entry LDS stack would.
constexpr int kMaxStack = 64;     // binary tree: depth <= 64
constexpr int kMaxStack4 = 192;   // four-wide tree: at most 3 deferred children per level, <= 64 levels

template <int LDS_DEPTH>
struct TravStack {
    int* __restrict__ lds;
    int stride;
    int* __restrict__ spill;
    uint32_t spillStride;
    __device__ __forceinline__ void put(int sp, int v) const {
        if (sp < LDS_DEPTH) lds[sp * stride] = v;
        else spill[(size_t)(sp - LDS_DEPTH) * spillStride] = v;
    }
    __device__ __forceinline__ int get(int sp) const {
        if (sp < LDS_DEPTH) return lds[sp * stride];
        return spill[(size_t)(sp - LDS_DEPTH) * spillStride];
    }
};

// Traverses the LBVH. Returns the closest hit below tmax or, with ANY, the
// first one found.
template <bool ANY, int LDS_DEPTH>
__device__ __forceinline__ Hit traverse(const BvhNode* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                        uint32_t nnodes, V3 o, V3 d, float tmax, const TravStack<LDS_DEPTH>& stack) {
    Hit h;
    h.t = tmax;
    h.leaf = 0xFFFFFFFFu;
    h.b1 = h.b2 = 0.0f;
    const RayShear sh = make_shear(d);
    const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    int node = nnodes ? 0 : ~0;
    int sp = 0;
    while (true) {
        if (node >= 0) {
            const float4* q = reinterpret_cast<const float4*>(nodes + node);
            const float4 a = q[0], b = q[1], c = q[2];
            const int4 e = reinterpret_cast<const int4*>(q)[3];
            const float best = h.t * 1.00000024f;
            // left child
            const float l0x = (a.x - o.x) * ix, l1x = (a.y - o.x) * ix;
            const float l0y = (a.z - o.y) * iy, l1y = (a.w - o.y) * iy;
            const float l0z = (c.x - o.z) * iz, l1z = (c.y - o.z) * iz;
            const float lnear = fmaxf(fmaxf(fminf(l0x, l1x), fminf(l0y, l1y)), fmaxf(fminf(l0z, l1z), 0.0f));
            const float lfar = fminf(fminf(fmaxf(l0x, l1x), fmaxf(l0y, l1y)), fminf(fmaxf(l0z, l1z), best));
            // right child
            const float r0x = (b.x - o.x) * ix, r1x = (b.y - o.x) * ix;
            const float r0y = (b.z - o.y) * iy, r1y = (b.w - o.y) * iy;
            const float r0z = (c.z - o.z) * iz, r1z = (c.w - o.z) * iz;
            const float rnear = fmaxf(fmaxf(fminf(r0x, r1x), fminf(r0y, r1y)), fmaxf(fminf(r0z, r1z), 0.0f));
            const float rfar = fminf(fminf(fmaxf(r0x, r1x), fmaxf(r0y, r1y)), fminf(fmaxf(r0z, r1z), best));
            const bool hl = lnear <= lfar * 1.00000024f, hr = rnear <= rfar * 1.00000024f;
            if (hl && hr) {
                const bool lfirst = lnear <= rnear;
                stack.put(sp, lfirst ? e.y : e.x);
                ++sp;
                node = lfirst ? e.x : e.y;
                continue;
            }
            if (hl) { node = e.x; continue; }
            if (hr) { node = e.y; continue; }
        } else {
            const TriRecord& tr = tris[~node];
            float t, b1, b2;
            if (tri_hit(sh, o, tr, h.t, t, b1, b2)) {
                h.t = t;
                h.b1 = b1;
                h.b2 = b2;
                h.leaf = (uint32_t)~node;
                if (ANY) return h;
            }
        }
        if (sp == 0) break;
        --sp;
        node = stack.get(sp);
    }
    return h;
}

// Four-wide traversal: a node fetch tests four child boxes; leaf children
// (single triangles) are intersected on the spot, inner children that
// survive the updated distance are visited nearest first, the others pushed.
// Four-wide traversal as a resumable state machine, so that a persistent
// kernel can give a lane a new ray as soon as its previous one finishes
// while the other lanes of the wave keep traversing.
struct TravRay {
    V3 o, d;
    float ix, iy, iz;
    RayShear sh;
    Hit h;
    int node;   // next node to visit (~i: the root is triangle i)
    int sp;     // entries on the traversal stack
};

__device__ __forceinline__ void trav_init(TravRay& r, V3 o, V3 d, float tmax, uint32_t nnodes) {
    r.o = o;
    r.d = d;
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    r.sh = make_shear(d);
    r.h.t = tmax;
    r.h.leaf = 0xFFFFFFFFu;
    r.h.b1 = r.h.b2 = 0.0f;
    r.node = nnodes ? 0 : ~0;
    r.sp = 0;
}

// Child boxes and references of one four-wide node.
struct Box4 {
    float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
    int ch[4];
};
__device__ __forceinline__ void load_node(const Bvh4Node* __restrict__ nodes, int i, Box4& b) {
    const float4* q = reinterpret_cast<const float4*>(nodes + i);
    const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
    const int4 ch4 = reinterpret_cast<const int4*>(q)[6];
    b.lx[0] = lx.x; b.lx[1] = lx.y; b.lx[2] = lx.z; b.lx[3] = lx.w;
    b.hx[0] = hx.x; b.hx[1] = hx.y; b.hx[2] = hx.z; b.hx[3] = hx.w;
    b.ly[0] = ly.x; b.ly[1] = ly.y; b.ly[2] = ly.z; b.ly[3] = ly.w;
    b.hy[0] = hy.x; b.hy[1] = hy.y; b.hy[2] = hy.z; b.hy[3] = hy.w;
    b.lz[0] = lz.x; b.lz[1] = lz.y; b.lz[2] = lz.z; b.lz[3] = lz.w;
    b.hz[0] = hz.x; b.hz[1] = hz.y; b.hz[2] = hz.z; b.hz[3] = hz.w;
    b.ch[0] = ch4.x; b.ch[1] = ch4.y; b.ch[2] = ch4.z; b.ch[3] = ch4.w;
}
__device__ __forceinline__ void load_node(const Bvh4QNode* __restrict__ nodes, int i, Box4& b) {
    const uint4* q = reinterpret_cast<const uint4*>(nodes + i);
    const uint4 h = q[0], c = q[1], a = q[2];
    const uint2 e = reinterpret_cast<const uint2*>(q)[6];
    const float ox = __uint_as_float(h.x), oy = __uint_as_float(h.y), oz = __uint_as_float(h.z);
    const float sx = qnode_step(h.w & 0xFFu), sy = qnode_step((h.w >> 8) & 0xFFu), sz = qnode_step((h.w >> 16) & 0xFFu);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        b.lx[k] = qnode_decode(ox, (a.x >> (8 * k)) & 0xFFu, sx);
        b.ly[k] = qnode_decode(oy, (a.y >> (8 * k)) & 0xFFu, sy);
        b.lz[k] = qnode_decode(oz, (a.z >> (8 * k)) & 0xFFu, sz);
        b.hx[k] = qnode_decode(ox, (a.w >> (8 * k)) & 0xFFu, sx);
        b.hy[k] = qnode_decode(oy, (e.x >> (8 * k)) & 0xFFu, sy);
        b.hz[k] = qnode_decode(oz, (e.y >> (8 * k)) & 0xFFu, sz);
    }
    b.ch[0] = (int)c.x; b.ch[1] = (int)c.y; b.ch[2] = (int)c.z; b.ch[3] = (int)c.w;
}

// Visits r.node: tests its triangle children (nearest hit so far shortens
// the ray), descends into the nearest box hit and defers the others,
// nearest last on the stack. Returns true when the query is complete.
// NODE: Bvh4Node or its 64-byte quantised form Bvh4QNode.
template <bool ANY, int LDS_DEPTH, typename NODE>
__device__ __forceinline__ bool trav_step(const NODE* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                          TravRay& r, const TravStack<LDS_DEPTH>& stack) {
    if (r.node < 0) {
        // a single-triangle scene: the root is the leaf
        const TriRecord& tr = tris[~r.node];
        float t, b1, b2;
        if (tri_hit(r.sh, r.o, tr, r.h.t, t, b1, b2)) {
            r.h.t = t;
            r.h.b1 = b1;
            r.h.b2 = b2;
            r.h.leaf = (uint32_t)~r.node;
        }
        return true;
    }
    Box4 bx;
    load_node(nodes, r.node, bx);
    const float* lxa = bx.lx; const float* hxa = bx.hx;
    const float* lya = bx.ly; const float* hya = bx.hy;
    const float* lza = bx.lz; const float* hza = bx.hz;
    const int* cha = bx.ch;
    float tn[4];
    bool in[4];
    {
        const float best = r.h.t * 1.00000024f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float x0 = (lxa[k] - r.o.x) * r.ix, x1 = (hxa[k] - r.o.x) * r.ix;
            const float y0 = (lya[k] - r.o.y) * r.iy, y1 = (hya[k] - r.o.y) * r.iy;
            const float z0 = (lza[k] - r.o.z) * r.iz, z1 = (hza[k] - r.o.z) * r.iz;
            const float tnear = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
            const float tfar = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), best));
            tn[k] = tnear;
            in[k] = cha[k] != kBvh4Empty && tnear <= tfar * 1.00000024f;
        }
    }
    // Triangle children first (a hit shortens the ray for the boxes), in
    // slot order. One loop over each lane's own triangle slots: a wave runs
    // it as often as its lane with the most triangle children needs, instead
    // of once for every slot in which some lane has one.
    uint32_t triMask = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) triMask |= (in[k] && cha[k] < 0) ? (1u << k) : 0u;
    while (triMask) {
        const uint32_t k = (uint32_t)__builtin_ctz(triMask);
        triMask &= triMask - 1u;
        const int c = k == 0u ? cha[0] : k == 1u ? cha[1] : k == 2u ? cha[2] : cha[3];
        const TriRecord& tr = tris[~c];
        float t, b1, b2;
        if (tri_hit(r.sh, r.o, tr, r.h.t, t, b1, b2)) {
            r.h.t = t;
            r.h.b1 = b1;
            r.h.b2 = b2;
            r.h.leaf = (uint32_t)~c;
            if (ANY) return true;
        }
    }
    // Surviving inner children nearest first. A child's key is the bit
    // pattern of its entry distance (>= 0, so ordered as an unsigned
    // integer; the sign bit cleared for -0) with its slot in the two low bits, so a five-comparator
    // network on unsigned min/max sorts them without branches or indexed
    // registers; children within 4 ulp of each other go in slot order.
    // Culled and empty slots sort last as ~0.
    const float limit = r.h.t * 1.00000024f * 1.00000024f;
    uint32_t key[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        key[k] = (in[k] && cha[k] >= 0 && tn[k] <= limit) ? ((__float_as_uint(tn[k]) & 0x7FFFFFFCu) | (uint32_t)k)
                                                          : ~0u;
    {
        auto cx = [](uint32_t& a, uint32_t& b) {
            const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            a = lo;
            b = hi;
        };
        cx(key[0], key[1]);
        cx(key[2], key[3]);
        cx(key[0], key[2]);
        cx(key[1], key[3]);
        cx(key[1], key[2]);
    }
    auto child = [&](uint32_t kk) {
        const uint32_t s = kk & 3u;
        return s == 0u ? cha[0] : s == 1u ? cha[1] : s == 2u ? cha[2] : cha[3];
    };
    if (key[0] != ~0u) {
        // farthest pushed first, so the nearer ones are popped first
        if (key[3] != ~0u) stack.put(r.sp++, child(key[3]));
        if (key[2] != ~0u) stack.put(r.sp++, child(key[2]));
        if (key[1] != ~0u) stack.put(r.sp++, child(key[1]));
        r.node = child(key[0]);
        return false;
    }
    if (r.sp == 0) return true;
    --r.sp;
    r.node = stack.get(r.sp);
    return false;
}

// Whole query in one call: the closest hit below tmax or, with ANY, the
// first one found.
template <bool ANY, int LDS_DEPTH, typename NODE>
__device__ __forceinline__ Hit traverse4(const NODE* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                         uint32_t nnodes, V3 o, V3 d, float tmax, const TravStack<LDS_DEPTH>& stack) {
    TravRay r;
    trav_init(r, o, d, tmax, nnodes);
    while (!trav_step<ANY>(nodes, tris, r, stack)) {
    }
    return r.h;
}

}  // namespace dev
}  // namespace rbh
