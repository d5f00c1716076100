// Linear BVH build on gfx950, replacing optixAccelBuild
// (src/core/rendering_interface.cpp:733-1064):
//   1. triangle centroid bounds (block reduction + ordered-int atomics)
//   2. Morton codes on the centroid grid, made unique with the triangle
//      index in the low bits of a 64-bit key: the index takes the bits n
//      needs, the Morton code as many of the rest as fit three axes (at most
//      21 per axis; 14 for 4.1M triangles, where 10 bits per axis would put
//      84% of the triangles in cells shared with others)
//   3. stable LSD radix sort of the keys (radix.hpp: ballot-ranked 8-bit digits)
//   4. Karras (HPG 2012) hierarchy: internal node i covers a key range and
//      splits at the highest differing bit
//   5. refit level by level from the deepest internal nodes up (one launch
//      per depth, so kernel boundaries order the box writes; no cross-XCD
//      fences or arrival counters); node depths by pointer jumping, the
//      depth order by a radix sort of (depth, node) pairs
//   6. emission of 64-byte two-child nodes (both child boxes in the parent)
//      and of triangles in leaf order
//   7. collapse into 128-byte four-wide nodes, top down one level per launch:
//      a four-wide node takes its binary node's children and keeps opening
//      the inner child with the largest box surface until it has four
// Every root-to-leaf path strictly lengthens the common key prefix, so the
// depth is at most 64 and a 64-entry traversal stack never overflows.
#include "../context.hpp"
#include "radix.hpp"

namespace rbh {

namespace {

constexpr int kBuildBlock = 256;

__device__ __forceinline__ uint32_t f2ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_bounds_init(uint32_t* __restrict__ b) {
    if (threadIdx.x < 3) b[threadIdx.x] = 0xFFFFFFFFu;           // min (ordered)
    else if (threadIdx.x < 6) b[threadIdx.x] = 0u;               // max (ordered)
    else if (threadIdx.x == 6) b[6] = 0u;                         // depth
}

__global__ __launch_bounds__(kBuildBlock) void k_centroid_bounds(const float* __restrict__ v, const uint32_t* __restrict__ idx,
                                                                 uint32_t n, uint32_t* __restrict__ b) {
    __shared__ float smin[3][kBuildBlock], smax[3][kBuildBlock];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = t; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t a = idx[3 * i], bb = idx[3 * i + 1], c = idx[3 * i + 2];
        for (int k = 0; k < 3; ++k) {
            const float lo = fminf(fminf(v[3 * a + k], v[3 * bb + k]), v[3 * c + k]);
            const float hi = fmaxf(fmaxf(v[3 * a + k], v[3 * bb + k]), v[3 * c + k]);
            const float ce = 0.5f * (lo + hi);
            mn[k] = fminf(mn[k], ce);
            mx[k] = fmaxf(mx[k], ce);
        }
    }
    for (int k = 0; k < 3; ++k) { smin[k][threadIdx.x] = mn[k]; smax[k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int s = kBuildBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + s]);
                smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&b[k], f2ord(smin[k][0]));
            atomicMax(&b[3 + k], f2ord(smax[k][0]));
        }
}

// bits 0..20 of x to bits 0, 3, 6, ..., 60
__device__ __forceinline__ unsigned long long spread21(uint32_t x) {
    unsigned long long y = x & 0x1FFFFFu;
    y = (y | (y << 32)) & 0x001F00000000FFFFull;
    y = (y | (y << 16)) & 0x001F0000FF0000FFull;
    y = (y | (y << 8)) & 0x100F00F00F00F00Full;
    y = (y | (y << 4)) & 0x10C30C30C30C30C3ull;
    y = (y | (y << 2)) & 0x1249249249249249ull;
    return y;
}

__global__ __launch_bounds__(kBuildBlock) void k_morton(const float* __restrict__ v, const uint32_t* __restrict__ idx,
                                                        uint32_t n, const uint32_t* __restrict__ b, int axisBits,
                                                        int idxBits, unsigned long long* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = ord2f(b[k]);
        ext[k] = ord2f(b[3 + k]) - lo[k];
    }
    const float cells = (float)(1u << axisBits);          // exact: a power of two
    const uint32_t a = idx[3 * i], bb = idx[3 * i + 1], c = idx[3 * i + 2];
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float l = fminf(fminf(v[3 * a + k], v[3 * bb + k]), v[3 * c + k]);
        const float h = fmaxf(fmaxf(v[3 * a + k], v[3 * bb + k]), v[3 * c + k]);
        const float ce = 0.5f * (l + h);
        const float u = ext[k] > 0.0f ? (ce - lo[k]) / ext[k] : 0.5f;
        q[k] = (uint32_t)fminf(fmaxf(u * cells, 0.0f), cells - 1.0f);
    }
    const unsigned long long m = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    keys[i] = (m << idxBits) | i;
}

__device__ __forceinline__ int common_prefix(const unsigned long long* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);
}

// Karras 2012, Figure 4. Internal nodes 0..n-2; child ids >= 0 internal,
// < 0 leaf ~sortedIndex. parent[] holds internal parents of internal nodes
// (slot i) and of leaves (slot n-1+j).
__global__ __launch_bounds__(kBuildBlock) void k_karras(const unsigned long long* __restrict__ k, int n,
                                                        int2* __restrict__ child, int* __restrict__ parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (common_prefix(k, n, i, i + 1) - common_prefix(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = common_prefix(k, n, i, i - d);
    int lmax = 2;
    while (common_prefix(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (common_prefix(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = common_prefix(k, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (common_prefix(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int split = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const int left = (first == split) ? ~split : split;
    const int right = (last == split + 1) ? ~(split + 1) : split + 1;
    child[i] = make_int2(left, right);
    if (left >= 0) parent[left] = i; else parent[n - 1 + split] = i;
    if (right >= 0) parent[right] = i; else parent[n - 1 + split + 1] = i;
    if (i == 0) parent[0] = -1;
}

struct Box { float lo[3], hi[3]; };

// Triangle records in leaf (Morton) order and their boxes.
__global__ __launch_bounds__(kBuildBlock) void k_leaves(const float* __restrict__ v, const uint32_t* __restrict__ idx,
                                                        const uint32_t* __restrict__ triMaterial,
                                                        const uint32_t* __restrict__ triLight,
                                                        const Material* __restrict__ mats, uint32_t nmats,
                                                        const unsigned long long* __restrict__ keys, int n, int idxBits,
                                                        float* __restrict__ lbox, TriRecord* __restrict__ tris) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t tri = (uint32_t)(keys[j] & ((1ull << idxBits) - 1ull));
    const uint32_t a = idx[3 * tri], b = idx[3 * tri + 1], c = idx[3 * tri + 2];
    TriRecord r;
    r.v0x = v[3 * a]; r.v0y = v[3 * a + 1]; r.v0z = v[3 * a + 2]; r.prim = tri;
    r.v1x = v[3 * b]; r.v1y = v[3 * b + 1]; r.v1z = v[3 * b + 2]; r.material = triMaterial[tri];
    r.v2x = v[3 * c]; r.v2y = v[3 * c + 1]; r.v2z = v[3 * c + 2]; r.light = triLight[tri];
    r.i0 = a; r.i1 = b; r.i2 = c;
    r.shadeClass = r.material < nmats ? shade_class(mats[r.material]) : (uint32_t)SC_SURFACE;
    tris[j] = r;
    float* lb = lbox + 6 * (size_t)j;
    lb[0] = fminf(fminf(r.v0x, r.v1x), r.v2x); lb[3] = fmaxf(fmaxf(r.v0x, r.v1x), r.v2x);
    lb[1] = fminf(fminf(r.v0y, r.v1y), r.v2y); lb[4] = fmaxf(fmaxf(r.v0y, r.v1y), r.v2y);
    lb[2] = fminf(fminf(r.v0z, r.v1z), r.v2z); lb[5] = fmaxf(fmaxf(r.v0z, r.v1z), r.v2z);
}

constexpr int kLevels = 65;   // internal-node depths 0..63 (64 collects anything deeper)

// Depth of every internal node (ancestors up to the root) and a histogram.
// Depth of every internal node by pointer jumping over the parent links:
// (ancestor, distance) pairs, each round adds the ancestor's distance and
// jumps to the ancestor's ancestor, so ceil(log2(64)) + 1 = 7 rounds reach the
// root from depth 64 (deeper trees are rejected after the build).
constexpr int kJumpRounds = 7;
__global__ __launch_bounds__(kBuildBlock) void k_jump_init(int nInt, const int* __restrict__ parent,
                                                           int2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInt) return;
    const int p = parent[i];
    out[i] = make_int2(p, p >= 0 ? 1 : 0);
}
__global__ __launch_bounds__(kBuildBlock) void k_jump(int nInt, const int2* __restrict__ in, int2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInt) return;
    const int2 a = in[i];
    if (a.x < 0) {
        out[i] = a;
        return;
    }
    const int2 b = in[a.x];
    out[i] = make_int2(b.x, a.y + b.y);
}
// sort keys: the depth (capped at kLevels - 1) and the node index
__global__ __launch_bounds__(kBuildBlock) void k_depth_keys(int nInt, const int2* __restrict__ jumped,
                                                            unsigned int* __restrict__ key, int* __restrict__ id) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInt) return;
    key[i] = (unsigned int)min(jumped[i].y, kLevels - 1);
    id[i] = i;
}
// start[d] = first position of depth d in the depth-sorted order (d <= kLevels)
__global__ void k_level_bounds(const unsigned int* __restrict__ sortedKey, int nInt, int* __restrict__ start) {
    const int d = threadIdx.x;
    if (d > kLevels) return;
    int lo = 0, hi = nInt;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)sortedKey[mid] < d) lo = mid + 1;
        else hi = mid;
    }
    start[d] = lo;
}
__global__ __launch_bounds__(kBuildBlock) void k_refit_level(int level, const int* __restrict__ start,
                                                             const int* __restrict__ order, const int2* __restrict__ child,
                                                             float* __restrict__ ibox, const float* __restrict__ lbox) {
    const int b = start[level], e = start[level + 1];
    for (int t = b + blockIdx.x * blockDim.x + threadIdx.x; t < e; t += gridDim.x * blockDim.x) {
        const int p = order[t];
        const int2 ch = child[p];
        const float* A = ch.x >= 0 ? ibox + 6 * (size_t)ch.x : lbox + 6 * (size_t)(~ch.x);
        const float* B = ch.y >= 0 ? ibox + 6 * (size_t)ch.y : lbox + 6 * (size_t)(~ch.y);
        float* o = ibox + 6 * (size_t)p;
        for (int k = 0; k < 3; ++k) {
            o[k] = fminf(A[k], B[k]);
            o[3 + k] = fmaxf(A[3 + k], B[3 + k]);
        }
    }
}

// Depth of every leaf (number of internal ancestors); the maximum bounds the
// traversal stack.
__global__ __launch_bounds__(kBuildBlock) void k_depth(int n, const int* __restrict__ parent, uint32_t* __restrict__ depthOut) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t d = 0;
    for (int p = n > 1 ? parent[n - 1 + j] : -1; p >= 0 && d < 1000; p = parent[p]) ++d;
    atomicMax(depthOut, d);
}

__global__ __launch_bounds__(kBuildBlock) void k_emit(int n, const int2* __restrict__ child, const float* __restrict__ ibox,
                                                      const float* __restrict__ lbox, BvhNode* __restrict__ nodes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int2 ch = child[i];
    const float* A = ch.x >= 0 ? ibox + 6 * (size_t)ch.x : lbox + 6 * (size_t)(~ch.x);
    const float* B = ch.y >= 0 ? ibox + 6 * (size_t)ch.y : lbox + 6 * (size_t)(~ch.y);
    BvhNode nd;
    nd.lminx = A[0]; nd.lmaxx = A[3]; nd.lminy = A[1]; nd.lmaxy = A[4];
    nd.rminx = B[0]; nd.rmaxx = B[3]; nd.rminy = B[1]; nd.rmaxy = B[4];
    nd.lminz = A[2]; nd.lmaxz = A[5]; nd.rminz = B[2]; nd.rmaxz = B[5];
    nd.left = ch.x; nd.right = ch.y;
    nd.lcnt = ch.x < 0 ? 1u : 0u;
    nd.rcnt = ch.y < 0 ? 1u : 0u;
    nodes[i] = nd;
}

__device__ __forceinline__ float half_area(const float* b) {
    const float x = b[3] - b[0], y = b[4] - b[1], z = b[5] - b[2];
    return x * y + y * z + z * x;
}

// Biased exponent of the smallest power-of-two step s (s >= 2^-126) with
// lo + 255 s >= hi in decode arithmetic.
__device__ __forceinline__ uint32_t qnode_exponent(float lo, float hi) {
    const float ext = hi - lo;
    if (!(ext > 0.0f)) return 1u;
    uint32_t e = (__float_as_uint(ext * (1.0f / 255.0f)) >> 23) & 0xFFu;
    if (e == 0u) e = 1u;
    while (e < 254u && qnode_decode(lo, 255u, qnode_step(e)) < hi) ++e;
    return e;
}
// conservative 8-bit bounds of [clo, chi] on the grid lo + q s
__device__ __forceinline__ void qnode_bounds(float lo, float s, float clo, float chi, uint32_t& qlo, uint32_t& qhi) {
    qlo = (uint32_t)fminf(fmaxf(floorf((clo - lo) / s), 0.0f), 255.0f);
    while (qlo > 0u && qnode_decode(lo, qlo, s) > clo) --qlo;
    qhi = (uint32_t)fminf(fmaxf(ceilf((chi - lo) / s), 0.0f), 255.0f);
    while (qhi < 255u && qnode_decode(lo, qhi, s) < chi) ++qhi;
}

// Quantised copy of a four-wide node (empty slots get zero bytes; their
// child reference masks them in traversal).
__device__ __forceinline__ Bvh4QNode quantise_node(const Bvh4Node& nd) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 4; ++k) {
        if (nd.child[k] == kBvh4Empty) continue;
        b[0] = fminf(b[0], nd.lox[k]); b[1] = fminf(b[1], nd.loy[k]); b[2] = fminf(b[2], nd.loz[k]);
        b[3] = fmaxf(b[3], nd.hix[k]); b[4] = fmaxf(b[4], nd.hiy[k]); b[5] = fmaxf(b[5], nd.hiz[k]);
    }
    Bvh4QNode q;
    q.ox = b[0]; q.oy = b[1]; q.oz = b[2];
    const uint32_t ex = qnode_exponent(b[0], b[3]), ey = qnode_exponent(b[1], b[4]), ez = qnode_exponent(b[2], b[5]);
    q.exps = ex | (ey << 8) | (ez << 16);
    const float sx = qnode_step(ex), sy = qnode_step(ey), sz = qnode_step(ez);
    q.lox = q.loy = q.loz = q.hix = q.hiy = q.hiz = 0u;
    for (int k = 0; k < 4; ++k) {
        q.child[k] = nd.child[k];
        if (nd.child[k] == kBvh4Empty) continue;
        uint32_t l, h;
        qnode_bounds(b[0], sx, nd.lox[k], nd.hix[k], l, h);
        q.lox |= l << (8 * k); q.hix |= h << (8 * k);
        qnode_bounds(b[1], sy, nd.loy[k], nd.hiy[k], l, h);
        q.loy |= l << (8 * k); q.hiy |= h << (8 * k);
        qnode_bounds(b[2], sz, nd.loz[k], nd.hiz[k], l, h);
        q.loz |= l << (8 * k); q.hiz |= h << (8 * k);
    }
    q.pad[0] = q.pad[1] = 0u;
    return q;
}

// One collapse level: in[] holds (binary node, four-wide index) pairs; each
// produces its four-wide node and appends its inner children to out[].
__global__ __launch_bounds__(kBuildBlock) void k_collapse(const int2* __restrict__ child, const float* __restrict__ ibox,
                                                          const float* __restrict__ lbox, const int2* __restrict__ in,
                                                          const unsigned int* __restrict__ inCount, int2* __restrict__ out,
                                                          unsigned int* __restrict__ outCount,
                                                          unsigned int* __restrict__ nodeCount, Bvh4Node* __restrict__ nodes4,
                                                          Bvh4QNode* __restrict__ nodes4q) {
    const unsigned int cnt = *inCount;
    for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x) {
        const int2 w = in[t];
        int slot[4];
        int ns = 2;
        slot[0] = child[w.x].x;
        slot[1] = child[w.x].y;
        while (ns < 4) {
            int best = -1;
            float area = -1.0f;
            for (int k = 0; k < ns; ++k)
                if (slot[k] >= 0) {
                    const float a = half_area(ibox + 6 * (size_t)slot[k]);
                    if (a > area) { area = a; best = k; }
                }
            if (best < 0) break;
            const int2 ch = child[slot[best]];
            slot[best] = ch.x;
            slot[ns++] = ch.y;
        }
        Bvh4Node nd;
        for (int k = 0; k < 4; ++k) {
            if (k >= ns) {
                nd.lox[k] = nd.loy[k] = nd.loz[k] = INFINITY;
                nd.hix[k] = nd.hiy[k] = nd.hiz[k] = -INFINITY;
                nd.child[k] = kBvh4Empty;
                continue;
            }
            const int c = slot[k];
            const float* b = c >= 0 ? ibox + 6 * (size_t)c : lbox + 6 * (size_t)(~c);
            nd.lox[k] = b[0]; nd.loy[k] = b[1]; nd.loz[k] = b[2];
            nd.hix[k] = b[3]; nd.hiy[k] = b[4]; nd.hiz[k] = b[5];
            if (c >= 0) {
                const int id = (int)atomicAdd(nodeCount, 1u);
                nd.child[k] = id;
                out[atomicAdd(outCount, 1u)] = make_int2(c, id);
            } else {
                nd.child[k] = c;
            }
        }
        nd.pad[0] = nd.pad[1] = nd.pad[2] = nd.pad[3] = 0;
        nodes4[w.y] = nd;
        nodes4q[w.y] = quantise_node(nd);
    }
}

__global__ void k_collapse_init(int2* __restrict__ list, unsigned int* __restrict__ counts) {
    if (threadIdx.x == 0) {
        list[0] = make_int2(0, 0);   // the binary root becomes four-wide node 0
        counts[0] = 1u;              // level list A
        counts[1] = 0u;              // level list B
        counts[2] = 1u;              // four-wide nodes allocated
    }
}

__global__ void k_reset_count(unsigned int* __restrict__ c) {
    if (threadIdx.x == 0) *c = 0u;
}


inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// p[i * stride + a] -= delta[a] for npoints points (stride 3: packed float3;
// stride 0: the 9 corner floats of 68-byte emissive records, 3 points each)
__global__ __launch_bounds__(kBuildBlock) void k_translate(float* __restrict__ p, size_t npoints, uint32_t stride,
                                                           float dx, float dy, float dz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoints) return;
    float* q = stride ? p + i * stride : p + (i / 3) * 17 + (i % 3) * 3;
    q[0] -= dx;
    q[1] -= dy;
    q[2] -= dz;
}

// Scratch layout of a build over n triangles (kept after the build: the
// sorted keys, the binary topology and the depth order are what a refit
// reuses).
struct BuildLayout {
    size_t oBounds, oKeys, oKeys2, oChild, oParent, oIbox, oLbox, oDepth, oOrder, oLevels, oListA, oListB, oCounts,
        oSort, sortBytes, total;
};

hipError_t build_layout(rbh_ctx c, int n, BuildLayout& L) {
    const size_t nInt = n > 1 ? (size_t)n - 1 : 1;
    size_t off = 0;
    L.oBounds = off; off = align_up(off + 64);
    L.oKeys = off; off = align_up(off + (size_t)n * 8);
    L.oKeys2 = off; off = align_up(off + (size_t)n * 8);
    L.oChild = off; off = align_up(off + nInt * 8);
    L.oParent = off; off = align_up(off + ((size_t)n + nInt) * 4);
    L.oIbox = off; off = align_up(off + nInt * 24);
    L.oLbox = off; off = align_up(off + (size_t)n * 24);
    L.oDepth = off; off = align_up(off + nInt * 4);
    L.oOrder = off; off = align_up(off + nInt * 4);
    L.oLevels = off; off = align_up(off + 3 * (kLevels + 1) * 4);
    L.oListA = off; off = align_up(off + nInt * 8);
    L.oListB = off; off = align_up(off + nInt * 8);
    L.oCounts = off; off = align_up(off + 16);
    L.sortBytes = std::max(radix::sort_scratch_bytes(n, 8, 0), radix::sort_scratch_bytes(nInt, 4, 4));
    L.oSort = off; off = align_up(off + L.sortBytes);
    L.total = off;
    return hipSuccess;
}

int cu_count_of(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

// The build's last stages, which a refit repeats over the kept topology:
// leaf boxes and leaf-order triangle records from the current vertices,
// internal boxes level by level, the two-child nodes and the four-wide
// collapse.
int build_boxes_and_nodes(rbh_ctx c, const BuildLayout& L) {
    Geometry& g = c->geom;
    Accel& acc = c->accel;
    const int n = (int)g.ntris;
    const size_t nInt = n > 1 ? (size_t)n - 1 : 1;
    char* s = acc.scratch.as<char>();
    auto* keys2 = reinterpret_cast<unsigned long long*>(s + L.oKeys2);
    int2* child = reinterpret_cast<int2*>(s + L.oChild);
    float* ibox = reinterpret_cast<float*>(s + L.oIbox);
    float* lbox = reinterpret_cast<float*>(s + L.oLbox);
    int* order = reinterpret_cast<int*>(s + L.oOrder);
    int* lstart = reinterpret_cast<int*>(s + L.oLevels);
    const int nb = (n + kBuildBlock - 1) / kBuildBlock;
    const int nbi = (int)((nInt + kBuildBlock - 1) / kBuildBlock);
    const int cus = cu_count_of(c->device);
    hipLaunchKernelGGL(k_leaves, dim3(nb), dim3(kBuildBlock), 0, c->stream, g.verts.as<float>(), g.indices.as<uint32_t>(),
                       g.triMaterial.as<uint32_t>(), g.triLight.as<uint32_t>(), c->materials.as<Material>(),
                       c->nmaterials, keys2, n, acc.keyIndexBits, lbox, acc.tris.as<TriRecord>());
    if (n > 1) {
        const int lg = std::min(nbi, cus * 8);
        for (int d = kLevels - 1; d >= 0; --d)
            hipLaunchKernelGGL(k_refit_level, dim3(lg), dim3(kBuildBlock), 0, c->stream, d, lstart, order, child, ibox, lbox);
        hipLaunchKernelGGL(k_emit, dim3(nbi), dim3(kBuildBlock), 0, c->stream, n, child, ibox, lbox, acc.nodes.as<BvhNode>());
        int2* la = reinterpret_cast<int2*>(s + L.oListA);
        int2* lb = reinterpret_cast<int2*>(s + L.oListB);
        unsigned int* counts = reinterpret_cast<unsigned int*>(s + L.oCounts);
        hipLaunchKernelGGL(k_collapse_init, dim3(1), dim3(64), 0, c->stream, la, counts);
        // each level consumes one list and fills the other; every level
        // descends at least one binary level, so 64 levels drain the tree
        for (int lvl = 0; lvl < 64; ++lvl) {
            const bool even = (lvl & 1) == 0;
            unsigned int* inC = counts + (even ? 0 : 1);
            unsigned int* outC = counts + (even ? 1 : 0);
            hipLaunchKernelGGL(k_reset_count, dim3(1), dim3(64), 0, c->stream, outC);
            hipLaunchKernelGGL(k_collapse, dim3(std::min(nbi, cus * 8)), dim3(kBuildBlock), 0, c->stream, child, ibox,
                               lbox, even ? la : lb, inC, even ? lb : la, outC, counts + 2, acc.nodes4.as<Bvh4Node>(),
                               acc.nodes4q.as<Bvh4QNode>());
        }
    }
    RBH_HIP(c, hipGetLastError());
    return RBH_OK;
}

// Synchronises, times the stages since ev0 and checks the collapse.
int finish_build(rbh_ctx c, const BuildLayout& L, float* ms) {
    Accel& acc = c->accel;
    const int n = (int)c->geom.ntris;
    const size_t nInt = n > 1 ? (size_t)n - 1 : 1;
    char* s = acc.scratch.as<char>();
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    RBH_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    // counts: pending entries of both level lists, four-wide nodes allocated
    uint32_t cnt4[3] = {0u, 0u, 0u};
    if (n > 1) RBH_HIP(c, hipMemcpy(cnt4, s + L.oCounts, 12, hipMemcpyDeviceToHost));
    if (n > 1 && (cnt4[0] != 0u || cnt4[2] > nInt))
        return fail(c, RBH_ERR_RANGE, "four-wide collapse left %u nodes unwritten (%u allocated)", cnt4[0], cnt4[2]);
    acc.nnodes4 = cnt4[2];
    return RBH_OK;
}

int lbvh_build(rbh_ctx c) {
    Geometry& g = c->geom;
    Accel& acc = c->accel;
    const int n = (int)g.ntris;
    if (n <= 0) return fail(c, RBH_ERR_STATE, "no geometry uploaded");
    const size_t nInt = n > 1 ? (size_t)n - 1 : 1;
    BuildLayout L;
    RBH_HIP(c, build_layout(c, n, L));
    acc.built = false;
    RBH_ALLOC(c, acc.scratch, L.total);
    RBH_ALLOC(c, acc.nodes, nInt * sizeof(BvhNode));
    RBH_ALLOC(c, acc.tris, (size_t)n * sizeof(TriRecord));
    RBH_ALLOC(c, acc.nodes4, nInt * sizeof(Bvh4Node));
    RBH_ALLOC(c, acc.nodes4q, nInt * sizeof(Bvh4QNode));
    char* s = acc.scratch.as<char>();
    acc.keysOffset = L.oKeys2;
    uint32_t* bounds = reinterpret_cast<uint32_t*>(s + L.oBounds);
    auto* keys = reinterpret_cast<unsigned long long*>(s + L.oKeys);
    auto* keys2 = reinterpret_cast<unsigned long long*>(s + L.oKeys2);
    int2* child = reinterpret_cast<int2*>(s + L.oChild);
    int* parent = reinterpret_cast<int*>(s + L.oParent);
    int* ndepth = reinterpret_cast<int*>(s + L.oDepth);
    int* order = reinterpret_cast<int*>(s + L.oOrder);
    int* lstart = reinterpret_cast<int*>(s + L.oLevels);

    const float* v = g.verts.as<float>();
    const uint32_t* idx = g.indices.as<uint32_t>();
    const int nb = (n + kBuildBlock - 1) / kBuildBlock;
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, c->stream, bounds);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min(nb, 2048)), dim3(kBuildBlock), 0, c->stream, v, idx, (uint32_t)n, bounds);
    // key bits: the triangle index below, the Morton code above
    int idxBits = 1;
    while (idxBits < 32 && ((uint64_t)1 << idxBits) < (uint64_t)n) ++idxBits;
    const int axisBits = std::min(21, (64 - idxBits) / 3);
    acc.keyIndexBits = idxBits;
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kBuildBlock), 0, c->stream, v, idx, (uint32_t)n, bounds, axisBits,
                       idxBits, keys);
    RBH_HIP(c, (radix::sort<unsigned long long, uint32_t>(s + L.oSort, keys, keys2, nullptr, nullptr, (uint32_t)n, 0, 64,
                                                          c->stream)));
    const int nbi = (int)((nInt + kBuildBlock - 1) / kBuildBlock);
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(nbi), dim3(kBuildBlock), 0, c->stream, keys2, n, child, parent);
        // internal nodes in depth order (list buffers are free until the collapse)
        int2* ja = reinterpret_cast<int2*>(s + L.oListA);
        int2* jb = reinterpret_cast<int2*>(s + L.oListB);
        hipLaunchKernelGGL(k_jump_init, dim3(nbi), dim3(kBuildBlock), 0, c->stream, (int)nInt, parent, ja);
        for (int r = 0; r < kJumpRounds; ++r) {
            hipLaunchKernelGGL(k_jump, dim3(nbi), dim3(kBuildBlock), 0, c->stream, (int)nInt, ja, jb);
            std::swap(ja, jb);
        }
        unsigned int* dkey = reinterpret_cast<unsigned int*>(ndepth);
        int* ids = reinterpret_cast<int*>(jb);                       // free after the last round
        unsigned int* dsorted = reinterpret_cast<unsigned int*>(ids + nInt);
        hipLaunchKernelGGL(k_depth_keys, dim3(nbi), dim3(kBuildBlock), 0, c->stream, (int)nInt, ja, dkey, ids);
        RBH_HIP(c, (radix::sort<unsigned int, int>(s + L.oSort, dkey, dsorted, ids, order, (uint32_t)nInt, 0, 7, c->stream)));
        hipLaunchKernelGGL(k_level_bounds, dim3(1), dim3(128), 0, c->stream, dsorted, (int)nInt, lstart);
    }
    hipLaunchKernelGGL(k_depth, dim3(nb), dim3(kBuildBlock), 0, c->stream, n, parent, bounds + 6);
    const int rc = build_boxes_and_nodes(c, L);
    if (rc) return rc;
    float ms = 0.0f;
    const int fr = finish_build(c, L, &ms);
    if (fr) return fr;
    uint32_t depth = 0;
    RBH_HIP(c, hipMemcpy(&depth, bounds + 6, 4, hipMemcpyDeviceToHost));
    c->stats.bvhMs = ms;
    c->stats.bvhDepth = depth;
    acc.nnodes = n > 1 ? (uint32_t)(n - 1) : 0u;
    acc.depth = depth;
    acc.ntris = (uint32_t)n;
    if (depth > 64) return fail(c, RBH_ERR_RANGE, "LBVH depth %u exceeds the 64-entry traversal stack", depth);
    acc.built = true;
    acc.stale = false;
    return RBH_OK;
}

__global__ __launch_bounds__(kBuildBlock) void k_tri_classes(TriRecord* __restrict__ tris, int n,
                                                             const Material* __restrict__ mats, uint32_t nmats) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t m = tris[j].material;
    tris[j].shadeClass = m < nmats ? shade_class(mats[m]) : (uint32_t)SC_SURFACE;
}

// Re-derives the triangle records' shading classes after the materials
// changed (the classes only order the shade kernel's queue).
hipError_t refresh_tri_classes(rbh_ctx c) {
    const Accel& acc = c->accel;
    if (!acc.built || acc.ntris == 0) return hipSuccess;
    const int n = (int)acc.ntris;
    hipLaunchKernelGGL(k_tri_classes, dim3((n + kBuildBlock - 1) / kBuildBlock), dim3(kBuildBlock), 0, c->stream,
                       acc.tris.as<TriRecord>(), n, c->materials.as<Material>(), c->nmaterials);
    return hipGetLastError();
}

hipError_t translate_points(rbh_ctx c, float* p, size_t npoints, uint32_t strideFloats, const float* delta) {
    if (!npoints) return hipSuccess;
    const size_t blocks = (npoints + kBuildBlock - 1) / kBuildBlock;
    hipLaunchKernelGGL(k_translate, dim3((unsigned)blocks), dim3(kBuildBlock), 0, c->stream, p, npoints, strideFloats,
                       delta[0], delta[1], delta[2]);
    return hipGetLastError();
}

// Refit: new vertex positions under the kept topology (the triangles' leaf
// order, the binary hierarchy and its depth order stay; every box, the
// triangle records and the four-wide nodes are recomputed). The traversal
// stays exact for any positions; its cost grows as the positions drift from
// the ones the hierarchy was built for.
int lbvh_refit(rbh_ctx c) {
    Accel& acc = c->accel;
    const int n = (int)c->geom.ntris;
    if (!acc.built || acc.ntris != (uint32_t)n || n <= 0)
        return fail(c, RBH_ERR_STATE, "refit needs an acceleration structure built over the same triangles");
    BuildLayout L;
    RBH_HIP(c, build_layout(c, n, L));
    if (acc.scratch.bytes < L.total) return fail(c, RBH_ERR_STATE, "build scratch was released");
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    const int rc = build_boxes_and_nodes(c, L);
    if (rc) return rc;
    float ms = 0.0f;
    const int fr = finish_build(c, L, &ms);
    if (fr) return fr;
    c->stats.bvhMs = ms;
    acc.stale = false;
    return RBH_OK;
}

}  // namespace rbh
