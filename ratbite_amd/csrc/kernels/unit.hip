// Test-only kernels (include/ratbite_hip_test.h): each runs one device
// function of the hot path over a batch of inputs.
//
// BxDF record (64 x 32-bit words in, 16 out), the same as the CPU oracle's:
//  in  [0] material (u32) [1] path flags (u32) [2..4] ray direction
//      [5..13] shading frame columns T, B, N [14..17] light spectrum
//      [18..20] light direction [21] light pdf [22] occluded (u32)
//      [23] sample (u32) [24] dimension offset (u32) [25] pixel hash (u32)
//      [26..29] wavelengths [30..33] wavelength pdfs [34] layers (u32, bit 1 = sheen)
//      [35..37] base colour [38] metalness [39] ior [40] roughness
//      [41] transmission [42..44] sheen colour [45] sheen alpha
//      [46..49] throughput [50] refraction scale [51..54] L
//  out [0..3] L (+ unoccluded NEE term) [4..7] throughput [8] bxdf pdf
//      [9..11] new direction [12] flags (u32) [13] refraction scale
#include <cstring>
#include "../../../include/ratbite_hip_test.h"
#include "../context.hpp"
#include "../device/bxdf.hpp"
#include "../device/lighttree.hpp"
#include "../device/qrng.hpp"
#include "../device/traverse.hpp"
#include "radix.hpp"

namespace rbh {
using namespace dev;
namespace {

__global__ void k_sobol(const uint32_t* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* r = in + 5 * (size_t)i;
    const Qrng q{r[0], r[2], r[1]};
    float* o = out + 3 * (size_t)i;
    o[0] = o[1] = o[2] = 0.0f;
    if (r[4] == 1) {
        o[0] = q.s1(r[3]);
    } else if (r[4] == 2) {
        const V2 v = q.s2(r[3]);
        o[0] = v.x; o[1] = v.y;
    } else {
        const V3 v = q.s3(r[3]);
        o[0] = v.x; o[1] = v.y; o[2] = v.z;
    }
}

__global__ void k_hash(const uint32_t* xy, uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pixel_hash(xy[2 * i], xy[2 * i + 1]);
}

__global__ void k_wl(const float* u, float* lam, float* pdf, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Wl w = sample_wavelengths(u[i]);
    for (int k = 0; k < 4; ++k) { lam[4 * i + k] = w.lambda.v[k]; pdf[4 * i + k] = w.pdf.v[k]; }
}

__global__ void k_rgb(const float* colors, const float* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = in + 7 * (size_t)i;
    Wl w;
    for (int k = 0; k < 4; ++k) { w.lambda.v[k] = r[3 + k]; w.pdf.v[k] = 1.0f; }
    const S4 s = ColorTables{colors}.rgb(V3{r[0], r[1], r[2]}, w);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = s.v[k];
}

__global__ void k_oct(const float* v, uint32_t* enc, float* dec, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = oct_enc_u32(V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]});
    enc[i] = e;
    const V3 d = oct_dec_u32(e);
    dec[3 * i] = d.x; dec[3 * i + 1] = d.y; dec[3 * i + 2] = d.z;
}

template <bool ANY, int LDS_DEPTH>
__global__ __launch_bounds__(128) void k_trace_test(const BvhNode* nodes, const Bvh4Node* nodes4,
                                                    const Bvh4QNode* nodes4q, const TriRecord* tris, uint32_t nnodes,
                                                    uint32_t nnodes4, int tree, const float* rays, float* hits,
                                                    uint32_t n, int* spill, uint32_t spillStride) {
    __shared__ int stack[LDS_DEPTH * 128];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TravStack<LDS_DEPTH> st{stack + threadIdx.x, 128, spill + i, spillStride};
    const float* r = rays + 8 * (size_t)i;
    const V3 ro{r[0], r[1], r[2]}, rd{r[4], r[5], r[6]};
    const Hit h = tree == 1 ? traverse<ANY>(nodes, tris, nnodes, ro, rd, r[7], st)
                : tree == 2 ? traverse4<ANY>(nodes4q, tris, nnodes4, ro, rd, r[7], st)
                            : traverse4<ANY>(nodes4, tris, nnodes4, ro, rd, r[7], st);
    float* o = hits + 4 * (size_t)i;
    o[0] = h.leaf == 0xFFFFFFFFu ? -1.0f : h.t;
    o[1] = __uint_as_float(h.leaf == 0xFFFFFFFFu ? 0xFFFFFFFFu : tris[h.leaf].prim);
    o[2] = h.b1;
    o[3] = h.b2;
}

__global__ void k_lpdf(LightTree t, const uint32_t* lights, const float* pn, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = pn + 6 * (size_t)i;
    out[i] = t.pdf(LT_TRIANGLE, lights[i], V3{r[0], r[1], r[2]}, V3{r[3], r[4], r[5]});
}

__global__ void k_lsample(LightTree t, const float* in, uint32_t* ti, float* pdf, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = in + 7 * (size_t)i;
    uint32_t type = 0, idx = 0;
    float p = 0.0f;
    t.sample(r[0], V3{r[1], r[2], r[3]}, V3{r[4], r[5], r[6]}, type, idx, p);
    ti[2 * i] = type;
    ti[2 * i + 1] = idx;
    pdf[i] = p;
}

__global__ __launch_bounds__(64) void k_bxdf(const Material* mats, const float* spectra, const float* colors,
                                             const float* luts, const float* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = in + 64 * (size_t)i;
    auto U = [&](int k) { return __float_as_uint(r[k]); };
    const Material& m = mats[U(0)];
    PathIO P;
    P.flags = U(1);
    P.rd = V3{r[2], r[3], r[4]};
    const Frame fr{V3{r[5], r[6], r[7]}, V3{r[8], r[9], r[10]}, V3{r[11], r[12], r[13]}};
    Light dl;
    for (int k = 0; k < 4; ++k) dl.spectrum.v[k] = r[14 + k];
    dl.dir = V3{r[18], r[19], r[20]};
    dl.pdf = r[21];
    dl.occluded = U(22) != 0;
    const Qrng q{U(23), U(24), U(25)};
    for (int k = 0; k < 4; ++k) { P.wl.lambda.v[k] = r[26 + k]; P.wl.pdf.v[k] = r[30 + k]; }
    Surface s;
    s.sheen = (U(34) & 2u) != 0;
    s.base = V3{r[35], r[36], r[37]};
    s.metalness = r[38]; s.ior = r[39]; s.roughness = r[40]; s.transmission = r[41];
    s.sheenColor = V3{r[42], r[43], r[44]};
    s.sheenAlpha = r[45];
    S4 L;
    for (int k = 0; k < 4; ++k) { P.T.v[k] = r[46 + k]; L.v[k] = r[51 + k]; }
    P.refr = r[50];
    P.bxdfPdf = 1.0f;
    S4 Lnee = s4(0.0f);
    if (m.bxdf == 0)
        conductor_bxdf(spectra + (size_t)m.iorSpectrum * kDense, spectra + (size_t)m.acSpectrum * kDense, m.mfRoughness,
                       dl, q, fr, P, Lnee);
    else if (m.bxdf == 2)
        complex_bxdf(ColorTables{colors}, Luts{luts}, dl, q, fr, s, P, Lnee);
    float* o = out + 16 * (size_t)i;
    for (int k = 0; k < 4; ++k) { o[k] = L.v[k] + Lnee.v[k]; o[4 + k] = P.T.v[k]; }
    o[8] = P.bxdfPdf;
    o[9] = P.rd.x; o[10] = P.rd.y; o[11] = P.rd.z;
    o[12] = __uint_as_float(P.flags);
    o[13] = P.refr;
    o[14] = 0.0f;
    o[15] = 0.0f;
}

// Host helper: copies inputs in, runs, copies outputs back.
struct Staging {
    rbh_ctx c;
    std::vector<DevBuf*> bufs;
    ~Staging() { for (auto* b : bufs) delete b; }
    void* in(const void* src, size_t bytes) {
        DevBuf* b = new DevBuf();
        bufs.push_back(b);
        if (b->reserve(bytes ? bytes : 4) != hipSuccess) return nullptr;
        if (bytes && hipMemcpy(b->p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return b->p;
    }
    void* out(size_t bytes) {
        DevBuf* b = new DevBuf();
        bufs.push_back(b);
        if (b->reserve(bytes ? bytes : 4) != hipSuccess) return nullptr;
        return b->p;
    }
};

inline dim3 grid(uint32_t n, int b) { return dim3(n ? (n + b - 1) / b : 1u); }

struct PositiveAt {
    const float* w;
    __device__ __forceinline__ bool operator()(uint32_t i) const { return w[i] > 0.0f; }
};

// plain streaming copy, one 16-byte element per lane and no loop (a grid of
// n / 256 blocks): the reference point for the HBM byte counters and the
// measured bandwidth ceiling. On MI355X this form moves 6.3 TB/s read+write at
// 1 and 4 GiB, against 4.7-5.1 TB/s for grid-stride loops of 4 to 64 blocks
// per CU with four loads in flight per lane.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace
}  // namespace rbh

using namespace rbh;

#define STAGE_CHECK(p) \
    do { if (!(p)) return fail(c, RBH_ERR_NOMEM, "test staging allocation failed"); } while (0)
#define FINISH(outp, dptr, bytes)                                                          \
    do {                                                                                   \
        RBH_HIP(c, hipGetLastError());                                                     \
        RBH_HIP(c, hipDeviceSynchronize());                                                \
        RBH_HIP(c, hipMemcpy((outp), (dptr), (bytes), hipMemcpyDeviceToHost));             \
    } while (0)

extern "C" {

int rbh_test_sobol(rbh_ctx c, const uint32_t* in, float* out, uint32_t n) {
    if (!c || !in || !out) return RBH_ERR_INVALID;
    Staging s{c};
    void* din = s.in(in, (size_t)n * 20);
    void* dout = s.out((size_t)n * 12);
    STAGE_CHECK(din && dout);
    hipLaunchKernelGGL(k_sobol, grid(n, 256), dim3(256), 0, 0, (const uint32_t*)din, (float*)dout, n);
    FINISH(out, dout, (size_t)n * 12);
    return RBH_OK;
}

int rbh_test_pixel_hash(rbh_ctx c, const uint32_t* xy, uint32_t* out, uint32_t n) {
    if (!c || !xy || !out) return RBH_ERR_INVALID;
    Staging s{c};
    void* din = s.in(xy, (size_t)n * 8);
    void* dout = s.out((size_t)n * 4);
    STAGE_CHECK(din && dout);
    hipLaunchKernelGGL(k_hash, grid(n, 256), dim3(256), 0, 0, (const uint32_t*)din, (uint32_t*)dout, n);
    FINISH(out, dout, (size_t)n * 4);
    return RBH_OK;
}

int rbh_test_wavelengths(rbh_ctx c, const float* u, float* lam, float* pdf, uint32_t n) {
    if (!c || !u || !lam || !pdf) return RBH_ERR_INVALID;
    Staging s{c};
    void* du = s.in(u, (size_t)n * 4);
    void* dl = s.out((size_t)n * 16);
    void* dp = s.out((size_t)n * 16);
    STAGE_CHECK(du && dl && dp);
    hipLaunchKernelGGL(k_wl, grid(n, 256), dim3(256), 0, 0, (const float*)du, (float*)dl, (float*)dp, n);
    FINISH(lam, dl, (size_t)n * 16);
    RBH_HIP(c, hipMemcpy(pdf, dp, (size_t)n * 16, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_test_rgb_to_spectrum(rbh_ctx c, const float* in, float* out, uint32_t n) {
    if (!c || !in || !out) return RBH_ERR_INVALID;
    if (!c->haveColor) return fail(c, RBH_ERR_STATE, "colour tables not uploaded");
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) {
            const float l = in[7 * i + 3 + k];
            if (!(l >= 359.5f && l < 830.5f)) return fail(c, RBH_ERR_INVALID, "wavelength %g outside [360, 830]", l);
        }
    Staging s{c};
    void* din = s.in(in, (size_t)n * 28);
    void* dout = s.out((size_t)n * 16);
    STAGE_CHECK(din && dout);
    hipLaunchKernelGGL(k_rgb, grid(n, 256), dim3(256), 0, 0, c->colorTables.as<float>(), (const float*)din, (float*)dout, n);
    FINISH(out, dout, (size_t)n * 16);
    return RBH_OK;
}

int rbh_test_oct(rbh_ctx c, const float* v, uint32_t* enc, float* dec, uint32_t n) {
    if (!c || !v || !enc || !dec) return RBH_ERR_INVALID;
    Staging s{c};
    void* dv = s.in(v, (size_t)n * 12);
    void* de = s.out((size_t)n * 4);
    void* dd = s.out((size_t)n * 12);
    STAGE_CHECK(dv && de && dd);
    hipLaunchKernelGGL(k_oct, grid(n, 256), dim3(256), 0, 0, (const float*)dv, (uint32_t*)de, (float*)dd, n);
    FINISH(enc, de, (size_t)n * 4);
    RBH_HIP(c, hipMemcpy(dec, dd, (size_t)n * 12, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_test_trace(rbh_ctx c, const float* rays, float* hits, uint32_t n, int mode) {
    if (!c || !rays || !hits) return RBH_ERR_INVALID;
    if (!c->accel.built) return fail(c, RBH_ERR_STATE, "no acceleration structure");
    Staging s{c};
    void* dr = s.in(rays, (size_t)n * 32);
    void* dh = s.out((size_t)n * 16);
    const dim3 g = grid(n, 128);
    const uint32_t lanes = g.x * 128u;
    int* spill = (int*)s.out((size_t)lanes * kMaxStack4 * 4);
    STAGE_CHECK(dr && dh && spill);
    const BvhNode* nodes = c->accel.nodes.as<BvhNode>();
    const Bvh4Node* nodes4 = c->accel.nodes4.as<Bvh4Node>();
    const Bvh4QNode* nodes4q = c->accel.nodes4q.as<Bvh4QNode>();
    const TriRecord* tris = c->accel.tris.as<TriRecord>();
    const uint32_t nn = c->accel.nnodes, nn4 = c->accel.nnodes4;
    const bool any = (mode & 1) != 0, tiny = (mode & 2) != 0;
    const int binary = (mode & 4) ? 1 : ((mode & 8) ? 2 : 0);
    // mode bit 1: a 2-entry LDS stack, so the spill columns carry the rest;
    // bit 2: the binary tree instead of the four-wide one; bit 3: the
    // four-wide nodes quantised to 64 bytes
    if (any && tiny)
        hipLaunchKernelGGL((k_trace_test<true, 2>), g, dim3(128), 0, 0, nodes, nodes4, nodes4q, tris, nn, nn4, binary, (const float*)dr, (float*)dh, n, spill, lanes);
    else if (any)
        hipLaunchKernelGGL((k_trace_test<true, 16>), g, dim3(128), 0, 0, nodes, nodes4, nodes4q, tris, nn, nn4, binary, (const float*)dr, (float*)dh, n, spill, lanes);
    else if (tiny)
        hipLaunchKernelGGL((k_trace_test<false, 2>), g, dim3(128), 0, 0, nodes, nodes4, nodes4q, tris, nn, nn4, binary, (const float*)dr, (float*)dh, n, spill, lanes);
    else
        hipLaunchKernelGGL((k_trace_test<false, 16>), g, dim3(128), 0, 0, nodes, nodes4, nodes4q, tris, nn, nn4, binary, (const float*)dr, (float*)dh, n, spill, lanes);
    FINISH(hits, dh, (size_t)n * 16);
    return RBH_OK;
}

int rbh_test_copy(rbh_ctx c, uint64_t bytes, int repeats, float* ms) {
    if (!c || bytes < 16 || repeats < 1) return RBH_ERR_INVALID;
    if (bytes > (uint64_t(1) << 34)) return fail(c, RBH_ERR_RANGE, "copy size above 16 GiB");
    DevBuf a, b;
    RBH_ALLOC(c, a, bytes);
    RBH_ALLOC(c, b, bytes);
    RBH_HIP(c, hipMemsetAsync(a.p, 1, bytes, c->stream));
    const size_t n = bytes / 16;
    // one untimed launch first (first-touch page mapping)
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a.as<float4>(),
                       b.as<float4>(), n);
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a.as<float4>(),
                           b.as<float4>(), n);
    RBH_HIP(c, hipGetLastError());
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    float t = 0.0f;
    RBH_HIP(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    if (ms) *ms = t / repeats;
    return RBH_OK;
}

int rbh_test_radix_sort(rbh_ctx c, const uint64_t* keys, const uint32_t* vals, uint32_t n, int beginBit, int endBit,
                        uint64_t* outKeys, uint32_t* outVals) {
    if (!c || !keys || !outKeys || (vals && !outVals)) return RBH_ERR_INVALID;
    if (beginBit < 0 || endBit > 64 || beginBit > endBit) return fail(c, RBH_ERR_INVALID, "bad bit range");
    Staging s{c};
    const uint64_t* dk = static_cast<const uint64_t*>(s.in(keys, (size_t)n * 8));
    const uint32_t* dv = vals ? static_cast<const uint32_t*>(s.in(vals, (size_t)n * 4)) : nullptr;
    uint64_t* dko = static_cast<uint64_t*>(s.out((size_t)n * 8));
    uint32_t* dvo = vals ? static_cast<uint32_t*>(s.out((size_t)n * 4)) : nullptr;
    void* scratch = s.out(radix::sort_scratch_bytes(n, 8, vals ? 4 : 0));
    STAGE_CHECK(dk && dko && scratch && (!vals || (dv && dvo)));
    RBH_HIP(c, (radix::sort<uint64_t, uint32_t>(scratch, dk, dko, dv, dvo, n, beginBit, endBit, 0)));
    FINISH(outKeys, dko, (size_t)n * 8);
    if (vals) RBH_HIP(c, hipMemcpy(outVals, dvo, (size_t)n * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_test_select_positive(rbh_ctx c, const float* weights, uint32_t n, uint32_t* out, uint32_t* count) {
    if (!c || !weights || !out || !count) return RBH_ERR_INVALID;
    Staging s{c};
    const float* dw = static_cast<const float*>(s.in(weights, (size_t)n * 4));
    uint32_t* dout = static_cast<uint32_t*>(s.out((size_t)n * 4 + 4));
    void* scratch = s.out(radix::select_scratch_bytes(n));
    STAGE_CHECK(dw && dout && scratch);
    RBH_HIP(c, radix::select_indices(scratch, n, PositiveAt{dw}, dout, dout + n, 0));
    FINISH(count, dout + n, 4);
    if (*count > n) return fail(c, RBH_ERR_INVALID, "selection count %u above %u", *count, n);
    RBH_HIP(c, hipMemcpy(out, dout, (size_t)*count * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_test_lbvh_read(rbh_ctx c, uint64_t* keys, void* nodes, void* tris, uint32_t* nnodes) {
    if (!c || !nnodes) return RBH_ERR_INVALID;
    if (!c->accel.built) return fail(c, RBH_ERR_STATE, "no acceleration structure");
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    const size_t n = c->geom.ntris;
    *nnodes = c->accel.nnodes;
    if (keys) RBH_HIP(c, hipMemcpy(keys, c->accel.scratch.as<char>() + c->accel.keysOffset, n * 8, hipMemcpyDeviceToHost));
    if (nodes && c->accel.nnodes)
        RBH_HIP(c, hipMemcpy(nodes, c->accel.nodes.p, (size_t)c->accel.nnodes * sizeof(BvhNode), hipMemcpyDeviceToHost));
    if (tris) RBH_HIP(c, hipMemcpy(tris, c->accel.tris.p, n * sizeof(TriRecord), hipMemcpyDeviceToHost));
    return RBH_OK;
}

static LightTree tree_of(rbh_ctx c) {
    return LightTree{c->tree.nlights ? c->tree.nodesX.as<TreeNodeX>() : nullptr, c->tree.lightPointers.as<uint32_t>(),
                     c->tree.bitmasks.as<uint64_t>(), c->env.enabled};
}

int rbh_test_light_pdf(rbh_ctx c, const uint32_t* lights, const float* pn, float* pdf, uint32_t n) {
    if (!c || !lights || !pn || !pdf) return RBH_ERR_INVALID;
    if (!c->tree.nlights) return fail(c, RBH_ERR_STATE, "no light tree");
    for (uint32_t i = 0; i < n; ++i)
        if (lights[i] >= c->tree.nlights) return fail(c, RBH_ERR_INVALID, "light %u out of range", lights[i]);
    Staging s{c};
    void* dl = s.in(lights, (size_t)n * 4);
    void* dp = s.in(pn, (size_t)n * 24);
    void* dout = s.out((size_t)n * 4);
    STAGE_CHECK(dl && dp && dout);
    hipLaunchKernelGGL(k_lpdf, grid(n, 128), dim3(128), 0, 0, tree_of(c), (const uint32_t*)dl, (const float*)dp, (float*)dout, n);
    FINISH(pdf, dout, (size_t)n * 4);
    return RBH_OK;
}

int rbh_test_light_sample(rbh_ctx c, const float* in, uint32_t* ti, float* pdf, uint32_t n) {
    if (!c || !in || !ti || !pdf) return RBH_ERR_INVALID;
    if (!c->tree.nlights && !c->env.enabled) return fail(c, RBH_ERR_STATE, "no lights");
    Staging s{c};
    void* din = s.in(in, (size_t)n * 28);
    void* dt = s.out((size_t)n * 8);
    void* dp = s.out((size_t)n * 4);
    STAGE_CHECK(din && dt && dp);
    hipLaunchKernelGGL(k_lsample, grid(n, 128), dim3(128), 0, 0, tree_of(c), (const float*)din, (uint32_t*)dt, (float*)dp, n);
    FINISH(ti, dt, (size_t)n * 8);
    RBH_HIP(c, hipMemcpy(pdf, dp, (size_t)n * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_test_bxdf(rbh_ctx c, const float* in, float* out, uint32_t n) {
    if (!c || !in || !out) return RBH_ERR_INVALID;
    if (!c->haveColor || !c->haveLuts || !c->nmaterials) return fail(c, RBH_ERR_STATE, "tables/materials missing");
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t m;
        std::memcpy(&m, in + 64 * (size_t)i, 4);
        if (m >= c->nmaterials) return fail(c, RBH_ERR_INVALID, "record %u: material %u out of range", i, m);
    }
    Staging s{c};
    void* din = s.in(in, (size_t)n * 256);
    void* dout = s.out((size_t)n * 64);
    STAGE_CHECK(din && dout);
    hipLaunchKernelGGL(k_bxdf, grid(n, 64), dim3(64), 0, 0, c->materials.as<Material>(), c->spectra.as<float>(),
                       c->colorTables.as<float>(), c->luts.as<float>(), (const float*)din, (float*)dout, n);
    FINISH(out, dout, (size_t)n * 64);
    return RBH_OK;
}

}  // extern "C"
