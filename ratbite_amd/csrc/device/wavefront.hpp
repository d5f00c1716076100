// Per-path bodies of the wavefront integrator kernels (camera set-up, sky
// hits, surface shading, the NEE term of an unoccluded shadow ray, FP64
// accumulation), restating the OptiX megakernel __raygen__main
// (src/kernels/pathtrace.cu:1114-1813) one path slot at a time. The kernels
// in kernels/integrator.hip wrap them in persistent grids with wave-level
// queue compaction; every function here reads and writes only the slots it
// is given, so the same code also runs in a host build of the device
// headers.
#pragma once
#include "../guide.hpp"
#include "../integrator.hpp"
#include "bxdf.hpp"
#include "guided.hpp"
#include "lighttree.hpp"
#include "qrng.hpp"
#include "shapes.hpp"

namespace rbh {
namespace wf {

using namespace dev;

// The scene owning a path slot: the pass's only scene, or (batched passes)
// the last view whose slot range starts at or before the slot.
template <bool BATCH>
__device__ __forceinline__ const SceneView& scene_of(const Frame1& f, uint32_t slot) {
    if (!BATCH) return f.sc;
    uint32_t lo = 0, hi = f.nscenes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (f.scenes[mid].slot0 <= slot) lo = mid;
        else hi = mid;
    }
    return f.scenes[lo];
}

// Pixel of a scene-local slot index: bands of 8 rows, each band in blocks of
// 8 columns (the last block of a band as wide as what is left), a block row
// by row. A 64-lane wave of consecutive slots then covers an 8x8 tile
// instead of a 64x1 strip, so its rays share BVH nodes and textures. A
// bijection of [0, W*H) for any film size.
__device__ __forceinline__ void slot_pixel(const SceneView& S, uint32_t pix, uint32_t& x, uint32_t& y) {
    const uint32_t band = pix / (8u * S.W);
    const uint32_t y0 = band * 8u;
    const uint32_t h = min(8u, S.H - y0);
    const uint32_t r = pix - band * 8u * S.W;     // < h * W
    const uint32_t full = S.W / 8u;                // blocks 8 columns wide
    const uint32_t xb = r / (8u * h);
    if (xb < full) {
        const uint32_t w = r - xb * 8u * h;
        y = y0 + w / 8u;
        x = xb * 8u + w % 8u;
    } else {
        const uint32_t wl = S.W - full * 8u;      // width of the band's last block (> 0 here)
        const uint32_t w = r - full * 8u * h;
        y = y0 + w / wl;
        x = full * 8u + w % wl;
    }
}

// ------------------------------------------------------------------ camera
// The sample's hero wavelengths (spectral.h:227-241 from the pixel's
// wavelength dimension at the first bounce). They never change along the
// path, so every kernel recomputes them from the slot instead of keeping
// 32 bytes per path slot in HBM.
__device__ __forceinline__ Wl slot_wavelengths(const SceneView& S, uint32_t sk, uint32_t x, uint32_t y) {
    return sample_wavelengths(qrng_at(S.sample + sk, 0, pixel_hash(x, y)).s1(QD_WAVELENGTH));
}
__device__ __forceinline__ Wl path_wavelengths(const SceneView& S, uint32_t pid) {
    const uint32_t local = pid - S.slot0;
    const uint32_t sk = local / S.P, pix = local - sk * S.P;
    uint32_t x, y;
    slot_pixel(S, pix, x, y);
    return slot_wavelengths(S, sk, x, y);
}

// Camera ray of path slot i (pathtrace.cu:1114-1160).
__device__ __forceinline__ void camera_ray(const SceneView& S, uint32_t i, V3& o, V3& d) {
    const uint32_t local = i - S.slot0;
    const uint32_t k = local / S.P, pix = local - k * S.P;
    uint32_t x, y;
    slot_pixel(S, pix, x, y);
    const Qrng q = qrng_at(S.sample + k, 0, pixel_hash(x, y));
    const V2 fo = gaussian_offset(q.s2(QD_FILTER));
    const float cx = (float)x + fo.x, cy = (float)y + fo.y;
    o = V3{0.0f, 0.0f, 0.0f};
    if (S.dof) {
        // src/device/dir_gen.h:358-374
        const V2 ls = q.s2(QD_LENS);
        const float xs = (2.0f * cx * S.invW - 1.0f) * S.psW;
        const float zs = (2.0f * cy * S.invH - 1.0f) * S.psH;
        const V3 focus = V3{xs, zs, 1.0f} * S.focus;
        const V2 dk = disk_polar(ls);
        const V3 lens{dk.x * S.aperture, dk.y * S.aperture, 0.0f};
        o = S.camU * lens.x + S.camW * lens.y;
        const V3 rd = nrm(focus - lens);
        d = S.camU * rd.x + S.camW * rd.y + S.camV * rd.z;
    } else {
        // src/device/dir_gen.h:346-357
        const float xs = 2.0f * cx * S.invW - 1.0f;
        const float zs = 2.0f * cy * S.invH - 1.0f;
        d = nrm(S.camV + S.camU * xs * S.psW + S.camW * zs * S.psH);
    }
}

// The state a path starts with. The first bounce's kernels use it instead of
// reading it: only the camera ray's origin and direction are stored for
// every slot (by the first closest-hit pass), the rest is first written by
// the kernel that shades the slot's first hit or miss.
__device__ __forceinline__ float4 init_L() { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
__device__ __forceinline__ float4 init_T() { return make_float4(1.0f, 1.0f, 1.0f, 1.0f); }
// last shading normal 0, no state flags
__device__ __forceinline__ float4 init_nf() { return make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u)); }

// Camera ray and the whole fresh path state of slot i in memory (the host
// build of these bodies starts its passes with it).
__device__ __forceinline__ void camera_slot(const Frame1& f, const SceneView& S, uint32_t i) {
    V3 o, d;
    camera_ray(S, i, o, d);
    const PathBuffers& p = f.p;
    p.od[i] = make_float4(o.x, o.y, o.z, 1.0f);           // bxdfPdf = 1
    p.dr[i] = make_float4(d.x, d.y, d.z, 1.0f);           // refractionScale = 1
    p.L[i] = init_L();
    p.T[i] = init_T();
    p.nf[i] = init_nf();
    p.depth[i] = 0u;
    if (f.collect) p.prec[i] = kNoRec;
}

// Guided renders that collect training data: the radiance of the path just
// before the vertex its last recorded direction leads to.
__device__ __forceinline__ void snapshot_radiance(const Frame1& f, uint32_t pid, float4 L) {
    const uint32_t pr = f.p.prec[pid];
    if (pr != kNoRec && (pr & kRecPending)) {
        const uint32_t r = pr & ~kRecPending;
        float* rec = f.grec + (size_t)r * kPendingRec;
        rec[20] = L.x;
        rec[21] = L.y;
        rec[22] = L.z;
        rec[23] = L.w;
        f.p.prec[pid] = r;
    }
}

// A path whose ray left the scene: sky emission with the MIS weight of the
// previous vertex's BxDF sample, then the path ends (pathtrace.cu:1185-1274,
// the sky branch; 1240-1252 for the lookup and its pdf).
__device__ __forceinline__ void shade_miss(const Frame1& f, const SceneView& S, uint32_t pid) {
    const bool first = f.bounce == 0u;
    float4 nf = first ? init_nf() : f.p.nf[pid];
    uint32_t flags = __float_as_uint(nf.w);
    flags &= ~PF_TRIANGULAR;
    if (f.collect) snapshot_radiance(f, pid, first ? init_L() : f.p.L[pid]);
    if (first && !S.envOn) f.p.L[pid] = init_L();
    if (S.envOn) {
        const uint32_t depth = first ? 0u : f.p.depth[pid];
        const float4 od = f.p.od[pid], dr = f.p.dr[pid];
        const V3 rd{dr.x, dr.y, dr.z};
        const LightTree tree{S.treeNodes, S.treePointers, S.treeMasks, true};
        const bool prevNEE = depth != 0 && !(flags & PF_PREV_SPECULAR);
        const float treePdf = prevNEE ? tree.pdf(LT_SKY, 0, V3{od.x, od.y, od.z}, V3{nf.x, nf.y, nf.z}) : 1.0f;
        const float phi = atan2f(rd.y, rd.x);
        const float theta = acosf(rd.z);
        const V3 sky = env_lookup(S.envTex, S.envW, S.envH, S.envC, 0.5f - phi / (2.0f * 3.14159265358979f),
                                  theta / 3.14159265358979f);
        float lightPdf = (1.0f / (4.0f * 3.14159265358979f)) * ((sky.x + sky.y + sky.z) / 3.0f) * sinf(theta) /
                         S.envIntegral;
        const Wl wl = path_wavelengths(S, pid);
        const S4 Le = ColorTables{S.colors}.rgb(sky, wl);
        lightPdf *= treePdf;
        const float w = prevNEE ? power_heuristic(od.w, lightPdf) : 1.0f;
        const float4 Tv = first ? init_T() : f.p.T[pid];
        float4 L = first ? init_L() : f.p.L[pid];
        L.x += Tv.x * Le.v[0] * w;
        L.y += Tv.y * Le.v[1] * w;
        L.z += Tv.z * Le.v[2] * w;
        L.w += Tv.w * Le.v[3] * w;
        f.p.L[pid] = L;
    }
    flags |= PF_FINISHED;
    nf.w = __uint_as_float(flags);
    f.p.nf[pid] = nf;
}

// Network input of a hit: scene-normalised position, geometric normal turned
// towards the viewer, outgoing direction, roughness factor, hero wavelength.
__device__ __forceinline__ void guide_input(const Frame1& f, const SceneView& S, uint32_t pid, float* __restrict__ out) {
    const float4 hr = f.p.hit[pid], dr = f.p.dr[pid];
    if (__float_as_uint(hr.w) & kShapeHit) {
        // a shape-light hit ends the path: no vertex to guide
        for (int k = 0; k < 16; ++k) out[k] = 0.0f;
        return;
    }
    const TriRecord& tr = S.tris[__float_as_uint(hr.w)];
    const V3 v0{tr.v0x, tr.v0y, tr.v0z}, v1{tr.v1x, tr.v1y, tr.v1z}, v2{tr.v2x, tr.v2y, tr.v2z};
    const float b1 = hr.y, b2 = hr.z, b0 = 1.0f - b1 - b2;
    const V3 hp{v0.x * b0 + v1.x * b1 + v2.x * b2, v0.y * b0 + v1.y * b1 + v2.y * b2, v0.z * b0 + v1.z * b1 + v2.z * b2};
    V3 g = cross(v1 - v0, v2 - v0);
    g = g * rsqrtf(dot(g, g));
    const V3 wo{-dr.x, -dr.y, -dr.z};
    if (dot(g, wo) < 0.0f) g = -g;
    const Material& m = S.materials[tr.material];
    const float rough = m.bxdf == 0 ? m.mfRoughness : ((m.factors & 4u) ? m.roughness : 1.0f);
    out[0] = clampf((hp.x - f.gmin[0]) * f.ginv[0], 0.0f, 1.0f);
    out[1] = clampf((hp.y - f.gmin[1]) * f.ginv[1], 0.0f, 1.0f);
    out[2] = clampf((hp.z - f.gmin[2]) * f.ginv[2], 0.0f, 1.0f);
    out[3] = g.x; out[4] = g.y; out[5] = g.z;
    out[6] = wo.x; out[7] = wo.y; out[8] = wo.z;
    out[9] = rough;
    out[10] = (path_wavelengths(S, pid).lambda.v[0] - 360.0f) * (1.0f / 470.0f);
    for (int k = 11; k < 16; ++k) out[k] = 0.0f;
}

// The deferred NEE term of a path whose shadow ray reached the light.
__device__ __forceinline__ void add_nee(const Frame1& f, uint32_t pid) {
    const float4 a = f.p.Lnee[pid];
    float4 L = f.p.L[pid];
    L.x += a.x; L.y += a.y; L.z += a.z; L.w += a.w;
    f.p.L[pid] = L;
}

__device__ __forceinline__ S4 tri_emission(const SceneView& S, const EmissiveTri& tri, float ba, float bb, float bc,
                                           const Wl& wl, const ColorTables& ct) {
    const Material& m = S.materials[tri.material];
    if (m.flags & 4u) return dense4(S.spectra + (size_t)m.emissionSpectrum * kDense, wl) * m.emissive[0];
    V3 e{m.emissive[0], m.emissive[1], m.emissive[2]};
    if (m.textures & 64u) {
        const float u = ba * tri.uv[0] + bb * tri.uv[2] + bc * tri.uv[4];
        const float v = ba * tri.uv[1] + bb * tri.uv[3] + bc * tri.uv[5];
        const float4 t = tex2d(S.texDescs, S.texels, m.tex[6], u, v, S.srgbLut);
        e = V3{e.x * t.x, e.y * t.y, e.z * t.z};
    }
    return ct.rgb(e, wl);
}

// Emission of a shape light: its material's emission (spectral, or RGB
// lifted to a spectrum; no texture) times the light's power scale.
__device__ __forceinline__ S4 shape_emission(const SceneView& S, const ShapeLight& sl, const Wl& wl,
                                             const ColorTables& ct) {
    const Material& m = S.materials[sl.material];
    const S4 e = (m.flags & 4u) ? dense4(S.spectra + (size_t)m.emissionSpectrum * kDense, wl) * m.emissive[0]
                                : ct.rgb(V3{m.emissive[0], m.emissive[1], m.emissive[2]}, wl);
    return e * sl.power;
}

// A path whose ray hit a shape light: its emission with the MIS weight of
// the previous vertex's BxDF sample (the light-tree and shape densities
// below are those the NEE sampler uses), then the path ends.
__device__ __forceinline__ void shade_shape_hit(const SceneView& S, const LightTree& tree, const ColorTables& ct,
                                                uint32_t si, float t, V3 ro, V3 lastN, uint32_t depth, PathIO& P,
                                                S4& L) {
    const ShapeLight& sl = S.shapes[si];
    const bool prevNEE = depth != 0 && !(P.flags & PF_PREV_SPECULAR);
    float lightPdf = 0.0f;
    bool noEm;
    if (sl.type == SHAPE_DISK) {
        const V3 th = P.rd * t;
        const float d2 = th.x * th.x + th.y * th.y + th.z * th.z;
        const float lcos = -dot(P.rd, shape_normal(sl));
        noEm = lcos <= 0.0f;
        lightPdf = (1.0f / (kShapePi * sl.radius * sl.radius)) * d2 / lcos;
    } else {
        const V3 oc = shape_center(sl) - ro;
        noEm = !(dot(oc, oc) > sl.radius * sl.radius);   // from inside: the inner face does not emit
        if (!noEm) lightPdf = sphere_cone_pdf(sl, ro);
    }
    if (!noEm) {
        const float treePdf = prevNEE ? tree.pdf(LT_SHAPE, S.shapeBase + si, ro, lastN) : 1.0f;
        lightPdf *= treePdf;
        const float w = prevNEE ? power_heuristic(P.bxdfPdf, lightPdf) : 1.0f;
        L += P.T * shape_emission(S, sl, P.wl, ct) * w;
    }
    P.flags |= PF_FINISHED;
}

__device__ __forceinline__ V2 attr2(const uint8_t* __restrict__ a, uint32_t stride, uint32_t off, uint32_t vi) {
    const float2 v = *reinterpret_cast<const float2*>(a + (size_t)stride * vi + off);
    return V2{v.x, v.y};
}

// glm::mat3_cast columns of a unit quaternion (x, y, z, w)
__device__ __forceinline__ void quat_cols(float4 q, V3& c0, V3& c2) {
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    c0 = V3{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y)};
    c2 = V3{2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)};
}

// Shades path slot pid, whose ray hit a triangle (hit-queue entry i):
// emission + MIS, surface attributes, light-tree NEE, BxDF, Russian
// roulette, next origin, depth caps (pathtrace.cu:1179-1785). Sets alive when
// the path goes on and shadow when it has a shadow ray to test; in guided
// renders that collect training data, recordSlot with the recorded direction,
// its BSDF-sampler density and the throughput before the vertex. The outputs
// must be initialised by the caller (false / zero).
template <bool GUIDED>
__device__ __forceinline__ void shade_hit(const Frame1& f, const SceneView& S, uint32_t pid, uint32_t i, bool& alive,
                                          bool& shadow, bool& recordSlot, V3& recDir, float& recPb, S4& recT) {
    const ColorTables ct{S.colors};
    const Luts lut{S.luts};
    const LightTree tree{S.treeNodes, S.treePointers, S.treeMasks, S.envOn != 0};
    const bool treeOn = S.treeNodes != nullptr;
    const bool first = f.bounce == 0u;
    const float4 od = f.p.od[pid], dr = f.p.dr[pid], nf = first ? init_nf() : f.p.nf[pid], hr = f.p.hit[pid];
    PathIO P;
    P.rd = V3{dr.x, dr.y, dr.z};
    P.refr = dr.w;
    P.bxdfPdf = od.w;
    P.flags = __float_as_uint(nf.w);
    float4 Lv = first ? init_L() : f.p.L[pid], Tv = first ? init_T() : f.p.T[pid];
    if (GUIDED && f.collect) snapshot_radiance(f, pid, Lv);
    S4 L{{Lv.x, Lv.y, Lv.z, Lv.w}};
    P.T = S4{{Tv.x, Tv.y, Tv.z, Tv.w}};
    V3 ro{od.x, od.y, od.z};
    V3 lastN{nf.x, nf.y, nf.z};
    uint32_t depth = first ? 0u : f.p.depth[pid];
    const uint32_t local = pid - S.slot0;
    const uint32_t sk = local / S.P, pix = local - sk * S.P;
    uint32_t px, py;
    slot_pixel(S, pix, px, py);
    P.wl = slot_wavelengths(S, sk, px, py);
    if (__float_as_uint(hr.w) & kShapeHit) {
        shade_shape_hit(S, tree, ct, __float_as_uint(hr.w) & ~kShapeHit, hr.x, ro, lastN, depth, P, L);
        f.p.L[pid] = make_float4(L.v[0], L.v[1], L.v[2], L.v[3]);
        f.p.nf[pid] = make_float4(lastN.x, lastN.y, lastN.z, __uint_as_float(P.flags));
        return;
    }
    const Qrng q = qrng_at(S.sample + sk, depth, pixel_hash(px, py));

    // ---- interaction (pathtrace.cu:86-149, 906-953): the hit's
    // leaf-order triangle record holds its vertices, vertex
    // indices, material and light index in one 64-byte line
    const TriRecord& tr = S.tris[__float_as_uint(hr.w)];
    // k_trace queues only paths that hit a triangle
    uint32_t ltype = LT_NONE, lindex = 0;
    const float b1 = hr.y, b2 = hr.z;
    const uint32_t vi0 = tr.i0, vi1 = tr.i1, vi2 = tr.i2;
    const V3 v0{tr.v0x, tr.v0y, tr.v0z}, v1{tr.v1x, tr.v1y, tr.v1z}, v2{tr.v2x, tr.v2y, tr.v2z};
    V3 gn;
    {
        V3 g = cross(v1 - v0, v2 - v0);
        g = g * rsqrtf(dot(g, g));
        gn = oct_dec_u32(oct_enc_u32(g));
    }
    const uint32_t mat = tr.material;
    {
        const uint32_t tl = tr.light;
        if (tl != 0xFFFFFFFFu) { ltype = LT_TRIANGLE; lindex = tl; }
    }
    P.flags |= PF_TRIANGULAR;
    V3 hitPos;
    {
        const float b0 = 1.0f - b1 - b2;
        hitPos = V3{v0.x * b0 + v1.x * b1 + v2.x * b2, v0.y * b0 + v1.y * b1 + v2.y * b2,
                    v0.z * b0 + v1.z * b1 + v2.z * b2};
    }
    V3 hitInterp = hitPos;
    const bool backHit = dot(gn, P.rd) > 0.0f;

    // ---- emission + MIS (pathtrace.cu:1185-1274)
    bool finished = false;
    if (ltype != LT_NONE) {
        const bool prevNEE = depth != 0 && !(P.flags & PF_PREV_SPECULAR);
        const float treePdf = prevNEE ? tree.pdf(ltype, lindex, ro, lastN) : 1.0f;
        float lightPdf = 0.0f;
        S4 Le = s4(0.0f);
        bool noEm = false;
        if (ltype == LT_TRIANGLE) {
            const EmissiveTri& tri = S.lights[lindex];
            const V3 a = ld3(tri.v), b = ld3(tri.v + 3), c = ld3(tri.v + 6);
            V3 n = cross(b - a, c - a);
            const float nl = len(n);
            n = n / nl;
            const V3 th = hitPos - ro;
            const float d2 = th.x * th.x + th.y * th.y + th.z * th.z;
            const float lcos = -dot(P.rd, n);
            noEm = lcos <= 0.0f;
            lightPdf = (1.0f / (nl / 2.0f)) * d2 / lcos;
            Le = tri_emission(S, tri, 1.0f - b1 - b2, b1, b2, P.wl, ct);
        }
        if (!noEm) {
            lightPdf *= treePdf;
            const float w = prevNEE ? power_heuristic(P.bxdfPdf, lightPdf) : 1.0f;
            L += P.T * Le * w;
        }
    }
    if (P.flags & PF_FINISHED) finished = true;

    if (!finished) {
        // ---- surface attributes (pathtrace.cu:1279-1567)
        const Material& m = S.materials[mat];
        const float bw0 = 1.0f - b1 - b2, bw1 = b1, bw2 = b2;
        V2 tc1{0, 0}, tc2{0, 0};
        V3 normal, tangent;
        bool flipT = false;
        const uint8_t* ab = S.attribs;
        const uint32_t stride = m.attributeStride;
        if (m.attributes & 3u) {
            V3 sN, sT, n0, n1, n2;
            if (m.attributes & 2u) {
                const float4 q0 = *reinterpret_cast<const float4*>(ab + (size_t)stride * vi0 + m.frameOffset);
                const float4 q1 = *reinterpret_cast<const float4*>(ab + (size_t)stride * vi1 + m.frameOffset);
                const float4 q2 = *reinterpret_cast<const float4*>(ab + (size_t)stride * vi2 + m.frameOffset);
                flipT = signbit(q0.w);
                V3 t0, t1, t2;
                quat_cols(q0, t0, n0);
                quat_cols(q1, t1, n1);
                quat_cols(q2, t2, n2);
                sN = bw0 * n0 + bw1 * n1 + bw2 * n2;
                sT = bw0 * t0 + bw1 * t1 + bw2 * t2;
            } else {
                const V2 e0 = attr2(ab, stride, m.normalOffset, vi0), e1 = attr2(ab, stride, m.normalOffset, vi1),
                         e2 = attr2(ab, stride, m.normalOffset, vi2);
                n0 = oct_dec(e0.x, e0.y); n1 = oct_dec(e1.x, e1.y); n2 = oct_dec(e2.x, e2.y);
                sN = bw0 * n0 + bw1 * n1 + bw2 * n2;
                const float sg = copysignf(1.0f, sN.z);
                const float aa = -1.0f / (sg + sN.z);
                const float bb = sN.x * sN.y * aa;
                sT = V3{1.0f + sg * (sN.x * sN.x) * aa, sg * bb, -sg * sN.x};
            }
            // Hanika 2021 shadow-terminator offset (pathtrace.cu:1380-1399)
            const V3 nA = nrm(n0), nB = nrm(n1), nC = nrm(n2);
            V3 tu = hitPos - v0, tv = hitPos - v1, tw = hitPos - v2;
            tu = tu - fminf(0.0f, dot(tu, nA)) * nA;
            tv = tv - fminf(0.0f, dot(tv, nB)) * nB;
            tw = tw - fminf(0.0f, dot(tw, nC)) * nC;
            hitInterp = hitPos + bw0 * tu + bw1 * tv + bw2 * tw;
            normal = nrm(sN);
            tangent = nrm(sT);
        } else {
            normal = gn;
            const float sg = copysignf(1.0f, gn.z);
            const float aa = -1.0f / (sg + gn.z);
            const float bb = gn.x * gn.y * aa;
            tangent = V3{1.0f + sg * (gn.x * gn.x) * aa, sg * bb, -sg * gn.x};
        }
        if (m.attributes & 8u) {
            const V2 t0 = attr2(ab, stride, m.texCoord1Offset, vi0), t1 = attr2(ab, stride, m.texCoord1Offset, vi1),
                     t2 = attr2(ab, stride, m.texCoord1Offset, vi2);
            tc1 = V2{bw0 * t0.x + bw1 * t1.x + bw2 * t2.x, bw0 * t0.y + bw1 * t1.y + bw2 * t2.y};
        }
        if (m.attributes & 16u) {
            const V2 t0 = attr2(ab, stride, m.texCoord2Offset, vi0), t1 = attr2(ab, stride, m.texCoord2Offset, vi1),
                     t2 = attr2(ab, stride, m.texCoord2Offset, vi2);
            tc2 = V2{bw0 * t0.x + bw1 * t1.x + bw2 * t2.x, bw0 * t0.y + bw1 * t1.y + bw2 * t2.y};
        }
        auto uvOf = [&](int slot) { return ((m.texCoordSets >> slot) & 1u) ? tc2 : tc1; };
        const V3 bit = cross(normal, tangent) * (flipT ? -1.0f : 1.0f);
        const Frame tbn{tangent, bit, normal};
        V3 n{0.0f, 0.0f, 1.0f};
        if (m.textures & 2u) {
            const V2 uv = uvOf(1);
            const float4 t = tex2d(S.texDescs, S.texels, m.tex[1], uv.x, uv.y, S.srgbLut);
            n = nrm(V3{t.x * 2.0f - 1.0f, t.y * 2.0f - 1.0f, t.z * 2.0f - 1.0f});
        }
        n = tbn.to_world(n);
        {
            const float dNs = dot(n, P.rd), dNg = dot(gn, P.rd);
            const bool fromOut = dNs > 0.0f && dNg < 0.0f;
            const bool fromIn = dNs < 0.0f && dNg > 0.0f;
            if (fromOut || fromIn) n = nrm(n - (P.rd * (dNs + (fromOut ? 0.001f : -0.001f))));
        }
        V3 sB = nrm(cross(n, tangent));
        const V3 sT2 = nrm(cross(sB, n));
        sB = sB * (flipT ? -1.0f : 1.0f);
        const Frame local{sT2, sB, n};
        lastN = n;

        Surface s;
        s.sheen = false;
        s.base = V3{1.0f, 1.0f, 1.0f};
        s.metalness = 1.0f;
        s.ior = 1.5f;
        s.roughness = 1.0f;
        s.transmission = 0.0f;
        s.sheenColor = V3{1.0f, 1.0f, 1.0f};
        s.sheenAlpha = 0.0f;
        bool skipped = false;
        {
            const bool bcTex = (m.textures & 1u) != 0, bcFac = (m.factors & 1u) != 0;
            float alphaV = 0.0f;
            if (bcTex) {
                const V2 uv = uvOf(1);  // the reference reads the normal map's set here (pathtrace.cu:1453)
                const float4 t = tex2d(S.texDescs, S.texels, m.tex[0], uv.x, uv.y, S.srgbLut);
                s.base = V3{t.x, t.y, t.z};
                alphaV = t.w;
                if (bcFac) {
                    s.base = s.base * V3{m.baseColor[0], m.baseColor[1], m.baseColor[2]};
                    alphaV *= m.baseColor[3];
                }
            } else if (bcFac) {
                s.base = V3{m.baseColor[0], m.baseColor[1], m.baseColor[2]};
                alphaV = m.baseColor[3];
            }
            if (m.factors & 16u) {
                if (alphaV < m.alphaCutoff) skipped = true;
            } else if (m.factors & 32u) {
                if (q.s1(QD_ALPHA_BLEND) > alphaV) skipped = true;
            }
            const bool metF = (m.factors & 2u) != 0, rF = (m.factors & 4u) != 0;
            if (m.textures & 4u) {
                const V2 uv = uvOf(2);
                const float4 t = tex2d(S.texDescs, S.texels, m.tex[2], uv.x, uv.y, S.srgbLut);
                s.metalness = t.z * (metF ? m.metalness : 1.0f);
                s.roughness = t.y * (rF ? m.roughness : 1.0f);
            } else {
                if (metF) s.metalness = m.metalness;
                if (rF) s.roughness = m.roughness;
            }
            const bool trF = (m.factors & 8u) != 0;
            if (m.textures & 8u) {
                const V2 uv = uvOf(3);
                const float4 t = tex2d(S.texDescs, S.texels, m.tex[3], uv.x, uv.y, S.srgbLut);
                s.transmission = t.x * (trF ? m.transmission : 1.0f);
            } else if (trF) {
                s.transmission = m.transmission;
            }
            s.ior = m.ior;
            if (m.flags & 2u) {
                s.sheen = true;
                const bool shcF = (m.factors & 64u) != 0;
                if (m.textures & 16u) {
                    const V2 uv = uvOf(4);
                    const float4 t = tex2d(S.texDescs, S.texels, m.tex[4], uv.x, uv.y, S.srgbLut);
                    s.sheenColor = V3{t.x * (shcF ? m.sheenColor[0] : 1.0f), t.y * (shcF ? m.sheenColor[1] : 1.0f),
                                      t.z * (shcF ? m.sheenColor[2] : 1.0f)};
                } else if (shcF) {
                    s.sheenColor = V3{m.sheenColor[0], m.sheenColor[1], m.sheenColor[2]};
                }
                const bool shrF = (m.factors & 128u) != 0;
                if (m.textures & 32u) {
                    const V2 uv = uvOf(5);
                    const float4 t = tex2d(S.texDescs, S.texels, m.tex[5], uv.x, uv.y, S.srgbLut);
                    const float r = t.w * (shrF ? m.sheenRoughness : 1.0f);
                    s.sheenAlpha = r * r;
                } else if (shrF) {
                    s.sheenAlpha = m.sheenRoughness;
                }
            }
        }

        // ---- NEE sample (pathtrace.cu:1569-1720) + BxDF
        if (!skipped) {
            Light dl;
            dl.spectrum = s4(0.0f);
            dl.dir = V3{0, 0, 0};
            dl.pdf = 0.0f;
            dl.occluded = false;
            V3 lro{0, 0, 0};
            float dToL = 0.0f;
            if (!S.envOn && !treeOn) {
                dl.occluded = true;
            } else {
                const V3 rnd = q.s3(QD_LIGHT);
                uint32_t type = LT_NONE, index = 0;
                float treePdf = 0.0f;
                tree.sample(rnd.z, hitPos, lastN, type, index, treePdf);
                float lightPdf = 0.0f;
                if (type == LT_TRIANGLE) {
                    const EmissiveTri& tri = S.lights[index];
                    const V3 a = ld3(tri.v), b = ld3(tri.v + 3), c = ld3(tri.v + 6);
                    V3 nl = cross(b - a, c - a);
                    const float nlen = len(nl);
                    nl = nl / nlen;
                    V2 u2{rnd.x, rnd.y};
                    if (u2.x + u2.y > 1.0f) u2 = V2{1.0f - u2.x, 1.0f - u2.y};
                    const float ba = 1.0f - u2.x - u2.y, bb = u2.x, bc = u2.y;
                    const V3 lp = a + (b - a) * u2.x + (c - a) * u2.y;
                    const bool inSample = dot(lastN, lp - hitPos) < 0.0f;
                    lro = offset_point(inSample ? hitPos : hitInterp, inSample ? -gn : gn);
                    const V3 rl = lp - lro;
                    const float s2 = rl.x * rl.x + rl.y * rl.y + rl.z * rl.z;
                    dToL = sqrtf(s2);
                    dl.dir = rl / dToL;
                    const float lcos = -dot(nl, dl.dir);
                    dl.occluded = lcos <= 0.0f;
                    lightPdf = (1.0f / (nlen / 2.0f)) * s2 / lcos;
                    dl.spectrum = tri_emission(S, tri, ba, bb, bc, P.wl, ct);
                } else if (type == LT_SHAPE) {
                    // a disk point (area measure) or a direction in the cone a
                    // sphere subtends (solid-angle measure)
                    const ShapeLight& sl = S.shapes[index - S.shapeBase];
                    V3 lp;
                    bool inside = false;
                    if (sl.type == SHAPE_DISK) lp = disk_sample(sl, V2{rnd.x, rnd.y});
                    else lp = sphere_sample(sl, hitPos, V2{rnd.x, rnd.y}, lightPdf, inside);
                    if (inside) {
                        dl.occluded = true;
                    } else {
                        const bool inSample = dot(lastN, lp - hitPos) < 0.0f;
                        lro = offset_point(inSample ? hitPos : hitInterp, inSample ? -gn : gn);
                        const V3 rl = lp - lro;
                        const float s2 = rl.x * rl.x + rl.y * rl.y + rl.z * rl.z;
                        dToL = sqrtf(s2);
                        dl.dir = rl / dToL;
                        if (sl.type == SHAPE_DISK) {
                            const float lcos = -dot(shape_normal(sl), dl.dir);
                            dl.occluded = lcos <= 0.0f;
                            lightPdf = (1.0f / (kShapePi * sl.radius * sl.radius)) * s2 / lcos;
                        }
                        dl.spectrum = shape_emission(S, sl, P.wl, ct);
                    }
                } else if (type == LT_SKY) {
                    const float W = (float)S.envW, H = (float)S.envH;
                    const float fcx = fminf(rnd.x * (W - 1.0f), W - 1.0001f);
                    const float fcy = fminf(rnd.y * (H - 1.0f), H - 1.0001f);
                    const float flx = floorf(fcx), fly = floorf(fcy);
                    const uint32_t ix = (uint32_t)flx, iy = (uint32_t)fly;
                    const float m0 = (float)S.envMarg[iy], m1 = (float)S.envMarg[iy + 1];
                    // the reference's float row offset (pathtrace.cu:1670), capped at
                    // the last row: where m0 * W rounds (W with a large odd factor,
                    // products above 2^24) it could otherwise read past the table
                    const size_t row = min((size_t)(uint32_t)(m0 * W), (size_t)(S.envH - 1) * (size_t)S.envW);
                    const float c0 = (float)S.envCond[row + ix], c1 = (float)S.envCond[row + ix + 1];
                    const float ux = lerpf(c0, c1, fcx - flx) * (1.0f / W);
                    const float uy = lerpf(m0, m1, fcy - fly) * (1.0f / H);
                    float sp, cp, st, ctt;
                    sincosf(2.0f * 3.14159265358979f * ux, &sp, &cp);
                    sincosf(3.14159265358979f * uy, &st, &ctt);
                    dl.dir = V3{-st * cp, st * sp, ctt};
                    const V3 rgb = env_lookup(S.envTex, S.envW, S.envH, S.envC, ux, uy);
                    dl.spectrum = ct.rgb(rgb, P.wl);
                    lightPdf = (1.0f / (4.0f * 3.14159265358979f)) * ((rgb.x + rgb.y + rgb.z) / 3.0f) * st / S.envIntegral;
                    dToL = 3.402823466e+38f;
                    lro = offset_point(hitInterp, backHit ? -gn : gn);
                } else {
                    dl.occluded = true;
                }
                dl.pdf = lightPdf * treePdf;
            }
            S4 Lnee = s4(0.0f);
            const bool guideHere = GUIDED && m.bxdf == 2 && s.roughness >= kGuideMinRough &&
                                   dot(local.n, -P.rd) > 0.0f;
            if (guideHere) {
                V3 wiw;
                float pb = 0.0f;
                const S4 Tb = P.T;
                const float pmix = complex_guided(ct, lut, dl, q, local, s, P, Lnee, f.gmix + (size_t)i * kMixFloats,
                                                  wiw, pb);
                if (pmix == 0.0f) {
                    P.flags |= PF_FINISHED;
                } else if (f.collect && pixel_hash(pid, depth * 0x9E3779B9u + S.sample) < f.recThreshold) {
                    recordSlot = true;
                    recDir = wiw;
                    recPb = pb;
                    recT = Tb;
                }
            } else if (m.bxdf == 0) {
                conductor_bxdf(S.spectra + (size_t)m.iorSpectrum * kDense, S.spectra + (size_t)m.acSpectrum * kDense,
                               m.mfRoughness, dl, q, local, P, Lnee);
            } else if (m.bxdf == 2) {
                complex_bxdf(ct, lut, dl, q, local, s, P, Lnee);
            } else {
                dielectric_bxdf(S.spectra + (size_t)m.iorSpectrum * kDense, m.mfRoughness, dl, q, local, P, Lnee);
            }
            if (f.noDrop && (P.flags & PF_TERMINATED)) P.flags = (P.flags & ~PF_TERMINATED) | PF_FINISHED;
            if (!dl.occluded) {
                shadow = true;
                f.p.Lnee[pid] = make_float4(Lnee.v[0], Lnee.v[1], Lnee.v[2], Lnee.v[3]);
                f.p.sro[pid] = make_float4(lro.x, lro.y, lro.z, dToL - 0.1f);
                f.p.srd[pid] = make_float4(dl.dir.x, dl.dir.y, dl.dir.z, 0.0f);
            }
        }

        // ---- Russian roulette (pathtrace.cu:1736-1744)
        bool rr = false;
        {
            const float tMax = smax4(P.T) * P.refr;
            if (tMax < 1.0f && depth > 0) {
                float qq = fmaxf(0.0f, 1.0f - tMax);
                qq = qq * qq;
                if (q.s1(QD_ROULETTE) < qq) rr = true;
                else P.T = P.T * (1.0f / (1.0f - qq));
            }
        }
        if (!rr) {
            // ---- next origin + direction fix (pathtrace.cu:1746-1766)
            const bool refracted = (P.flags & PF_RAY_REFRACTED) != 0;
            const bool triG = (P.flags & PF_TRIANGULAR) != 0;
            const bool in = refracted || skipped;
            const V3 org = ((in == backHit) && triG) ? hitInterp : hitPos;
            const V3 off = in == backHit ? gn : -gn;
            ro = offset_point(org, off);
            const float dNgRo = dot(gn, P.rd);
            if (dNgRo > 0.0f && (backHit != refracted) && !skipped) P.rd = nrm(P.rd - (gn * (dNgRo + 0.001f)));
            // ---- depth caps (pathtrace.cu:1768-1779)
            bool cont = ++depth < S.maxDepth;
            if (!(P.flags & PF_CUR_SPECULAR) && !skipped)
                cont = (P.flags & PF_REFRACTION_HAPPENED) ? depth < S.maxTrans : depth < S.maxRefl;
            if (!cont) P.flags |= PF_FINISHED;
            // ---- updateStateFlags (pathtrace.cu:954-972)
            const uint32_t excl = PF_CUR_SPECULAR | PF_PREV_SPECULAR | PF_TRIANGULAR | PF_RIGHT_HANDED | PF_RAY_REFRACTED;
            uint32_t incl = (P.flags & PF_CUR_SPECULAR) ? PF_PREV_SPECULAR : PF_REGULARIZED;
            if (P.flags & PF_RAY_REFRACTED) incl |= PF_REFRACTION_HAPPENED;
            P.flags = (P.flags & ~excl) | incl;
            alive = !(P.flags & PF_FINISHED) && !(P.flags & PF_TERMINATED);
        }
    }
    f.p.L[pid] = make_float4(L.v[0], L.v[1], L.v[2], L.v[3]);
    f.p.T[pid] = make_float4(P.T.v[0], P.T.v[1], P.T.v[2], P.T.v[3]);
    f.p.od[pid] = make_float4(ro.x, ro.y, ro.z, P.bxdfPdf);
    f.p.dr[pid] = make_float4(P.rd.x, P.rd.y, P.rd.z, P.refr);
    f.p.nf[pid] = make_float4(lastN.x, lastN.y, lastN.z, __uint_as_float(P.flags));
    f.p.depth[pid] = depth;
}

// Turns the pending records of a finished path into training records. The
// weight is the sampled contribution over its density, f |cos| L_i / pdf =
// radiance gathered after the vertex / throughput before it, averaged over
// the wavelengths; the BSDF sampler's density of the direction goes along for
// the one-sample MIS density the loss fits.
__device__ __forceinline__ void finish_records(const Frame1& f, uint32_t slot, float4 L) {
    uint32_t r = f.p.prec[slot];
    for (int guard = 0; r != kNoRec && guard < 1024; ++guard) {
        const bool pending = (r & kRecPending) != 0;
        r &= ~kRecPending;
        const float* d = f.grec + (size_t)r * kPendingRec;
        float* t = f.gtrain + (size_t)r * guide::kRec;
        for (int k = 0; k < 14; ++k) t[k] = d[k];
        float w = 0.0f;
        if (!pending) {
            const float dl[4] = {L.x - d[20], L.y - d[21], L.z - d[22], L.w - d[23]};
            float li = 0.0f;
            for (int k = 0; k < 4; ++k) li += d[16 + k] > 0.0f ? dl[k] / d[16 + k] : 0.0f;
            w = fmaxf(0.0f, 0.25f * li);
            if (!(w < 3.0e38f)) w = 0.0f;
        }
        t[14] = w;
        t[15] = d[14];
        r = __float_as_uint(d[15]);
    }
}

// Adds the S samples of the pass for scene-local pixel slot i to the FP64
// accumulation (pathtrace.cu:1789-1812); they are added one at a time in
// sample order, so any split of a sample range into passes gives the same sums.
__device__ __forceinline__ void accumulate_slot(const Frame1& f, const SceneView& S, uint32_t i) {
    uint32_t px, py;
    slot_pixel(S, i, px, py);
    const uint32_t pixel = py * S.W + px;
    double4 a = S.first ? make_double4(0.0, 0.0, 0.0, 0.0) : S.accum[pixel];
    const ColorTables ct{S.colors};
    uint32_t bad = 0;
    for (uint32_t k = 0; k < S.S; ++k) {
        const uint32_t j = S.slot0 + k * S.P + i;
        if (f.collect) finish_records(f, j, f.p.L[j]);
        const uint32_t flags = __float_as_uint(f.p.nf[j].w);
        if (!(flags & PF_TERMINATED)) {
            const float4 Lv = f.p.L[j];
            const Wl wl = slot_wavelengths(S, k, px, py);
            const S4 pdf = wl_effective_pdf(wl, flags);
            S4 L{{Lv.x, Lv.y, Lv.z, Lv.w}};
            for (int c = 0; c < kNL; ++c) L.v[c] = pdf.v[c] != 0.0f ? L.v[c] / pdf.v[c] : 0.0f;
            double xyz[3];
            ct.xyz(wl, L, xyz);
            if (isfinite(xyz[0]) && isfinite(xyz[1]) && isfinite(xyz[2])) {
                a.x += xyz[0];
                a.y += xyz[1];
                a.z += xyz[2];
            } else {
                ++bad;   // counted and added as a black sample instead of poisoning the pixel
            }
        }
        a.w += 1.0;
    }
    S.accum[pixel] = a;
    if (bad) atomicAdd(f.stats + STAT_NONFINITE, (unsigned long long)bad);
}

}  // namespace wf
}  // namespace rbh
