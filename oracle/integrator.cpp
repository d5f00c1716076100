// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// The megakernel integrator __raygen__main (src/kernels/pathtrace.cu:1114-1813)
// restated as scalar C++, one pixel at a time, OpenMP over 16x16 tiles. The
// closest-hit payload (pathtrace.cu:86-149) and interaction unpacking
// (pathtrace.cu:906-953) are restated for a flattened, camera-relative scene
// with identity instance transforms. This is both the image-level oracle and
// the CPU throughput baseline.
#include <chrono>
#include "render.hpp"
#include "rb_bvh.hpp"
#include <vector>
#include <atomic>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

namespace {

struct Ctx {
    const orc_scene* sc;
    const orc_params* p;
    CpuBvh bvh;
    LightTreeView tree;
    const float* cA; const float* cB; const float* cC;
};

inline Spec rgb_spec(const orc_scene* sc, V3 rgb, const Wavelengths& wl) {
    return rgb_to_spectrum(rgb, wl, sc->basis, sc->basis + NDENSE, sc->basis + 2 * NDENSE);
}

// Emission of an emissive triangle at barycentrics (b1, b2).
// pathtrace.cu:1210-1220 / 1634-1651; spectral emission (flag 4) replaces the
// RGB factor by emissive[0] x spectra[emissionSpectrum] — the reference
// declares MaterialData::emissionSpectrumDataIndex (material.h:77) but only
// emits RGB.
Spec triangle_emission(const orc_scene* sc, const orc_emissive_tri& tri, float ba, float bb, float bc, const Wavelengths& wl) {
    const orc_material& m = sc->materials[tri.materialIndex];
    if (m.flags & MF_SPECTRAL_EMISSION)
        return dense_sample(sc->spectra + (size_t)m.emissionSpectrum * NDENSE, wl) * m.emissive[0];
    float e[3] = {m.emissive[0], m.emissive[1], m.emissive[2]};
    if (m.textures & TX_EMISSION) {
        float u = ba * tri.uvs[0] + bb * tri.uvs[2] + bc * tri.uvs[4];
        float v = ba * tri.uvs[1] + bb * tri.uvs[3] + bc * tri.uvs[5];
        float t[4];
        tex_sample(sc, m.tex[TS_EM], u, v, t);
        e[0] *= t[0]; e[1] *= t[1]; e[2] *= t[2];
    }
    return rgb_spec(sc, {e[0], e[1], e[2]}, wl);
}

// Environment-map texel lookup (clamp, linear), normalised (u, v).
void env_lookup(const orc_scene* sc, float u, float v, float out[3]) {
    int W = (int)sc->envWidth, H = (int)sc->envHeight, ch = (int)sc->envChannels;
    float x = u * W - 0.5f, y = v * H - 0.5f;
    float fx = std::floor(x), fy = std::floor(y);
    float ax = x - fx, ay = y - fy;
    auto cl = [](int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
    int x0 = cl((int)fx, W), x1 = cl((int)fx + 1, W), y0 = cl((int)fy, H), y1 = cl((int)fy + 1, H);
    for (int k = 0; k < 3; ++k) {
        auto at = [&](int xi, int yi) { return sc->envTexels[((size_t)yi * W + xi) * ch + k]; };
        float a = at(x0, y0) * (1.0f - ax) + at(x1, y0) * ax;
        float b = at(x0, y1) * (1.0f - ax) + at(x1, y1) * ax;
        out[k] = a * (1.0f - ay) + b * ay;
    }
}

V3 vread3(const uint8_t* p) { float f[3]; std::memcpy(f, p, 12); return {f[0], f[1], f[2]}; }
V2 vread2(const uint8_t* p) { float f[2]; std::memcpy(f, p, 8); return {f[0], f[1]}; }

// glm::mat3_cast of a unit quaternion (x, y, z, w)
M3 quat_to_mat(const float q[4]) {
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float xx = x * x, yy = y * y, zz = z * z, xz = x * z, xy = x * y, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
    M3 m;
    m.c[0] = {1.0f - 2.0f * (yy + zz), 2.0f * (xy + wz), 2.0f * (xz - wy)};
    m.c[1] = {2.0f * (xy - wz), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + wx)};
    m.c[2] = {2.0f * (xz + wy), 2.0f * (yz - wx), 1.0f - 2.0f * (xx + yy)};
    return m;
}

// ---- shape lights: disks and spheres (DiskLightData / SphereLightData,
// light_tree_types.h:24-39; the reference's programs for them are commented
// out, pathtrace.cu:150-228, its samplers are sampling.h:32-110). A disk
// emits along its frame's third column and occludes from both sides; a
// sphere emits outwards; a path that hits either ends there.
V3 shape_center(const orc_shape_light& s) { return {s.position[0], s.position[1], s.position[2]}; }

// ray parameter of the hit in (0, tmax), else tmax (pathtrace.cu:150-190)
float disk_t(const orc_shape_light& s, V3 o, V3 d, float tmax) {
    V3 n = quat_to_mat(s.frame).c[2];
    V3 oc = o - shape_center(s);
    float t = -dot(n, oc) / dot(d, n);
    V3 p = oc + d * t;
    return (t > 0.0f && t < tmax && dot(p, p) < s.radius * s.radius) ? t : tmax;
}
// nearest root in (0, tmax) of |o + t d - c| = r, roots from the
// closest-approach discriminant and the non-cancelling quadratic form
float sphere_t(const orc_shape_light& s, V3 o, V3 d, float tmax) {
    V3 oc = o - shape_center(s);
    float b = dot(oc, d);
    V3 qc = oc - d * b;
    float r2 = s.radius * s.radius;
    float disc = r2 - dot(qc, qc);
    if (!(disc >= 0.0f)) return tmax;
    float q = -b - std::copysign(std::sqrt(disc), b);
    float c = dot(oc, oc) - r2;
    float t0 = c / q, t1 = q;
    if (t0 > t1) std::swap(t0, t1);
    float t = t0 > 0.0f ? t0 : t1;
    return (t > 0.0f && t < tmax) ? t : tmax;
}
float shape_t(const orc_shape_light& s, V3 o, V3 d, float tmax) {
    return s.type == 0u ? disk_t(s, o, d, tmax) : sphere_t(s, o, d, tmax);
}

// 1 - cos(theta_max) of the cone a sphere subtends from distance dist, with
// the small-angle form below sin^2 = 0.00068523 (sampling.h:79-92)
float sphere_omc(float dist, float r, float& sinMax, float& sin2Max, float& cosMax) {
    sinMax = r / dist;
    sin2Max = sinMax * sinMax;
    cosMax = std::max(0.0f, std::sqrt(std::max(0.0f, 1.0f - sin2Max)));
    return sin2Max < 0.00068523f ? sin2Max / 2.0f : 1.0f - cosMax;
}
float sphere_cone_pdf(const orc_shape_light& s, V3 o) {
    float a, b, c;
    float omc = sphere_omc(length(shape_center(s) - o), s.radius, a, b, c);
    return 1.0f / (2.0f * PI * omc);
}

// concentric square-to-disk map (sampling.h:34-51), centre to centre
V2 concentric(V2 u) {
    float a = u.x * 2.0f - 1.0f, b = u.y * 2.0f - 1.0f;
    if (a == 0.0f && b == 0.0f) return {0.0f, 0.0f};
    float r, th;
    if (std::fabs(a) > std::fabs(b)) { r = a; th = (PI / 4.0f) * (b / a); }
    else { r = b; th = (PI / 2.0f) - (PI / 4.0f) * (a / b); }
    return {r * std::cos(th), r * std::sin(th)};
}

// cone sampling of a sphere seen from o (sampling.h:71-110)
V3 sphere_sample(const orc_shape_light& s, V3 o, V2 u, float& pdf, bool& inside) {
    V3 c = shape_center(s);
    V3 w = o - c;
    float dist = length(w);
    inside = !(dist > s.radius);
    if (inside) { pdf = 0.0f; return c; }
    w = w / dist;
    float sinMax, sin2Max, cosMax;
    float omc = sphere_omc(dist, s.radius, sinMax, sin2Max, cosMax);
    float cosT = (cosMax - 1.0f) * u.x + 1.0f;
    float sin2T = 1.0f - cosT * cosT;
    if (sin2Max < 0.00068523f) {
        sin2T = sin2Max * u.x;
        cosT = std::sqrt(1.0f - sin2T);
    }
    float cosA = sin2T / sinMax + cosT * std::sqrt(std::max(0.0f, 1.0f - sin2T / (sinMax * sinMax)));
    float sinA = std::sqrt(std::max(0.0f, 1.0f - cosA * cosA));
    float phi = u.y * 2.0f * PI;
    // util.h:31-38 basis around w
    float sign = std::copysign(1.0f, w.z);
    float aa = -1.0f / (sign + w.z);
    float bb = w.x * w.y * aa;
    V3 bu{1.0f + sign * (w.x * w.x) * aa, sign * bb, -sign * w.x};
    V3 bv{bb, sign + (w.y * w.y) * aa, -w.y};
    V3 dir = w * cosA + bu * (sinA * std::cos(phi)) + bv * (sinA * std::sin(phi));
    pdf = 1.0f / (2.0f * PI * omc);
    return c + dir * s.radius;
}

// the material's emission (spectral, or RGB lifted; no texture) x power scale
Spec shape_emission(const orc_scene* sc, const orc_shape_light& s, const Wavelengths& wl) {
    const orc_material& m = sc->materials[s.materialIndex];
    Spec e = (m.flags & MF_SPECTRAL_EMISSION) ? dense_sample(sc->spectra + (size_t)m.emissionSpectrum * NDENSE, wl) * m.emissive[0]
                                              : rgb_spec(sc, {m.emissive[0], m.emissive[1], m.emissive[2]}, wl);
    return e * s.powerScale;
}

struct Stats { uint64_t closest = 0, shadow = 0, vertices = 0; };

void render_pixel(const Ctx& C, uint32_t px, uint32_t py, double res[4], Stats& st) {
    const orc_scene* sc = C.sc;
    const orc_params* P = C.p;
    const float invW = 1.0f / (float)P->width, invH = 1.0f / (float)P->height;
    QState q = QState::make(P->sampleOffset, pixel_hash(px, py));
    const bool envOn = sc->envEnabled != 0;
    const bool treeOn = sc->nodes != nullptr && sc->nnodes > 0;
    for (uint32_t smp = 0; smp < P->sampleCount; ++smp) {
        PathIO path;
        path.wl = sample_visible(sobol1(q, D_WAVELENGTH));
        path.L = specs(0.0f);
        path.T = specs(1.0f);
        path.refrScale = 1.0f;
        path.bxdfPDF = 1.0f;
        path.flags = 0;
        uint32_t depth = 0;
        V3 ro{0, 0, 0}, lastN{0, 0, 0};
        {
            V2 sub = sobol2(q, D_FILTER);
            V2 fo = gaussian_filter_sample(sub);
            V2 c{(float)px + fo.x, (float)py + fo.y};
            V3 U{P->camU[0], P->camU[1], P->camU[2]}, Vv{P->camV[0], P->camV[1], P->camV[2]}, Wv{P->camW[0], P->camW[1], P->camW[2]};
            if (P->depthOfField) {
                // src/device/dir_gen.h:358-374
                V2 ls = sobol2(q, D_LENS);
                float xs = (2.0f * c.x * invW - 1.0f) * P->perspectiveScaleW;
                float zs = (2.0f * c.y * invH - 1.0f) * P->perspectiveScaleH;
                V3 focus = V3{xs, zs, 1.0f} * P->focusDistance;
                V2 d = disk_polar(ls);
                V3 lens{d.x * P->aperture, d.y * P->aperture, 0.0f};
                ro = U * lens.x + Wv * lens.y;
                V3 rd = normalize(focus - lens);
                path.rD = U * rd.x + Wv * rd.y + Vv * rd.z;
            } else {
                // src/device/dir_gen.h:346-357
                float xs = 2.0f * c.x * invW - 1.0f;
                float zs = 2.0f * c.y * invH - 1.0f;
                path.rD = normalize(Vv + U * xs * P->perspectiveScaleW + Wv * zs * P->perspectiveScaleH);
                ro = {0, 0, 0};
            }
        }
        bool finished = false;  // reached finishPath
        do {
            // ---- closest hit (pathtrace.cu:1162-1177 + 86-149 + 906-953)
            CpuBvh::Hit hit{};
            ++st.closest;
            ++st.vertices;
            // the nearest shape light bounds the triangle query
            float shT = FLT_MAX;
            uint32_t shIdx = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < sc->nshapes; ++k) {
                float t = shape_t(sc->shapes[k], ro, path.rD, shT);
                if (t < shT) { shT = t; shIdx = k; }
            }
            bool isHit = C.bvh.trace(ro, path.rD, shT, false, hit);
            if (!isHit && shIdx != 0xFFFFFFFFu) {
                // a shape light: its emission with the MIS weight of the previous
                // BxDF sample, then the path ends
                const orc_shape_light& sl = sc->shapes[shIdx];
                const bool prevNEE = depth != 0 && !(path.flags & PF_PREVIOUS_HIT_SPECULAR);
                float lightPdf = 0.0f;
                bool noEmission;
                if (sl.type == 0u) {
                    V3 th = path.rD * shT;
                    float d2 = th.x * th.x + th.y * th.y + th.z * th.z;
                    float lcos = -dot(path.rD, quat_to_mat(sl.frame).c[2]);
                    noEmission = lcos <= 0.0f;
                    lightPdf = (1.0f / (PI * sl.radius * sl.radius)) * d2 / lcos;
                } else {
                    V3 oc = shape_center(sl) - ro;
                    noEmission = !(dot(oc, oc) > sl.radius * sl.radius);
                    if (!noEmission) lightPdf = sphere_cone_pdf(sl, ro);
                }
                if (!noEmission) {
                    float treePdf = prevNEE ? light_tree_pdf(C.tree, LT_SHAPE, sc->nlights + shIdx, ro, lastN) : 1.0f;
                    lightPdf *= treePdf;
                    float w = prevNEE ? power_heuristic(path.bxdfPDF, lightPdf) : 1.0f;
                    path.L += path.T * shape_emission(sc, sl, path.wl) * w;
                }
                path.flags |= PF_FINISHED;
                finished = true;
                break;
            }
            uint32_t ltype = LT_NONE, lindex = 0;
            V3 gn{0, 0, 0}, hitPos{0, 0, 0}, hitInterp{0, 0, 0}, vtx[3];
            float b1 = 0, b2 = 0;
            uint32_t prim = 0, mat = 0;
            bool backHit = false;
            if (!isHit) {
                ltype = LT_SKY;
                path.flags &= ~PF_TRIANGULAR_GEOMETRY;
            } else {
                prim = hit.prim;
                for (int j = 0; j < 3; ++j) {
                    const float* v = sc->verts + 3 * (size_t)sc->indices[3 * (size_t)prim + j];
                    vtx[j] = {v[0], v[1], v[2]};
                }
                V3 e1 = vtx[1] - vtx[0], e2 = vtx[2] - vtx[0];
                V3 g = cross(e1, e2);
                g = g * (1.0f / std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z));
                gn = oct_decode_u32(oct_encode_u32(g));
                mat = sc->triMaterial[prim];
                if (sc->triLight[prim] != 0xFFFFFFFFu) { ltype = LT_TRIANGLE; lindex = sc->triLight[prim]; }
                path.flags |= PF_TRIANGULAR_GEOMETRY;
                b1 = hit.b1; b2 = hit.b2;
                hitPos.x = vtx[0].x * (1.0f - b1 - b2) + vtx[1].x * b1 + vtx[2].x * b2;
                hitPos.y = vtx[0].y * (1.0f - b1 - b2) + vtx[1].y * b1 + vtx[2].y * b2;
                hitPos.z = vtx[0].z * (1.0f - b1 - b2) + vtx[1].z * b1 + vtx[2].z * b2;
                hitInterp = hitPos;
                backHit = dot(gn, path.rD) > 0.0f;
            }
            const orc_material* M = isHit ? &sc->materials[mat] : nullptr;

            // ---- emission at the hit (pathtrace.cu:1185-1274)
            if (ltype != LT_NONE) {
                const bool prevNEE = depth != 0 && !(path.flags & PF_PREVIOUS_HIT_SPECULAR);
                float treePdf = prevNEE ? light_tree_pdf(C.tree, ltype, lindex, ro, lastN) : 1.0f;
                float lightPdf = 0.0f;
                Spec Le = specs(0.0f);
                bool noEmission = false;
                V3 toHit = hitPos - ro;
                float d2 = toHit.x * toHit.x + toHit.y * toHit.y + toHit.z * toHit.z;
                if (ltype == LT_TRIANGLE) {
                    const orc_emissive_tri& tri = sc->lights[lindex];
                    V3 a{tri.vertices[0], tri.vertices[1], tri.vertices[2]};
                    V3 b{tri.vertices[3], tri.vertices[4], tri.vertices[5]};
                    V3 c{tri.vertices[6], tri.vertices[7], tri.vertices[8]};
                    V3 n = cross(b - a, c - a);
                    float nl = length(n);
                    n = n / nl;
                    float area = nl / 2.0f;
                    float lcos = -dot(path.rD, n);
                    noEmission = lcos <= 0.0f;
                    lightPdf = (1.0f / area) * d2 / lcos;
                    Le = triangle_emission(sc, tri, 1.0f - b1 - b2, b1, b2, path.wl);
                } else {  // SKY
                    if (envOn) {
                        float phi = std::atan2(path.rD.y, path.rD.x);
                        float theta = std::acos(path.rD.z);
                        float sky[3];
                        env_lookup(sc, 0.5f - phi / (2.0f * PI), theta / PI, sky);
                        lightPdf = (1.0f / (4.0f * PI)) * ((sky[0] + sky[1] + sky[2]) / 3.0f) * std::sin(theta) / sc->envIntegral;
                        Le = rgb_spec(sc, {sky[0], sky[1], sky[2]}, path.wl);
                        path.flags |= PF_FINISHED;
                    } else {
                        path.flags |= PF_FINISHED;
                        noEmission = true;
                    }
                }
                if (!noEmission) {
                    lightPdf *= treePdf;
                    float w = prevNEE ? power_heuristic(path.bxdfPDF, lightPdf) : 1.0f;
                    path.L += path.T * Le * w;
                }
            }
            if (path.flags & PF_FINISHED) { finished = true; break; }

            // ---- surface attributes (pathtrace.cu:1279-1567)
            Local local;
            Surface surf;
            bool skipped = false;
            {
                const orc_material& m = *M;
                float bw[3] = {1.0f - b1 - b2, b1, b2};
                uint32_t vi[3] = {sc->indices[3 * (size_t)prim], sc->indices[3 * (size_t)prim + 1], sc->indices[3 * (size_t)prim + 2]};
                V2 tc1{0, 0}, tc2{0, 0};
                V3 normal, tangent;
                bool flipT = false;
                const uint8_t* ab = sc->attribs;
                const uint32_t stride = m.attributeStride;
                if (m.attributes & (AT_NORMAL | AT_FRAME)) {
                    V3 sN, sT, ns[3];
                    if (m.attributes & AT_FRAME) {
                        float qd[3][4];
                        for (int j = 0; j < 3; ++j) std::memcpy(qd[j], ab + (size_t)stride * vi[j] + m.frameOffset, 16);
                        flipT = std::signbit(qd[0][3]);
                        M3 fm[3] = {quat_to_mat(qd[0]), quat_to_mat(qd[1]), quat_to_mat(qd[2])};
                        for (int j = 0; j < 3; ++j) ns[j] = fm[j].c[2];
                        sN = bw[0] * ns[0] + bw[1] * ns[1] + bw[2] * ns[2];
                        sT = bw[0] * fm[0].c[0] + bw[1] * fm[1].c[0] + bw[2] * fm[2].c[0];
                    } else {
                        for (int j = 0; j < 3; ++j) ns[j] = oct_decode(vread2(ab + (size_t)stride * vi[j] + m.normalOffset));
                        sN = bw[0] * ns[0] + bw[1] * ns[1] + bw[2] * ns[2];
                        float sign = std::copysign(1.0f, sN.z);
                        float aa = -1.0f / (sign + sN.z);
                        float bb = sN.x * sN.y * aa;
                        sT = {1.0f + sign * (sN.x * sN.x) * aa, sign * bb, -sign * sN.x};
                    }
                    // Hanika 2021 terminator offset (pathtrace.cu:1380-1399)
                    V3 nA = normalize(ns[0]), nB = normalize(ns[1]), nC = normalize(ns[2]);
                    V3 tu = hitPos - vtx[0], tv = hitPos - vtx[1], tw = hitPos - vtx[2];
                    float du = std::min(0.0f, dot(tu, nA)), dv = std::min(0.0f, dot(tv, nB)), dw = std::min(0.0f, dot(tw, nC));
                    tu = tu - du * nA; tv = tv - dv * nB; tw = tw - dw * nC;
                    hitInterp = hitPos + bw[0] * tu + bw[1] * tv + bw[2] * tw;
                    normal = normalize(sN);
                    tangent = normalize(sT);
                } else {
                    normal = gn;
                    float sign = std::copysign(1.0f, gn.z);
                    float aa = -1.0f / (sign + gn.z);
                    float bb = gn.x * gn.y * aa;
                    tangent = {1.0f + sign * (gn.x * gn.x) * aa, sign * bb, -sign * gn.x};
                }
                if (m.attributes & AT_TEX1) {
                    V2 t0 = vread2(ab + (size_t)stride * vi[0] + m.texCoord1Offset), t1 = vread2(ab + (size_t)stride * vi[1] + m.texCoord1Offset),
                       t2 = vread2(ab + (size_t)stride * vi[2] + m.texCoord1Offset);
                    tc1 = {bw[0] * t0.x + bw[1] * t1.x + bw[2] * t2.x, bw[0] * t0.y + bw[1] * t1.y + bw[2] * t2.y};
                }
                if (m.attributes & AT_TEX2) {
                    V2 t0 = vread2(ab + (size_t)stride * vi[0] + m.texCoord2Offset), t1 = vread2(ab + (size_t)stride * vi[1] + m.texCoord2Offset),
                       t2 = vread2(ab + (size_t)stride * vi[2] + m.texCoord2Offset);
                    tc2 = {bw[0] * t0.x + bw[1] * t1.x + bw[2] * t2.x, bw[0] * t0.y + bw[1] * t1.y + bw[2] * t2.y};
                }
                auto uvFor = [&](int slot) { return (m.texCoordSets >> slot) & 1u ? tc2 : tc1; };
                V3 bit = cross(normal, tangent) * (flipT ? -1.0f : 1.0f);
                M3 frame{{tangent, bit, normal}};
                V3 n;
                if (m.textures & TX_NORMAL) {
                    V2 uv = uvFor(TS_NORMAL);
                    float t[4];
                    tex_sample(sc, m.tex[TS_NORMAL], uv.x, uv.y, t);
                    n = normalize(V3{t[0] * 2.0f - 1.0f, t[1] * 2.0f - 1.0f, t[2] * 2.0f - 1.0f});
                } else {
                    n = {0.0f, 0.0f, 1.0f};
                }
                n = mul(frame, n);
                // shading-normal correction (pathtrace.cu:1433-1440)
                float dNs = dot(n, path.rD), dNg = dot(gn, path.rD);
                bool fromOut = dNs > 0.0f && dNg < 0.0f;
                bool fromIn = dNs < 0.0f && dNg > 0.0f;
                if (fromOut || fromIn) n = normalize(n - (path.rD * (dNs + (fromOut ? 0.001f : -0.001f))));
                V3 sB = normalize(cross(n, tangent));
                V3 sT2 = normalize(cross(sB, n));
                sB = sB * (flipT ? -1.0f : 1.0f);
                local.f = M3{{sT2, sB, n}};
                lastN = n;

                // material inputs (pathtrace.cu:1450-1566); the base colour
                // texture reads the normal map's coordinate set, as the
                // reference does (pathtrace.cu:1453)
                bool bcTex = (m.textures & TX_BASE_COLOR) != 0, bcFac = (m.factors & FC_BASE_COLOR) != 0;
                float alphaV = 0.0f;
                if (bcTex) {
                    V2 uv = uvFor(TS_NORMAL);
                    float t[4];
                    tex_sample(sc, m.tex[TS_BASE], uv.x, uv.y, t);
                    surf.baseColor = {t[0], t[1], t[2]};
                    alphaV = t[3];
                    if (bcFac) {
                        surf.baseColor = surf.baseColor * V3{m.baseColor[0], m.baseColor[1], m.baseColor[2]};
                        alphaV *= m.baseColor[3];
                    }
                } else if (bcFac) {
                    surf.baseColor = {m.baseColor[0], m.baseColor[1], m.baseColor[2]};
                    alphaV = m.baseColor[3];
                }
                if (m.factors & FC_CUTOFF) {
                    if (alphaV < m.alphaCutoff) skipped = true;
                } else if (m.factors & FC_BLEND) {
                    if (sobol1(q, D_ALPHA_BLEND) > alphaV) skipped = true;
                }
                bool metF = (m.factors & FC_METALNESS) != 0, rF = (m.factors & FC_ROUGHNESS) != 0;
                if (m.textures & TX_MET_ROUGH) {
                    V2 uv = uvFor(TS_MR);
                    float t[4];
                    tex_sample(sc, m.tex[TS_MR], uv.x, uv.y, t);
                    surf.metalness = t[2] * (metF ? m.metalness : 1.0f);
                    surf.roughness = t[1] * (rF ? m.roughness : 1.0f);
                } else {
                    surf.metalness = metF ? m.metalness : surf.metalness;
                    surf.roughness = rF ? m.roughness : surf.roughness;
                }
                bool trF = (m.factors & FC_TRANSMISSION) != 0;
                if (m.textures & TX_TRANSMISSION) {
                    V2 uv = uvFor(TS_TRANS);
                    float t[4];
                    tex_sample(sc, m.tex[TS_TRANS], uv.x, uv.y, t);
                    surf.transmission = t[0] * (trF ? m.transmission : 1.0f);
                } else {
                    surf.transmission = trF ? m.transmission : surf.transmission;
                }
                surf.ior = m.ior;
                if (m.flags & MF_SHEEN) {
                    surf.layers |= 2u;
                    bool shcF = (m.factors & FC_SHEEN_COLOR) != 0;
                    if (m.textures & TX_SHEEN_COLOR) {
                        V2 uv = uvFor(TS_SHC);
                        float t[4];
                        tex_sample(sc, m.tex[TS_SHC], uv.x, uv.y, t);
                        surf.sheenColor = {t[0] * (shcF ? m.sheenColor[0] : 1.0f), t[1] * (shcF ? m.sheenColor[1] : 1.0f),
                                           t[2] * (shcF ? m.sheenColor[2] : 1.0f)};
                    } else if (shcF) {
                        surf.sheenColor = {m.sheenColor[0], m.sheenColor[1], m.sheenColor[2]};
                    }
                    bool shrF = (m.factors & FC_SHEEN_ROUGH) != 0;
                    if (m.textures & TX_SHEEN_ROUGH) {
                        V2 uv = uvFor(TS_SHR);
                        float t[4];
                        tex_sample(sc, m.tex[TS_SHR], uv.x, uv.y, t);
                        float r = t[3] * (shrF ? m.sheenRoughness : 1.0f);
                        surf.sheenAlpha = r * r;
                    } else if (shrF) {
                        surf.sheenAlpha = m.sheenRoughness;
                    }
                }
            }

            // ---- NEE + BxDF (pathtrace.cu:1569-1733)
            if (!skipped) {
                DirectLight dl;
                if (!envOn && !treeOn) {
                    dl.occluded = true;
                } else {
                    V3 rnd = sobol3(q, D_LIGHT);
                    const bool canTransmit = surf.transmission != 0.0f;
                    uint32_t type = LT_NONE, index = 0;
                    float treePdf = 0.0f;
                    light_tree_sample(C.tree, rnd.z, hitPos, lastN, type, index, treePdf);
                    V3 lro{0, 0, 0};
                    float dToL = 0.0f, lightPdf = 0.0f;
                    if (type == LT_TRIANGLE) {
                        const orc_emissive_tri& tri = sc->lights[index];
                        V3 a{tri.vertices[0], tri.vertices[1], tri.vertices[2]};
                        V3 b{tri.vertices[3], tri.vertices[4], tri.vertices[5]};
                        V3 c{tri.vertices[6], tri.vertices[7], tri.vertices[8]};
                        V3 n = cross(b - a, c - a);
                        float nl = length(n);
                        n = n / nl;
                        float area = nl / 2.0f;
                        float bw[3];
                        V3 lp = tri_sample({rnd.x, rnd.y}, a, b, c, bw[0], bw[1], bw[2]);
                        V3 sr = lp - hitPos;
                        const bool inSample = dot(lastN, sr) < 0.0f;
                        (void)canTransmit;  // the reference's occlusion test here is overwritten below
                        lro = offset_point(inSample ? hitPos : hitInterp, inSample ? -gn : gn);
                        V3 rl = lp - lro;
                        float s2 = rl.x * rl.x + rl.y * rl.y + rl.z * rl.z;
                        dToL = std::sqrt(s2);
                        dl.dir = rl / dToL;
                        float lcos = -dot(n, dl.dir);
                        dl.occluded = lcos <= 0.0f;
                        lightPdf = (1.0f / area) * s2 / lcos;
                        dl.spectrum = triangle_emission(sc, tri, bw[0], bw[1], bw[2], path.wl);
                    } else if (type == LT_SHAPE) {
                        // a disk point (area measure) or a direction in the cone a
                        // sphere subtends (solid-angle measure)
                        const orc_shape_light& sl = sc->shapes[index - sc->nlights];
                        V3 lp;
                        bool inside = false;
                        if (sl.type == 0u) {
                            M3 f = quat_to_mat(sl.frame);
                            V2 xy = concentric({rnd.x, rnd.y});
                            lp = shape_center(sl) + (f.c[0] * xy.x + f.c[1] * xy.y) * sl.radius;
                        } else {
                            lp = sphere_sample(sl, hitPos, {rnd.x, rnd.y}, lightPdf, inside);
                        }
                        if (inside) {
                            dl.occluded = true;
                        } else {
                            const bool inSample = dot(lastN, lp - hitPos) < 0.0f;
                            lro = offset_point(inSample ? hitPos : hitInterp, inSample ? -gn : gn);
                            V3 rl = lp - lro;
                            float s2 = rl.x * rl.x + rl.y * rl.y + rl.z * rl.z;
                            dToL = std::sqrt(s2);
                            dl.dir = rl / dToL;
                            if (sl.type == 0u) {
                                float lcos = -dot(quat_to_mat(sl.frame).c[2], dl.dir);
                                dl.occluded = lcos <= 0.0f;
                                lightPdf = (1.0f / (PI * sl.radius * sl.radius)) * s2 / lcos;
                            }
                            dl.spectrum = shape_emission(sc, sl, path.wl);
                        }
                    } else if (type == LT_SKY) {
                        float W = (float)sc->envWidth, H = (float)sc->envHeight;
                        float fcx = std::min(rnd.x * (W - 1.0f), W - 1.0001f);
                        float fcy = std::min(rnd.y * (H - 1.0f), H - 1.0001f);
                        float flx = std::floor(fcx), fly = std::floor(fcy);
                        float frx = fcx - flx, fry = fcy - fly;
                        uint32_t ix = (uint32_t)flx, iy = (uint32_t)fly;
                        float m0 = sc->envMarg[iy], m1 = sc->envMarg[iy + 1];
                        // pathtrace.cu:1670 forms the row offset as a float product; where
                        // it rounds (W with a large odd factor, m0 * W above 2^24) the read
                        // could pass the table's end, so the offset is capped at the last row
                        size_t row = std::min((size_t)(uint32_t)(m0 * W), (size_t)(sc->envHeight - 1) * sc->envWidth);
                        float c0 = sc->envCond[row + ix], c1 = sc->envCond[row + ix + 1];
                        float ux = mixf(c0, c1, frx) * (1.0f / W);
                        float uy = mixf(m0, m1, fry) * (1.0f / H);
                        float phi = 2.0f * PI * ux, theta = PI * uy;
                        dl.dir = {-std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
                        float rgb[3];
                        env_lookup(sc, ux, uy, rgb);
                        dl.spectrum = rgb_spec(sc, {rgb[0], rgb[1], rgb[2]}, path.wl);
                        lightPdf = (1.0f / (4.0f * PI)) * ((rgb[0] + rgb[1] + rgb[2]) / 3.0f) * std::sin(theta) / sc->envIntegral;
                        dToL = FLT_MAX;
                        lro = offset_point(hitInterp, backHit ? -gn : gn);
                    } else {
                        dl.occluded = true;
                    }
                    dl.pdf = lightPdf * treePdf;
                    if (!dl.occluded) {
                        ++st.shadow;
                        CpuBvh::Hit sh{};
                        const float tmax = dToL - 0.1f;
                        bool blocked = false;
                        for (uint32_t k = 0; k < sc->nshapes && !blocked; ++k)
                            blocked = shape_t(sc->shapes[k], lro, dl.dir, tmax) < tmax;
                        dl.occluded = blocked || C.bvh.trace(lro, dl.dir, tmax, true, sh);
                    }
                }
                if (M->bxdf == 0) bxdf_conductor(sc, *M, dl, q, local, path);
                else if (M->bxdf == 2) bxdf_complex(sc, dl, q, local, surf, path);
                else bxdf_dielectric(sc, *M, dl, q, local, path);
                // RBH_RENDER_NO_DROP (flag 16, no reference counterpart): an
                // invalid BxDF sample ends the path and keeps its radiance
                // instead of discarding the whole sample
                if ((P->flags & 16u) && (path.flags & PF_PATH_TERMINATED))
                    path.flags = (path.flags & ~PF_PATH_TERMINATED) | PF_FINISHED;
            }

            // ---- Russian roulette (pathtrace.cu:1736-1744)
            {
                float tMax = smax(path.T) * path.refrScale;
                if (tMax < 1.0f && depth > 0) {
                    float qq = std::max(0.0f, 1.0f - tMax);
                    qq = qq * qq;
                    if (sobol1(q, D_ROULETTE) < qq) { finished = true; break; }
                    path.T = path.T / (1.0f - qq);
                }
            }
            // ---- next origin and direction fix (pathtrace.cu:1746-1766)
            {
                const bool refracted = (path.flags & PF_RAY_REFRACTED) != 0;
                const bool tri = (path.flags & PF_TRIANGULAR_GEOMETRY) != 0;
                const bool in = refracted || skipped;
                V3 org = (in == backHit) && tri ? hitInterp : hitPos;
                V3 off = in == backHit ? gn : -gn;
                ro = offset_point(org, off);
                float dNgRo = dot(gn, path.rD);
                if (dNgRo > 0.0f && (backHit != refracted) && !skipped) path.rD = normalize(path.rD - (gn * (dNgRo + 0.001f)));
            }
            // ---- depth limits (pathtrace.cu:1768-1779)
            {
                bool cont = ++depth < P->maxPathDepth;
                if (!(path.flags & PF_CURRENT_HIT_SPECULAR) && !skipped)
                    cont = (path.flags & PF_REFRACTION_HAPPENED) ? depth < P->maxTransmittedPathDepth : depth < P->maxReflectedPathDepth;
                if (!cont) path.flags |= PF_FINISHED;
            }
            // ---- updateStateFlags (pathtrace.cu:954-972) + advanceBounce
            {
                uint32_t excl = PF_CURRENT_HIT_SPECULAR | PF_PREVIOUS_HIT_SPECULAR | PF_TRIANGULAR_GEOMETRY | PF_RIGHT_HANDED_FRAME | PF_RAY_REFRACTED;
                uint32_t incl = 0;
                if (path.flags & PF_CURRENT_HIT_SPECULAR) incl |= PF_PREVIOUS_HIT_SPECULAR; else incl |= PF_REGULARIZED;
                if (path.flags & PF_RAY_REFRACTED) incl |= PF_REFRACTION_HAPPENED;
                path.flags = (path.flags & ~excl) | incl;
                q.advance_bounce();
            }
        } while (!(path.flags & PF_FINISHED) && !(path.flags & PF_PATH_TERMINATED));
        (void)finished;
        // ---- finishPath (pathtrace.cu:1789-1809)
        q.advance_sample();
        if (!(path.flags & PF_PATH_TERMINATED)) {
            for (int i = 0; i < NL; ++i) path.L.v[i] = path.wl.pdf.v[i] != 0.0f ? path.L.v[i] / path.wl.pdf.v[i] : 0.0f;
            double xyz[3];
            to_xyz(C.cA, C.cB, C.cC, path.wl, path.L, xyz);
            // a non-finite sample is added as black (the product path does the same and counts it)
            if (std::isfinite(xyz[0]) && std::isfinite(xyz[1]) && std::isfinite(xyz[2])) {
                res[0] += xyz[0]; res[1] += xyz[1]; res[2] += xyz[2];
            }
            res[3] += 1.0;
        } else {
            res[3] += 1.0;
        }
    }
}

}  // namespace

int render(const orc_scene* sc, const orc_params* p, double* accum, int threads, orc_render_stats* stats) {
    if (!sc || !p || !accum) return -1;
    Ctx C;
    C.sc = sc;
    C.p = p;
    const auto tb0 = std::chrono::steady_clock::now();
    C.bvh.build(sc->verts, sc->indices, sc->ntris);
    const auto tb1 = std::chrono::steady_clock::now();
    C.tree.nodes = (sc->nodes && sc->nnodes) ? (const PackedNode*)sc->nodes : nullptr;
    C.tree.lightPointers = sc->lightPointers;
    C.tree.bitmasks = sc->bitmasks;
    C.tree.envEnabled = sc->envEnabled != 0;
    C.cA = sc->xyz; C.cB = sc->xyz + NDENSE; C.cC = sc->xyz + 2 * NDENSE;
    uint32_t x0 = p->cropW ? p->cropX : 0, y0 = p->cropH ? p->cropY : 0;
    uint32_t cw = p->cropW ? p->cropW : p->width, ch = p->cropH ? p->cropH : p->height;
    const uint32_t TS = 16;
    uint32_t tx = (cw + TS - 1) / TS, ty = (ch + TS - 1) / TS;
    std::atomic<uint64_t> aC{0}, aS{0}, aV{0};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
        Stats st;
#pragma omp for schedule(dynamic, 1)
        for (long t = 0; t < (long)tx * ty; ++t) {
            uint32_t bx = (uint32_t)(t % tx) * TS, by = (uint32_t)(t / tx) * TS;
            for (uint32_t yy = by; yy < std::min(by + TS, ch); ++yy)
                for (uint32_t xx = bx; xx < std::min(bx + TS, cw); ++xx) {
                    uint32_t px = x0 + xx, py = y0 + yy;
                    double* r = accum + 4 * ((size_t)py * p->width + px);
                    double res[4] = {0, 0, 0, 0};
                    if (p->sampleOffset != 0) { res[0] = r[0]; res[1] = r[1]; res[2] = r[2]; res[3] = r[3]; }
                    render_pixel(C, px, py, res, st);
                    r[0] = res[0]; r[1] = res[1]; r[2] = res[2]; r[3] = res[3];
                }
        }
        aC += st.closest; aS += st.shadow; aV += st.vertices;
    }
    const auto tr1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->closestRays = aC;
        stats->shadowRays = aS;
        stats->pathVertices = aV;
        stats->bvhSeconds = std::chrono::duration<double>(tb1 - tb0).count();
        stats->renderSeconds = std::chrono::duration<double>(tr1 - tb1).count();
    }
    return 0;
}

}  // namespace orc
