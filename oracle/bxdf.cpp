// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// The two live BxDF callables of src/kernels/pathtrace.cu:
//   PureConductorBxDF   pathtrace.cu:464-547
//   ComplexSurface_BxDF pathtrace.cu:548-905
// plus software filtering of the LUT/material textures. The reference reads
// LUTs through CUDA texture units (clamp, linear, normalised coordinates,
// src/core/rendering_interface.cpp:528-622) whose interpolation weights are
// 8-bit fixed point; this oracle filters with exact FP32 weights, as does the
// product.
#include "render.hpp"

namespace orc {

static inline int addr_index(int i, int n, uint32_t mode) {
    if (mode == 0) { int r = i % n; return r < 0 ? r + n : r; }             // wrap
    if (mode == 2) {                                                         // mirror
        int p = 2 * n; int r = i % p; if (r < 0) r += p;
        return r < n ? r : p - 1 - r;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);                                 // clamp
}

// Single-channel LUT, clamp addressing, linear filter, texel centres at
// (i + 0.5) / n.
float tex_sample1(const orc_scene*, const float* lut, int nx, int ny, int nz, float u, float v, float w) {
    float x = u * nx - 0.5f, y = v * ny - 0.5f, z = w * nz - 0.5f;
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    float ax = x - fx, ay = y - fy, az = z - fz;
    int x0 = addr_index((int)fx, nx, 1), x1 = addr_index((int)fx + 1, nx, 1);
    int y0 = addr_index((int)fy, ny, 1), y1 = addr_index((int)fy + 1, ny, 1);
    auto at = [&](int xi, int yi, int zi) { return lut[((size_t)zi * ny + yi) * nx + xi]; };
    if (nz <= 1) {
        float a = at(x0, y0, 0) * (1.0f - ax) + at(x1, y0, 0) * ax;
        float b = at(x0, y1, 0) * (1.0f - ax) + at(x1, y1, 0) * ax;
        return a * (1.0f - ay) + b * ay;
    }
    int z0 = addr_index((int)fz, nz, 1), z1 = addr_index((int)fz + 1, nz, 1);
    float r[2];
    for (int k = 0; k < 2; ++k) {
        int zi = k ? z1 : z0;
        float a = at(x0, y0, zi) * (1.0f - ax) + at(x1, y0, zi) * ax;
        float b = at(x0, y1, zi) * (1.0f - ax) + at(x1, y1, zi) * ax;
        r[k] = a * (1.0f - ay) + b * ay;
    }
    return r[0] * (1.0f - az) + r[1] * az;
}

static void fetch_texel(const orc_scene* sc, const orc_texture& t, int x, int y, float o[4]) {
    const uint8_t* base = sc->texels + t.offset;
    size_t i = (size_t)y * t.width + x;
    switch (t.format) {
        case 0: { const uint8_t* p = base + 4 * i; for (int k = 0; k < 4; ++k) o[k] = p[k] / 255.0f; break; }
        case 1: { o[0] = base[i] / 255.0f; o[1] = 0.0f; o[2] = 0.0f; o[3] = 1.0f; break; }
        case 2: { const float* p = (const float*)base + 4 * i; for (int k = 0; k < 4; ++k) o[k] = p[k]; break; }
        case 3: { const float* p = (const float*)base + 3 * i; o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = 1.0f; break; }
        default: { o[0] = ((const float*)base)[i]; o[1] = 0.0f; o[2] = 0.0f; o[3] = 1.0f; break; }
    }
    if (t.srgb) for (int k = 0; k < 3; ++k) o[k] = srgb_to_linear(o[k]);
}

void tex_sample(const orc_scene* sc, int32_t ti, float u, float v, float out[4]) {
    const orc_texture& t = sc->textures[ti];
    int W = (int)t.width, H = (int)t.height;
    if (t.filter == 0) {
        int x = addr_index((int)std::floor(u * W), W, t.addr[0]);
        int y = addr_index((int)std::floor(v * H), H, t.addr[1]);
        fetch_texel(sc, t, x, y, out);
        return;
    }
    float x = u * W - 0.5f, y = v * H - 0.5f;
    float fx = std::floor(x), fy = std::floor(y);
    float ax = x - fx, ay = y - fy;
    int x0 = addr_index((int)fx, W, t.addr[0]), x1 = addr_index((int)fx + 1, W, t.addr[0]);
    int y0 = addr_index((int)fy, H, t.addr[1]), y1 = addr_index((int)fy + 1, H, t.addr[1]);
    float c00[4], c10[4], c01[4], c11[4];
    fetch_texel(sc, t, x0, y0, c00); fetch_texel(sc, t, x1, y0, c10);
    fetch_texel(sc, t, x0, y1, c01); fetch_texel(sc, t, x1, y1, c11);
    for (int k = 0; k < 4; ++k) {
        float a = c00[k] * (1.0f - ax) + c10[k] * ax;
        float b = c01[k] * (1.0f - ax) + c11[k] * ax;
        out[k] = a * (1.0f - ay) + b * ay;
    }
}

static inline Spec rgb_spec(const orc_scene* sc, V3 rgb, const Wavelengths& wl) {
    return rgb_to_spectrum(rgb, wl, sc->basis, sc->basis + NDENSE, sc->basis + 2 * NDENSE);
}

// src/kernels/pathtrace.cu:464-547
void bxdf_conductor(const orc_scene* sc, const orc_material& m, const DirectLight& dl, const QState& q,
                    const Local& local, PathIO& p) {
    V3 wo = local.to_local(-p.rD);
    V3 li = local.to_local(dl.dir);
    if (wo.z == 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
    Spec eta = dense_sample(sc->spectra + (size_t)m.iorSpectrum * NDENSE, p.wl);
    Spec k = dense_sample(sc->spectra + (size_t)m.acSpectrum * NDENSE, p.wl);
    float alpha = m.mfRoughness * m.mfRoughness;
    Alpha a{alpha, alpha};
    if (p.flags & PF_REGULARIZED) a.regularize();
    CtxW co = ctx_w(wo);
    if (alpha < 0.001f) {
        V3 wi{-wo.x, -wo.y, wo.z};
        float act = std::fabs(wi.z);
        Spec F;
        for (int i = 0; i < NL; ++i) F.v[i] = fresnel_complex1(act, eta.v[i], k.v[i]);
        p.T *= F;
        p.bxdfPDF = 1.0f;
        p.flags |= PF_CURRENT_HIT_SPECULAR;
        p.rD = normalize(local.from_local(wi));
        return;
    }
    if (!dl.occluded) {
        V3 wi = li;
        if (wo.z * wi.z > 0.0f) {
            CtxW ci = ctx_w(wi);
            V3 wm = normalize(wi + wo);
            CtxM cm = ctx_m(wm);
            float awm = std::fabs(dot(wo, wm));
            Spec F;
            for (int i = 0; i < NL; ++i) F.v[i] = fresnel_complex1(awm, eta.v[i], k.v[i]);
            Spec f = ggx_D(cm, a) * F * ggx_G(ci, co, a) / std::fabs(4.0f * wo.z * wi.z);
            float cosF = std::fabs(wi.z);
            float lpdf = vndf_pdf_bounded(wo, awm, co, cm, a) / (4.0f * awm);
            p.L += dl.spectrum * f * p.T * cosF * power_heuristic(dl.pdf, lpdf) / dl.pdf;
        }
    }
    V2 rnd = sobol2(q, D_BXDF0);
    V3 wm = vndf_sample_bounded(wo, a, rnd);
    CtxM cm = ctx_m(wm);
    V3 wi = reflect(wo, wm);
    CtxW ci = ctx_w(wi);
    if (wo.z * wi.z <= 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
    float awm = std::fabs(dot(wo, wm));
    Spec F;
    for (int i = 0; i < NL; ++i) F.v[i] = fresnel_complex1(awm, eta.v[i], k.v[i]);
    Spec f = ggx_D(cm, a) * F * ggx_G(ci, co, a) / std::fabs(4.0f * wo.z * wi.z);
    float cosF = std::fabs(wi.z);
    p.bxdfPDF = vndf_pdf_bounded(wo, awm, co, cm, a) / (4.0f * awm);
    p.T *= f * cosF / p.bxdfPDF;
    p.rD = normalize(local.from_local(wi));
}

// PureDielectric with spectral dispersion: the body pathtrace.cu:238-463
// keeps commented out, with two readings of its text: a secondary wavelength
// is terminated when its configuration is invalid (the text's condition is
// inverted), and each wavelength's transmission is divided by its own
// relative eta squared (the text divides by the last one computed).
void bxdf_dielectric(const orc_scene* sc, const orc_material& m, const DirectLight& dl, const QState& q,
                     const Local& local, PathIO& p) {
    V3 wo = local.to_local(-p.rD);
    V3 li = local.to_local(dl.dir);
    V3 rnd = sobol3(q, D_BXDF0);
    float alpha = m.mfRoughness * m.mfRoughness;
    Alpha a{alpha, alpha};
    if (p.flags & PF_REGULARIZED) a.regularize();
    const float cosO = wo.z;
    CtxW co = ctx_w(wo);
    Spec eta = dense_sample(sc->spectra + (size_t)m.iorSpectrum * NDENSE, p.wl);
    if (alpha < 0.001f || eta.v[0] == 1.0f) {
        Spec R;
        for (int i = 0; i < NL; ++i) R.v[i] = fresnel_real(cosO, eta.v[i]);
        float pr = R.v[0];
        V3 wi;
        if (rnd.z < pr) {
            wi = {-wo.x, -wo.y, wo.z};
            p.bxdfPDF = pr;
            float pdfSum = p.bxdfPDF;
            for (int i = 1; i < NL; ++i)
                if (p.wl.pdf.v[i] != 0.0f) pdfSum += R.v[i];
            p.T *= R * p.wl.active * (1.0f / pdfSum);
        } else {
            wl_terminate_all_secondary(p.wl);
            float T = clampf(1.0f - R.v[0], 0.0f, 1.0f);
            bool valid = false;
            float etaRel = 1.0f;
            wi = refract_rel(wo, {0.0f, 0.0f, 1.0f}, eta.v[0], valid, etaRel);
            if (!valid) { p.flags |= PF_PATH_TERMINATED; return; }
            p.refrScale = etaRel;
            p.bxdfPDF = std::max(0.0f, 1.0f - pr);
            p.T = p.T * (T / (etaRel * etaRel) / p.bxdfPDF);
            p.flags |= PF_RAY_REFRACTED;
        }
        p.flags |= PF_CURRENT_HIT_SPECULAR;
        p.rD = local.from_local(wi);
        return;
    }
    if (!dl.occluded) {
        V3 wi = li;
        CtxW ci = ctx_w(wi);
        float cosI = wi.z;
        float cosF = std::fabs(cosI);
        float t = cosO * cosI;
        bool refl = t > 0.0f;
        float G = ggx_G(ci, co, a);
        float lpdf = 0.0f;
        Spec f{};
        for (int i = 0; i < NL; ++i) {
            if (p.wl.pdf.v[i] == 0.0f) continue;
            float etaRel = 1.0f;
            if (!refl) etaRel = cosO > 0.0f ? eta.v[i] : 1.0f / eta.v[i];
            V3 wm = normalize(wi * etaRel + wo);
            wm = wm.z > 0.0f ? wm : -wm;
            float dmo = dot(wm, wo), dmi = dot(wm, wi);
            if (t != 0.0f && !(dmo * cosO < 0.0f || dmi * cosI < 0.0f)) {
                CtxM cm = ctx_m(wm);
                float R = fresnel_real(dmo, eta.v[i]);
                float T = clampf(1.0f - R, 0.0f, 1.0f);
                float pR = R / (R + T);
                float pT = std::max(0.0f, 1.0f - pR);
                float awm = std::fabs(dmo);
                if (refl) {
                    f.v[i] = ggx_D(cm, a) * R * G / (4.0f * cosO * cosI);
                    if (i == 0) lpdf = vndf_pdf_cap(wo, awm, co, cm, a) / (4.0f * awm) * pR;
                } else {
                    float tt = dmi + dmo / etaRel;
                    float den = tt * tt;
                    float dwmdwi = std::fabs(dmi) / den;
                    f.v[i] = ggx_D(cm, a) * T * G * std::fabs(dmo * dmi / (cosO * cosI * den)) / (etaRel * etaRel);
                    if (i == 0) lpdf = vndf_pdf_cap(wo, awm, co, cm, a) * dwmdwi * pT;
                }
            }
        }
        p.L += dl.spectrum * f * p.T * (cosF * power_heuristic(dl.pdf, lpdf) / dl.pdf);
    }
    V3 wm = vndf_sample_cap(wo, a, {rnd.x, rnd.y});
    CtxM cm = ctx_m(wm);
    float dmo = dot(wo, wm);
    float awm = std::fabs(dmo);
    float heroR = fresnel_real(dmo, eta.v[0]);
    float heroT = clampf(1.0f - heroR, 0.0f, 1.0f);
    float pR = heroR;
    Spec f{};
    float cosF = 0.0f, condCount = 0.0f, condSum = 0.0f;
    V3 wi;
    if (rnd.z < pR) {
        wi = reflect(wo, wm);
        if (wo.z * wi.z <= 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
        CtxW ci = ctx_w(wi);
        float cosI = wi.z;
        Spec R;
        R.v[0] = heroR;
        for (int i = 1; i < NL; ++i) R.v[i] = fresnel_real(dmo, eta.v[i]);
        f = ggx_D(cm, a) * R * ggx_G(ci, co, a) / (4.0f * cosO * cosI);
        float pdfTerm = vndf_pdf_cap(wo, awm, co, cm, a) / (4.0f * awm);
        p.bxdfPDF = pdfTerm * heroR;
        condCount = p.wl.active;
        condSum = p.bxdfPDF;
        for (int i = 1; i < NL; ++i)
            if (p.wl.pdf.v[i] != 0.0f) condSum += pdfTerm * R.v[i];
        cosF = std::fabs(cosI);
    } else {
        bool valid = false;
        float etaRel = 1.0f;
        wi = refract_rel(wo, wm, eta.v[0], valid, etaRel);
        if (wo.z * wi.z >= 0.0f || !valid) { p.flags |= PF_PATH_TERMINATED; return; }
        CtxW ci = ctx_w(wi);
        float cosI = wi.z;
        float dmi = dot(wm, wi);
        float tt = dmi + dmo / etaRel;
        float den = tt * tt;
        float dwmdwi = std::fabs(dmi) / den;
        float G = ggx_G(ci, co, a);
        f.v[0] = ggx_D(cm, a) * heroT * G * std::fabs(dmo * dmi / (cosO * cosI * den)) / (etaRel * etaRel);
        float pT = std::max(0.0f, 1.0f - pR);
        p.bxdfPDF = vndf_pdf_cap(wo, awm, co, cm, a) * dwmdwi * pT;
        p.refrScale = etaRel;
        condSum = p.bxdfPDF;
        cosF = std::fabs(cosI);
        for (int i = 1; i < NL; ++i) {
            if (p.wl.pdf.v[i] == 0.0f) continue;
            float er = cosO > 0.0f ? eta.v[i] : 1.0f / eta.v[i];
            V3 wmi = normalize(wi * er + wo);
            wmi = wmi.z > 0.0f ? wmi : -wmi;
            CtxM cmi = ctx_m(wmi);
            float dmoi = dot(wmi, wo), dmii = dot(wmi, wi);
            float awmi = std::fabs(dmoi);
            float secR = fresnel_real(dmoi, eta.v[i]);
            float secT = clampf(1.0f - secR, 0.0f, 1.0f);
            float ti = dmii + dmoi / er;
            float deni = ti * ti;
            float dwi = std::fabs(dmii) / deni;
            float fi = ggx_D(cmi, a) * secT * G * std::fabs(dmoi * dmii / (cosO * cosI * deni)) / (er * er);
            float pdf = vndf_pdf_cap(wo, awmi, co, cmi, a) * dwi * secT;
            if (pdf <= 0.0f || dmoi * cosO < 0.0f || dmii * cosI < 0.0f) {
                wl_terminate_secondary(p.wl, i);
            } else {
                f.v[i] = fi;
                condSum += pdf;
            }
        }
        condCount = p.wl.active;
        p.flags |= PF_RAY_REFRACTED;
    }
    if (p.bxdfPDF == 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
    p.T *= condCount * (1.0f / condSum) * f * cosF;
    p.rD = normalize(local.from_local(wi));
}

// src/kernels/pathtrace.cu:548-905
void bxdf_complex(const orc_scene* sc, const DirectLight& dl, const QState& q, const Local& local,
                  const Surface& s, PathIO& p) {
    V3 wo = local.to_local(-p.rD);
    V3 li = local.to_local(dl.dir);
    if (wo.z == 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
    CtxW co = ctx_w(wo);
    V3 wm{0, 0, 0}, wi{0, 0, 0};
    float roughness = s.roughness;
    bool perfectSpecular = false;
    if (roughness < 0.001f) { roughness = 0.001f; perfectSpecular = true; }
    float alpha = roughness * roughness;
    Alpha a{alpha, alpha};
    float eta = wo.z > 0.0f ? s.ior / 1.0f : 1.0f / s.ior;
    float f0Sr = (1.0f - eta) / (1.0f + eta);
    float f0 = f0Sr * f0Sr;
    float absCosO = std::fabs(wo.z);
    float zc = std::sqrt(std::fabs(f0Sr));
    float eCompDiel = 1.0f / tex_sample1(sc, eta > 1.0f ? sc->lutDielOuter : sc->lutDielInner, 32, 32, 32, absCosO, roughness, zc);
    float ePresDiff = 1.0f - tex_sample1(sc, eta > 1.0f ? sc->lutReflOuter : sc->lutReflInner, 32, 32, 32, absCosO, roughness, zc);
    float albCond = tex_sample1(sc, sc->lutConductor, 32, 32, 1, absCosO, roughness, 0.0f);
    float albSheen = 0.0f, ltcA = 0.0f, ltcB = 0.0f;
    if (s.layers & 2u) {
        // 4-channel sheen LUT (A, B, albedo, unused)
        float c[4];
        float x = absCosO * 32 - 0.5f, y = s.sheenAlpha * 32 - 0.5f;
        float fx = std::floor(x), fy = std::floor(y);
        float ax = x - fx, ay = y - fy;
        int x0 = addr_index((int)fx, 32, 1), x1 = addr_index((int)fx + 1, 32, 1);
        int y0 = addr_index((int)fy, 32, 1), y1 = addr_index((int)fy + 1, 32, 1);
        for (int kk = 0; kk < 4; ++kk) {
            auto at = [&](int xi, int yi) { return sc->lutSheen[((size_t)yi * 32 + xi) * 4 + kk]; };
            float u0 = at(x0, y0) * (1.0f - ax) + at(x1, y0) * ax;
            float u1 = at(x0, y1) * (1.0f - ax) + at(x1, y1) * ax;
            c[kk] = u0 * (1.0f - ay) + u1 * ay;
        }
        ltcA = c[0]; ltcB = c[1]; albSheen = c[2];
        if (albSheen < 0.0001f) albSheen = 0.0f;
    }

    if (!dl.occluded) {
        wi = li;
        CtxW ci = ctx_w(wi);
        const float ct_i = wi.z, ct_o = wo.z;
        const float cosF = std::fabs(ct_i);
        const bool refl = ct_i * ct_o > 0.0f;
        float etaR = 1.0f;
        if (!refl) etaR = eta;
        wm = normalize(wi * etaR + wo);
        if (wm.z < 0.0f) wm = -wm;
        CtxM cm = ctx_m(wm);
        const float dmo = dot(wm, wo), dmi = dot(wm, wi);
        const float awm = std::fabs(dot(wo, wm));
        const float Fd = schlick(f0, awm);
        const float Ds = ggx_D(cm, a);
        const float Gs = ggx_G(ci, co, a);
        const float sheen = albSheen, base = 1.0f - sheen;
        const float metal = s.metalness, transl = s.transmission;
        const float wSheen = sheen, wCSpec = base * metal, wDSpec = base * (1.0f - metal) * Fd;
        const float wDTrans = base * (1.0f - metal) * transl * (1.0f - Fd);
        const float wDiff = base * (1.0f - metal) * (1.0f - transl) * (1.0f - Fd);
        Spec flSheen = specs(0.0f);
        float pSheen = 1.0f;
        if (sheen != 0.0f) {
            float fns = 0.0f;
            ltc_eval(wo, wi, ltcA, ltcB, fns, pSheen);
            flSheen = rgb_spec(sc, s.sheenColor, p.wl) * fns;
        }
        if (refl) {
            Spec flCond = specs(0.0f);
            float flDiel = 0.0f, pSpec = 0.0f;
            if (!perfectSpecular) {
                V3 fc = f82(s.baseColor, awm) * (v3s(1.0f) + f_avg_f82(s.baseColor) * ((1.0f - albCond) / albCond));
                Spec FC = rgb_spec(sc, fc, p.wl);
                flCond = Ds * Gs * FC / std::fabs(4.0f * wo.z * wi.z);
                flDiel = Ds * Gs * eCompDiel / std::fabs(4.0f * wo.z * wi.z);
                pSpec = vndf_pdf_cap(wo, awm, co, cm, a) / (4.0f * awm);
            }
            Spec flDiff = ePresDiff * rgb_spec(sc, s.baseColor, p.wl) / PI;
            const float pDiff = cos_pdf(wi.z);
            const float pdf = wSheen * pSheen + wCSpec * pSpec + wDSpec * pSpec + wDiff * pDiff;
            const float mis = power_heuristic(dl.pdf, pdf);
            p.L += dl.spectrum * p.T * cosF * mis * (wSheen * flSheen + wCSpec * flCond + specs(wDSpec * flDiel) + wDiff * flDiff) / dl.pdf;
        } else if (s.transmission != 0.0f && !(dmo * ct_o <= 0.0f || dmi * ct_i <= 0.0f)) {
            float pTrans = 0.0f;
            Spec flTrans = specs(0.0f);
            if (!perfectSpecular) {
                const float t = dmi + dmo / eta;
                const float den = t * t;
                const float dwmdwi = std::fabs(dmi) / den;
                pTrans = vndf_pdf_cap(wo, awm, co, cm, a) * dwmdwi;
                flTrans = rgb_spec(sc, s.baseColor, p.wl) * Ds * (1.0f - Fd) * Gs * eCompDiel *
                          std::fabs(dmo * dmi / (ct_o * ct_i * den));
            }
            const float pdf = wSheen * pSheen + wDTrans * pTrans;
            const float mis = power_heuristic(dl.pdf, pdf);
            flTrans = flTrans / (eta * eta);
            p.L += dl.spectrum * p.T * cosF * mis * (wSheen * flSheen + wDTrans * flTrans) / dl.pdf;
        }
    }

    {
        V3 rnd = sobol3(q, D_BXDF0);
        const float sheen = albSheen, base = 1.0f - sheen;
        const float sheenP = sheen, baseP = base;
        const float baseCorr = 1.0f - albSheen;
        if (rnd.z > baseP) {
            float fns = 0.0f, pdfP = 0.0f;
            wi = ltc_sample(wo, ltcA, ltcB, {rnd.x, rnd.y});
            ltc_eval(wo, wi, ltcA, ltcB, fns, pdfP);
            Spec f = rgb_spec(sc, s.sheenColor, p.wl) * fns * sheen;
            p.bxdfPDF = pdfP * sheenP;
            p.T *= f * std::fabs(wi.z) / p.bxdfPDF;
        } else {
            wm = vndf_sample_cap(wo, a, {rnd.x, rnd.y});
            CtxM cm = ctx_m(wm);
            const float awm = std::fabs(dot(wo, wm));
            float Ds = ggx_D(cm, a);
            float Fd = schlick(f0, awm);
            float s2i = std::max(0.0f, 1.0f - awm * awm);
            float s2t = s2i / (eta * eta);
            if (s2t >= 1.0f) Fd = 1.0f;
            const float metal = s.metalness, transl = s.transmission;
            const float condP = base * metal;
            const float dSpecP = base * (1.0f - metal) * Fd;
            const float transP = base * transl * (1.0f - metal) * (1.0f - Fd);
            const float diffP = base * (1.0f - transl) * (1.0f - metal) * (1.0f - Fd);
            if (rnd.z < condP + dSpecP) {
                wi = normalize(reflect(wo, wm));
                CtxW ci = ctx_w(wi);
                if (wo.z * wi.z <= 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
                float Gs = ggx_G(ci, co, a);
                float spec = Ds * Gs / std::fabs(4.0f * wo.z * wi.z);
                Spec f;
                float pdfP;
                if (rnd.z < condP) {
                    V3 fc = f82(s.baseColor, awm) * (v3s(1.0f) + f_avg_f82(s.baseColor) * ((1.0f - albCond) / albCond));
                    f = spec * rgb_spec(sc, fc, p.wl) * metal;
                    pdfP = condP;
                } else {
                    f = specs(spec * Fd * eCompDiel * (1.0f - metal));
                    pdfP = dSpecP;
                }
                float cosF = std::fabs(wi.z);
                p.bxdfPDF = vndf_pdf_cap(wo, awm, co, cm, a) / (4.0f * awm) * pdfP;
                p.T *= f * cosF * baseCorr / p.bxdfPDF;
            } else if (rnd.z < condP + dSpecP + transP) {
                wi = refract(wo, wm, eta);
                p.refrScale *= eta;
                if (wo.z * wi.z >= 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
                CtxW ci = ctx_w(wi);
                float Gs = ggx_G(ci, co, a);
                const float ct_i = wi.z, ct_o = wo.z;
                float dmo = dot(wm, wo), dmi = dot(wm, wi);
                float t = dmi + dmo / eta;
                float den = t * t;
                float dwmdwi = std::fabs(dmi) / den;
                float f = Ds * (1.0f - Fd) * Gs * (1.0f - metal) * transl * std::fabs(dmo * dmi / (ct_o * ct_i * den));
                f /= (eta * eta);
                p.bxdfPDF = vndf_pdf_cap(wo, awm, co, cm, a) * dwmdwi * transP;
                float cosF = std::fabs(ct_i);
                p.T *= rgb_spec(sc, s.baseColor, p.wl) * f * eCompDiel * cosF * baseCorr / p.bxdfPDF;
                p.flags |= PF_RAY_REFRACTED;
            } else {
                V2 rd = sobol2(q, D_BXDF1);
                wi = cos_sample(rd);
                if (wo.z * wi.z <= 0.0f) { p.flags |= PF_PATH_TERMINATED; return; }
                float cosF = std::fabs(wi.z);
                Spec f = rgb_spec(sc, s.baseColor, p.wl) / PI * (1.0f - metal) * (1.0f - transl);
                p.bxdfPDF = cos_pdf(wi.z) * diffP;
                p.T *= f * ePresDiff * cosF * baseCorr / p.bxdfPDF;
            }
        }
    }
    if (p.bxdfPDF <= 0.0f || std::isinf(p.bxdfPDF)) { p.flags |= PF_PATH_TERMINATED; return; }
    p.rD = normalize(local.from_local(wi));
}

}  // namespace orc
