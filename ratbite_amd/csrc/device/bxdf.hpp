// BxDFs of the hot path: GGX microfacet pieces (src/device/surface.h:92-378),
// the OpenPBR "ComplexSurface" callable (src/kernels/pathtrace.cu:548-905) and
// the spectral "PureConductor" callable (pathtrace.cu:464-547).
//
// The callables are split here into the two halves the wavefront needs:
// the NEE evaluation (whose contribution the shadow-ray kernel adds only if
// the light is visible) and the BSDF sampling half. Their arithmetic is the
// reference's; the split only defers the visibility test.
#pragma once
#include "geom.hpp"
#include "spectral.hpp"
#include "texture.hpp"
#include "qrng.hpp"

namespace rbh {
namespace dev {

constexpr uint32_t PF_PREV_SPECULAR = 1u << 0, PF_CUR_SPECULAR = 1u << 1, PF_REGULARIZED = 1u << 2,
                   PF_RAY_REFRACTED = 1u << 3, PF_TERMINATED = 1u << 4, PF_TRIANGULAR = 1u << 5,
                   PF_RIGHT_HANDED = 1u << 6, PF_REFRACTION_HAPPENED = 1u << 7, PF_FINISHED = 1u << 8;

// --------------------------------------------------------------- GGX parts
__device__ __forceinline__ float tan2_of(V3 w) {
    const float c2 = w.z * w.z;
    return fmaxf(0.0f, 1.0f - c2) / c2;
}
__device__ __forceinline__ void cs_phi(V3 w, float& cp, float& sp) {
    const float st = sqrtf(fmaxf(0.0f, 1.0f - w.z * w.z));
    cp = st == 0.0f ? 1.0f : clampf(w.x / st, -1.0f, 1.0f);
    sp = st == 0.0f ? 0.0f : clampf(w.y / st, -1.0f, 1.0f);
}
struct DirCtx { float tan2, cp, sp; };
__device__ __forceinline__ DirCtx dir_ctx(V3 w) {
    DirCtx c;
    c.tan2 = tan2_of(w);
    cs_phi(w, c.cp, c.sp);
    return c;
}
struct MCtx { float cos2, tan2, cp, sp; };
__device__ __forceinline__ MCtx m_ctx(V3 m) {
    MCtx c;
    c.cos2 = m.z * m.z;
    c.tan2 = tan2_of(m);
    cs_phi(m, c.cp, c.sp);
    return c;
}
struct Rough { float ax, ay; };
// surface.h:138-146
__device__ __forceinline__ float smith_lambda(const Rough& a, const DirCtx& c) {
    if (isinf(c.tan2)) return 0.0f;
    const float x = c.cp * a.ax, y = c.sp * a.ay;
    return (sqrtf(1.0f + (x * x + y * y) * c.tan2) - 1.0f) / 2.0f;
}
__device__ __forceinline__ float smith_G1(const DirCtx& o, const Rough& a) { return 1.0f / (1.0f + smith_lambda(a, o)); }
__device__ __forceinline__ float smith_G(const DirCtx& i, const DirCtx& o, const Rough& a) {
    return 1.0f / (1.0f + smith_lambda(a, i) + smith_lambda(a, o));
}
// surface.h:161-174
__device__ __forceinline__ float ggx_D(const MCtx& m, const Rough& a) {
    if (isinf(m.tan2)) return 0.0f;
    const float c4 = m.cos2 * m.cos2;
    if (c4 < 1e-16f) return 0.0f;
    const float x = m.cp / a.ax, y = m.sp / a.ay;
    const float ope = 1.0f + m.tan2 * (x * x + y * y);
    return 1.0f / (3.14159265358979f * a.ax * a.ay * c4 * ope * ope);
}
// Complex Fresnel of a conductor, one wavelength (surface.h:176-195)
__device__ __forceinline__ float fresnel_conductor(float ct, float er, float ei) {
    ct = clampf(ct, 0.0f, 1.0f);
    const float s2 = 1.0f - ct * ct;
    const float e2r = er * er - ei * ei, e2i = 2.0f * er * ei;
    const float den = e2r * e2r + e2i * e2i;
    const float zr = 1.0f - s2 * e2r / den, zi = s2 * e2i / den;   // 1 - s2/eta^2
    const float mag = sqrtf(sqrtf(zr * zr + zi * zi));
    float sa, ca;
    sincosf(0.5f * atan2f(zi, zr), &sa, &ca);
    const float tr = mag * ca, ti = mag * sa;                        // cos(theta_t)
    const float ar = er * ct - tr, ai = ei * ct - ti, br = er * ct + tr, bi = ei * ct + ti;
    const float rpa = (ar * ar + ai * ai) / (br * br + bi * bi);
    const float er2 = er * tr - ei * ti, ei2 = er * ti + ei * tr;    // eta * cos(theta_t)
    const float cr = ct - er2, dr = ct + er2;
    const float rpe = (cr * cr + ei2 * ei2) / (dr * dr + ei2 * ei2);
    return (rpa + rpe) / 2.0f;
}
// (1 - mu)^5 of surface.h:245-252 (the reference's pow(x, 5)) as a double
// product rounded once: a few FP64 multiplies instead of a general powf
__device__ __forceinline__ float pow5(float x) {
    const double a = (double)x * (double)x;
    return (float)(a * a * (double)x);
}
__device__ __forceinline__ float schlick(float f0, float mu) { return f0 + (1.0f - f0) * pow5(1.0f - mu); }
// F82 with specularColor = 1 is Schlick (surface.h:254-264), times the
// Turquin multiple-scattering factor (1 + F_avg (1 - E) / E)
__device__ __forceinline__ V3 conductor_color(V3 f0, float mu, float E) {
    const float p = pow5(1.0f - mu);
    const V3 fav = 0.877105f * f0 + (V3{0.0648148f, 0.0648148f, 0.0648148f} + 0.0752755f * f0) - V3{0.0171958f, 0.0171958f, 0.0171958f};
    const float k = (1.0f - E) / E;
    return V3{f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p} *
           V3{1.0f + fav.x * k, 1.0f + fav.y * k, 1.0f + fav.z * k};
}

// Dupuy & Benyoub 2023 spherical-cap VNDF (surface.h:308-322)
__device__ __forceinline__ V3 vndf_cap(V3 wo, const Rough& a, V2 u) {
    V3 ws = nrm(V3{wo.x * a.ax, wo.y * a.ay, wo.z});
    if (ws.z < 0.0f) ws = -ws;
    const float z = fmaf(1.0f - u.y, 1.0f + ws.z, -ws.z);
    const float st = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    float s, c;
    sincosf(2.0f * 3.14159265358979f * u.x, &s, &c);
    const V3 m = ws + V3{st * c, st * s, z};
    return nrm(V3{m.x * a.ax, m.y * a.ay, m.z});
}
// Eto & Tokuyoshi bounded VNDF (surface.h:323-340)
__device__ __forceinline__ V3 vndf_bounded(V3 wo, const Rough& a, V2 u) {
    const V3 ws = nrm(V3{wo.x * a.ax, wo.y * a.ay, wo.z});
    const float a1 = fminf(a.ax, a.ay);
    const float sl = 1.0f + sqrtf(wo.x * wo.x + wo.y * wo.y);
    const float a2 = a1 * a1, s2 = sl * sl;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
    const float b = wo.z > 0.0f ? k * ws.z : ws.z;
    const float z = fmaf(1.0f - u.y, 1.0f + b, -b);
    const float st = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
    float s, c;
    sincosf(2.0f * 3.14159265358979f * u.x, &s, &c);
    const V3 m = ws + V3{st * c, st * s, z};
    return nrm(V3{m.x * a.ax, m.y * a.ay, m.z});
}
__device__ __forceinline__ float vndf_cap_pdf(V3 wo, float awm, const DirCtx& o, const MCtx& m, const Rough& a) {
    return smith_G1(o, a) / fabsf(wo.z) * ggx_D(m, a) * awm;
}
__device__ __forceinline__ float vndf_bounded_pdf(V3 wo, float awm, const MCtx& m, const Rough& a) {
    const float ndf = ggx_D(m, a);
    const float ox = wo.x * a.ax, oy = wo.y * a.ay;
    const float len2 = ox * ox + oy * oy;
    const float t = sqrtf(len2 + wo.z * wo.z);
    if (wo.z >= 0.0f) {
        const float a1 = clampf(fminf(a.ax, a.ay), 0.0f, 1.0f);
        const float sl = 1.0f + sqrtf(wo.x * wo.x + wo.y * wo.y);
        const float a2 = a1 * a1, s2 = sl * sl;
        const float k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
        return ndf * 2.0f * awm / (k * wo.z + t);
    }
    return ndf * (t - wo.z) * 2.0f * awm / len2;
}

// Zeltner 2022 LTC sheen (surface.h:53-88)
__device__ __forceinline__ void ltc_phi(V3 wo, float& cp, float& sp) {
    const float st = sqrtf(fmaxf(0.0f, 1.0f - wo.z * wo.z));
    cp = st == 0.0f ? 1.0f : clampf(wo.x / st, -1.0f, 1.0f);
    sp = st == 0.0f ? 0.0f : clampf(wo.y / st, -1.0f, 1.0f);
}
__device__ __forceinline__ V3 ltc_sample(V3 wo, float A, float B, V2 u) {
    const V3 w0 = cosine_hemisphere(u);
    const V3 ws{w0.x / A - w0.z * B / A, w0.y / A, w0.z};
    float cp, sp;
    ltc_phi(wo, cp, sp);
    const V3 wi{cp * ws.x - sp * ws.y, cp * ws.y + sp * ws.x, ws.z};
    return nrm(wo.z > 0.0f ? wi : -wi);
}
__device__ __forceinline__ void ltc_eval(V3 wo, V3 wi, float A, float B, float& f, float& pdf) {
    float cp, sp;
    ltc_phi(wo, cp, sp);
    const V3 ws{cp * wi.x + sp * wi.y, cp * wi.y - sp * wi.x, wi.z};
    V3 w0{A * ws.x + B * ws.z, A * ws.y, ws.z};
    const float l = len(w0);
    w0 = w0 / l;
    const float r = (w0.z / 3.14159265358979f) * (A * A / (l * l * l));
    f = r / fabsf(wi.z);
    pdf = r;
}

// Dielectric Fresnel reflectance (microfacet::FReal, src/device/surface.h:
// 197-216): eta is inverted for a cosine below zero; 1 under total internal
// reflection.
__device__ __forceinline__ float fresnel_real(float c, float eta) {
    c = clampf(c, -1.0f, 1.0f);
    if (c < 0.0f) {
        eta = 1.0f / eta;
        c = -c;
    }
    const float s2 = 1.0f - c * c;
    const float s2t = s2 / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - s2t));
    const float rParl = (eta * c - ct) / (eta * c + ct);
    const float rPerp = (c - eta * ct) / (c + eta * ct);
    return (rParl * rParl + rPerp * rPerp) / 2.0f;
}

// ------------------------------------------------------------------ inputs
struct Surface {
    bool sheen;
    V3 base;
    float metalness, ior, roughness, transmission;
    V3 sheenColor;
    float sheenAlpha;
};
struct Light {
    S4 spectrum;
    V3 dir;
    float pdf;
    bool occluded;
};
struct PathIO {
    S4 T;
    Wl wl;
    float bxdfPdf;
    V3 rd;
    uint32_t flags;
    float refr;
};
struct Luts { const float* __restrict__ t; };

// Secondary wavelengths a dispersive refraction has terminated
// (SampledWavelengths::terminateSecondary / terminateAllSecondary,
// src/device/spectral.h:250-276): bit i-1 of PF_WL_TERMINATED marks
// wavelength i (1..3). The hero wavelength is never terminated. The
// accumulation gives a terminated wavelength pdf 0 and scales the active ones
// by (active count) / (initially active count), the product of the
// reference's per-termination factors.
constexpr uint32_t PF_WL_SHIFT = 9, PF_WL_TERMINATED = 7u << PF_WL_SHIFT;
__device__ __forceinline__ bool wl_active(const Wl& wl, uint32_t flags, int i) {
    return wl.pdf.v[i] != 0.0f && (i == 0 || !((flags >> (PF_WL_SHIFT + i - 1)) & 1u));
}
__device__ __forceinline__ float wl_active_count(const Wl& wl, uint32_t flags) {
    float n = 0.0f;
    for (int i = 0; i < kNL; ++i) n += wl_active(wl, flags, i) ? 1.0f : 0.0f;
    return n;
}
__device__ __forceinline__ void wl_terminate(uint32_t& flags, int i) { flags |= 1u << (PF_WL_SHIFT + i - 1); }
__device__ __forceinline__ void wl_terminate_all(uint32_t& flags) { flags |= PF_WL_TERMINATED; }
// The pdfs the accumulation divides by.
__device__ __forceinline__ S4 wl_effective_pdf(const Wl& wl, uint32_t flags) {
    S4 pdf = wl.pdf;
    if (!(flags & PF_WL_TERMINATED)) return pdf;
    float initial = 0.0f;
    for (int i = 0; i < kNL; ++i) initial += wl.pdf.v[i] != 0.0f ? 1.0f : 0.0f;
    const float keep = wl_active_count(wl, flags) / initial;
    for (int i = 0; i < kNL; ++i) pdf.v[i] = wl_active(wl, flags, i) ? pdf.v[i] * keep : 0.0f;
    return pdf;
}

// Conductor callable (pathtrace.cu:464-547). Adds the unoccluded NEE term
// to Lnee and samples the next direction.
__device__ inline void conductor_bxdf(const float* __restrict__ eta_tab, const float* __restrict__ k_tab, float rough,
                                      const Light& dl, const Qrng& q, const Frame& fr, PathIO& p, S4& Lnee) {
    const V3 wo = fr.to_local(-p.rd);
    const V3 li = fr.to_local(dl.dir);
    if (wo.z == 0.0f) { p.flags |= PF_TERMINATED; return; }
    const S4 eta = dense4(eta_tab, p.wl), kk = dense4(k_tab, p.wl);
    const float alpha = rough * rough;
    Rough a{alpha, alpha};
    if (p.flags & PF_REGULARIZED) {
        if (a.ax < 0.3f) a.ax = clampf(2.0f * a.ax, 0.1f, 0.3f);
        if (a.ay < 0.3f) a.ay = clampf(2.0f * a.ay, 0.1f, 0.3f);
    }
    const DirCtx co = dir_ctx(wo);
    if (alpha < 0.001f) {
        const V3 wi{-wo.x, -wo.y, wo.z};
        const float act = fabsf(wi.z);
        S4 F;
        for (int i = 0; i < kNL; ++i) F.v[i] = fresnel_conductor(act, eta.v[i], kk.v[i]);
        p.T *= F;
        p.bxdfPdf = 1.0f;
        p.flags |= PF_CUR_SPECULAR;
        p.rd = nrm(fr.to_world(wi));
        return;
    }
    if (!dl.occluded && wo.z * li.z > 0.0f) {
        const DirCtx ci = dir_ctx(li);
        const V3 wm = nrm(li + wo);
        const MCtx cm = m_ctx(wm);
        const float awm = fabsf(dot(wo, wm));
        S4 F;
        for (int i = 0; i < kNL; ++i) F.v[i] = fresnel_conductor(awm, eta.v[i], kk.v[i]);
        const S4 f = ggx_D(cm, a) * F * smith_G(ci, co, a) / fabsf(4.0f * wo.z * li.z);
        const float lpdf = vndf_bounded_pdf(wo, awm, cm, a) / (4.0f * awm);
        Lnee += dl.spectrum * f * p.T * fabsf(li.z) * power_heuristic(dl.pdf, lpdf) / dl.pdf;
    }
    const V3 wm = vndf_bounded(wo, a, q.s2(QD_BXDF0));
    const MCtx cm = m_ctx(wm);
    const V3 wi = reflect_about(wo, wm);
    const DirCtx ci = dir_ctx(wi);
    if (wo.z * wi.z <= 0.0f) { p.flags |= PF_TERMINATED; return; }
    const float awm = fabsf(dot(wo, wm));
    S4 F;
    for (int i = 0; i < kNL; ++i) F.v[i] = fresnel_conductor(awm, eta.v[i], kk.v[i]);
    const S4 f = ggx_D(cm, a) * F * smith_G(ci, co, a) / fabsf(4.0f * wo.z * wi.z);
    p.bxdfPdf = vndf_bounded_pdf(wo, awm, cm, a) / (4.0f * awm);
    p.T *= f * fabsf(wi.z) / p.bxdfPdf;
    p.rd = nrm(fr.to_world(wi));
}

// PureDielectric callable with spectral dispersion: the body the reference
// keeps commented out (pathtrace.cu:238-463), a rough or smooth dielectric
// whose index of refraction is a dense spectrum sampled at the path's
// wavelengths. Reflection keeps every wavelength and weights by the hero
// MIS over the active wavelengths' Fresnel terms (smooth) or pdfs (rough);
// a smooth refraction follows the hero wavelength alone and terminates the
// secondaries; a rough refraction keeps each secondary whose own half vector
// for the sampled direction is a valid configuration and terminates the
// others. Two places where the commented-out text cannot be what it means
// are read as intended: a secondary is terminated when its configuration is
// INVALID (the text terminates the valid ones), and each wavelength's
// transmission is divided by its own relative eta squared (the text divides
// all of them by the last one computed); the refraction scale is the hero's.
__device__ inline void dielectric_bxdf(const float* __restrict__ eta_tab, float rough, const Light& dl, const Qrng& q,
                                       const Frame& fr, PathIO& p, S4& Lnee) {
    const V3 wo = fr.to_local(-p.rd);
    const V3 li = fr.to_local(dl.dir);
    const V3 rnd = q.s3(QD_BXDF0);
    const float alpha = rough * rough;
    Rough a{alpha, alpha};
    if (p.flags & PF_REGULARIZED) {
        if (a.ax < 0.3f) a.ax = clampf(2.0f * a.ax, 0.1f, 0.3f);
        if (a.ay < 0.3f) a.ay = clampf(2.0f * a.ay, 0.1f, 0.3f);
    }
    const float cosO = wo.z;
    const DirCtx co = dir_ctx(wo);
    const S4 eta = dense4(eta_tab, p.wl);
    if (alpha < 0.001f || eta.v[0] == 1.0f) {
        S4 R;
        for (int i = 0; i < kNL; ++i) R.v[i] = fresnel_real(cosO, eta.v[i]);
        const float pr = R.v[0];
        V3 wi;
        if (rnd.z < pr) {
            wi = V3{-wo.x, -wo.y, wo.z};
            p.bxdfPdf = pr;
            float pdfSum = p.bxdfPdf;
            for (int i = 1; i < kNL; ++i)
                if (wl_active(p.wl, p.flags, i)) pdfSum += R.v[i];
            p.T *= R * wl_active_count(p.wl, p.flags) * (1.0f / pdfSum);
        } else {
            wl_terminate_all(p.flags);
            const float T = clampf(1.0f - R.v[0], 0.0f, 1.0f);
            bool valid = false;
            float etaRel = 1.0f;
            wi = refract_rel(wo, V3{0.0f, 0.0f, 1.0f}, eta.v[0], valid, etaRel);
            if (!valid) { p.flags |= PF_TERMINATED; return; }
            p.refr = etaRel;
            p.bxdfPdf = fmaxf(0.0f, 1.0f - pr);
            p.T *= s4(T / (etaRel * etaRel) / p.bxdfPdf);
            p.flags |= PF_RAY_REFRACTED;
        }
        p.flags |= PF_CUR_SPECULAR;
        p.rd = fr.to_world(wi);
        return;
    }
    if (!dl.occluded) {
        const V3 wi = li;
        const DirCtx ci = dir_ctx(wi);
        const float cosI = wi.z;
        const float cosF = fabsf(cosI);
        const float t = cosO * cosI;
        const bool reflect = t > 0.0f;
        const float G = smith_G(ci, co, a);
        float lpdf = 0.0f;
        S4 f = s4(0.0f);
        for (int i = 0; i < kNL; ++i) {
            if (!wl_active(p.wl, p.flags, i)) continue;
            float etaRel = 1.0f;
            if (!reflect) etaRel = cosO > 0.0f ? eta.v[i] : 1.0f / eta.v[i];
            V3 wm = nrm(wi * etaRel + wo);
            wm = wm.z > 0.0f ? wm : -wm;
            const float dmo = dot(wm, wo), dmi = dot(wm, wi);
            if (t != 0.0f && !(dmo * cosO < 0.0f || dmi * cosI < 0.0f)) {
                const MCtx cm = m_ctx(wm);
                const float R = fresnel_real(dmo, eta.v[i]);
                const float T = clampf(1.0f - R, 0.0f, 1.0f);
                const float pR = R / (R + T);
                const float pT = fmaxf(0.0f, 1.0f - pR);
                const float awm = fabsf(dmo);
                if (reflect) {
                    f.v[i] = ggx_D(cm, a) * R * G / (4.0f * cosO * cosI);
                    if (i == 0) lpdf = vndf_cap_pdf(wo, awm, co, cm, a) / (4.0f * awm) * pR;
                } else {
                    const float tt = dmi + dmo / etaRel;
                    const float den = tt * tt;
                    const float dwmdwi = fabsf(dmi) / den;
                    f.v[i] = ggx_D(cm, a) * T * G * fabsf(dmo * dmi / (cosO * cosI * den)) / (etaRel * etaRel);
                    if (i == 0) lpdf = vndf_cap_pdf(wo, awm, co, cm, a) * dwmdwi * pT;
                }
            }
        }
        Lnee += dl.spectrum * f * p.T * (cosF * power_heuristic(dl.pdf, lpdf) / dl.pdf);
    }
    V3 wm = vndf_cap(wo, a, V2{rnd.x, rnd.y});
    MCtx cm = m_ctx(wm);
    float dmo = dot(wo, wm);
    float awm = fabsf(dmo);
    const float heroR = fresnel_real(dmo, eta.v[0]);
    const float heroT = clampf(1.0f - heroR, 0.0f, 1.0f);
    const float pR = heroR;
    S4 f = s4(0.0f);
    float cosF = 0.0f, condCount = 0.0f, condSum = 0.0f;
    V3 wi;
    if (rnd.z < pR) {
        wi = reflect_about(wo, wm);
        if (wo.z * wi.z <= 0.0f) { p.flags |= PF_TERMINATED; return; }
        const DirCtx ci = dir_ctx(wi);
        const float cosI = wi.z;
        S4 R;
        R.v[0] = heroR;
        for (int i = 1; i < kNL; ++i) R.v[i] = fresnel_real(dmo, eta.v[i]);
        f = ggx_D(cm, a) * R * smith_G(ci, co, a) / (4.0f * cosO * cosI);
        const float pdfTerm = vndf_cap_pdf(wo, awm, co, cm, a) / (4.0f * awm);
        p.bxdfPdf = pdfTerm * heroR;
        condCount = wl_active_count(p.wl, p.flags);
        condSum = p.bxdfPdf;
        for (int i = 1; i < kNL; ++i)
            if (wl_active(p.wl, p.flags, i)) condSum += pdfTerm * R.v[i];
        cosF = fabsf(cosI);
    } else {
        bool valid = false;
        float etaRel = 1.0f;
        wi = refract_rel(wo, wm, eta.v[0], valid, etaRel);
        if (wo.z * wi.z >= 0.0f || !valid) { p.flags |= PF_TERMINATED; return; }
        const DirCtx ci = dir_ctx(wi);
        const float cosI = wi.z;
        float dmi = dot(wm, wi);
        float tt = dmi + dmo / etaRel;
        float den = tt * tt;
        float dwmdwi = fabsf(dmi) / den;
        const float G = smith_G(ci, co, a);
        f.v[0] = ggx_D(cm, a) * heroT * G * fabsf(dmo * dmi / (cosO * cosI * den)) / (etaRel * etaRel);
        const float pT = fmaxf(0.0f, 1.0f - pR);
        p.bxdfPdf = vndf_cap_pdf(wo, awm, co, cm, a) * dwmdwi * pT;
        p.refr = etaRel;
        condSum = p.bxdfPdf;
        cosF = fabsf(cosI);
        for (int i = 1; i < kNL; ++i) {
            if (!wl_active(p.wl, p.flags, i)) continue;
            const float er = cosO > 0.0f ? eta.v[i] : 1.0f / eta.v[i];
            V3 wmi = nrm(wi * er + wo);
            wmi = wmi.z > 0.0f ? wmi : -wmi;
            const MCtx cmi = m_ctx(wmi);
            const float dmoi = dot(wmi, wo), dmii = dot(wmi, wi);
            const float awmi = fabsf(dmoi);
            const float secR = fresnel_real(dmoi, eta.v[i]);
            const float secT = clampf(1.0f - secR, 0.0f, 1.0f);
            const float ti = dmii + dmoi / er;
            const float deni = ti * ti;
            const float dwi = fabsf(dmii) / deni;
            const float fi = ggx_D(cmi, a) * secT * G * fabsf(dmoi * dmii / (cosO * cosI * deni)) / (er * er);
            const float pdf = vndf_cap_pdf(wo, awmi, co, cmi, a) * dwi * secT;
            if (pdf <= 0.0f || dmoi * cosO < 0.0f || dmii * cosI < 0.0f) {
                wl_terminate(p.flags, i);
            } else {
                f.v[i] = fi;
                condSum += pdf;
            }
        }
        condCount = wl_active_count(p.wl, p.flags);
        p.flags |= PF_RAY_REFRACTED;
    }
    if (p.bxdfPdf == 0.0f) { p.flags |= PF_TERMINATED; return; }
    p.T *= condCount * (1.0f / condSum) * f * cosF;
    p.rd = nrm(fr.to_world(wi));
}

// ComplexSurface callable (pathtrace.cu:548-905).
__device__ inline void complex_bxdf(const ColorTables& ct, const Luts& lut, const Light& dl, const Qrng& q,
                                    const Frame& fr, const Surface& s, PathIO& p, S4& Lnee) {
    const V3 wo = fr.to_local(-p.rd);
    const V3 li = fr.to_local(dl.dir);
    if (wo.z == 0.0f) { p.flags |= PF_TERMINATED; return; }
    const DirCtx co = dir_ctx(wo);
    float rough = s.roughness;
    bool perfect = false;
    if (rough < 0.001f) { rough = 0.001f; perfect = true; }
    const float alpha = rough * rough;
    const Rough a{alpha, alpha};
    const float eta = wo.z > 0.0f ? s.ior / 1.0f : 1.0f / s.ior;
    const float f0sr = (1.0f - eta) / (1.0f + eta);
    const float f0 = f0sr * f0sr;
    const float aco = fabsf(wo.z);
    const float zc = sqrtf(fabsf(f0sr));
    const float eCompDiel = 1.0f / lut3(lut.t + (eta > 1.0f ? kLutDielOut : kLutDielIn), aco, rough, zc);
    const float ePresDiff = 1.0f - lut3(lut.t + (eta > 1.0f ? kLutReflOut : kLutReflIn), aco, rough, zc);
    const float albCond = lut2(lut.t + kLutCond, aco, rough);
    float albSheen = 0.0f, A = 0.0f, B = 0.0f;
    if (s.sheen) {
        const float4 c = lut2x4(lut.t + kLutSheen, aco, s.sheenAlpha);
        A = c.x; B = c.y; albSheen = c.z;
        if (albSheen < 0.0001f) albSheen = 0.0f;
    }
    const float sheen = albSheen, base = 1.0f - sheen;
    const float metal = s.metalness, transl = s.transmission;

    if (!dl.occluded) {
        const V3 wi = li;
        const DirCtx ci = dir_ctx(wi);
        const float cti = wi.z, cto = wo.z;
        const float cosF = fabsf(cti);
        const bool refl = cti * cto > 0.0f;
        const float etaR = refl ? 1.0f : eta;
        V3 wm = nrm(wi * etaR + wo);
        if (wm.z < 0.0f) wm = -wm;
        const MCtx cm = m_ctx(wm);
        const float dmo = dot(wm, wo), dmi = dot(wm, wi);
        const float awm = fabsf(dot(wo, wm));
        const float Fd = schlick(f0, awm);
        const float Ds = ggx_D(cm, a);
        const float Gs = smith_G(ci, co, a);
        const float wSheen = sheen, wC = base * metal, wDS = base * (1.0f - metal) * Fd;
        const float wDT = base * (1.0f - metal) * transl * (1.0f - Fd);
        const float wDiff = base * (1.0f - metal) * (1.0f - transl) * (1.0f - Fd);
        S4 flSheen = s4(0.0f);
        float pSheen = 1.0f;
        if (sheen != 0.0f) {
            float fns = 0.0f;
            ltc_eval(wo, wi, A, B, fns, pSheen);
            flSheen = ct.rgb(s.sheenColor, p.wl) * fns;
        }
        if (refl) {
            S4 flC = s4(0.0f);
            float flD = 0.0f, pSpec = 0.0f;
            if (!perfect) {
                flC = Ds * Gs * ct.rgb(conductor_color(s.base, awm, albCond), p.wl) / fabsf(4.0f * wo.z * wi.z);
                flD = Ds * Gs * eCompDiel / fabsf(4.0f * wo.z * wi.z);
                pSpec = vndf_cap_pdf(wo, awm, co, cm, a) / (4.0f * awm);
            }
            const S4 flDiff = ePresDiff * ct.rgb(s.base, p.wl) / 3.14159265358979f;
            const float pdf = wSheen * pSheen + wC * pSpec + wDS * pSpec + wDiff * (wi.z / 3.14159265358979f);
            const float mis = power_heuristic(dl.pdf, pdf);
            Lnee += dl.spectrum * p.T * cosF * mis * (wSheen * flSheen + wC * flC + s4(wDS * flD) + wDiff * flDiff) / dl.pdf;
        } else if (transl != 0.0f && !(dmo * cto <= 0.0f || dmi * cti <= 0.0f)) {
            float pT = 0.0f;
            S4 flT = s4(0.0f);
            if (!perfect) {
                const float t = dmi + dmo / eta;
                const float den = t * t;
                pT = vndf_cap_pdf(wo, awm, co, cm, a) * (fabsf(dmi) / den);
                flT = ct.rgb(s.base, p.wl) * Ds * (1.0f - Fd) * Gs * eCompDiel * fabsf(dmo * dmi / (cto * cti * den));
            }
            const float pdf = wSheen * pSheen + wDT * pT;
            const float mis = power_heuristic(dl.pdf, pdf);
            flT = flT / (eta * eta);
            Lnee += dl.spectrum * p.T * cosF * mis * (wSheen * flSheen + wDT * flT) / dl.pdf;
        }
    }

    V3 wi;
    {
        const V3 u = q.s3(QD_BXDF0);
        const float baseCorr = 1.0f - albSheen;
        if (u.z > base) {
            float fns = 0.0f, pdfP = 0.0f;
            wi = ltc_sample(wo, A, B, V2{u.x, u.y});
            ltc_eval(wo, wi, A, B, fns, pdfP);
            const S4 f = ct.rgb(s.sheenColor, p.wl) * fns * sheen;
            p.bxdfPdf = pdfP * sheen;
            p.T *= f * fabsf(wi.z) / p.bxdfPdf;
        } else {
            const V3 wm = vndf_cap(wo, a, V2{u.x, u.y});
            const MCtx cm = m_ctx(wm);
            const float awm = fabsf(dot(wo, wm));
            const float Ds = ggx_D(cm, a);
            float Fd = schlick(f0, awm);
            if (fmaxf(0.0f, 1.0f - awm * awm) / (eta * eta) >= 1.0f) Fd = 1.0f;
            const float condP = base * metal;
            const float dSpecP = base * (1.0f - metal) * Fd;
            const float transP = base * transl * (1.0f - metal) * (1.0f - Fd);
            const float diffP = base * (1.0f - transl) * (1.0f - metal) * (1.0f - Fd);
            if (u.z < condP + dSpecP) {
                wi = nrm(reflect_about(wo, wm));
                const DirCtx ci = dir_ctx(wi);
                if (wo.z * wi.z <= 0.0f) { p.flags |= PF_TERMINATED; return; }
                const float spec = Ds * smith_G(ci, co, a) / fabsf(4.0f * wo.z * wi.z);
                S4 f;
                float pdfP;
                if (u.z < condP) {
                    f = spec * ct.rgb(conductor_color(s.base, awm, albCond), p.wl) * metal;
                    pdfP = condP;
                } else {
                    f = s4(spec * Fd * eCompDiel * (1.0f - metal));
                    pdfP = dSpecP;
                }
                p.bxdfPdf = vndf_cap_pdf(wo, awm, co, cm, a) / (4.0f * awm) * pdfP;
                p.T *= f * fabsf(wi.z) * baseCorr / p.bxdfPdf;
            } else if (u.z < condP + dSpecP + transP) {
                wi = refract_through(wo, wm, eta);
                p.refr *= eta;
                if (wo.z * wi.z >= 0.0f) { p.flags |= PF_TERMINATED; return; }
                const DirCtx ci = dir_ctx(wi);
                const float Gs = smith_G(ci, co, a);
                const float dmo = dot(wm, wo), dmi = dot(wm, wi);
                const float t = dmi + dmo / eta;
                const float den = t * t;
                float f = Ds * (1.0f - Fd) * Gs * (1.0f - metal) * transl * fabsf(dmo * dmi / (wo.z * wi.z * den));
                f /= (eta * eta);
                p.bxdfPdf = vndf_cap_pdf(wo, awm, co, cm, a) * (fabsf(dmi) / den) * transP;
                p.T *= ct.rgb(s.base, p.wl) * f * eCompDiel * fabsf(wi.z) * baseCorr / p.bxdfPdf;
                p.flags |= PF_RAY_REFRACTED;
            } else {
                wi = cosine_hemisphere(q.s2(QD_BXDF1));
                if (wo.z * wi.z <= 0.0f) { p.flags |= PF_TERMINATED; return; }
                const S4 f = ct.rgb(s.base, p.wl) / 3.14159265358979f * (1.0f - metal) * (1.0f - transl);
                p.bxdfPdf = (wi.z / 3.14159265358979f) * diffP;
                p.T *= f * ePresDiff * fabsf(wi.z) * baseCorr / p.bxdfPdf;
            }
        }
    }
    if (p.bxdfPdf <= 0.0f || isinf(p.bxdfPdf)) { p.flags |= PF_TERMINATED; return; }
    p.rd = nrm(fr.to_world(wi));
}

}  // namespace dev
}  // namespace rbh
