// Guided direction sampling for ComplexSurface vertices: one-sample MIS
// between the learned von Mises-Fisher mixture and a BSDF sampler whose
// density has a closed form. No reference counterpart; it replaces, in guided
// renders only, the lobe-selective sampling of complex_bxdf
// (src/kernels/pathtrace.cu:548-905). The BSDF value is the reference's
// ComplexSurface lobe sum for a given direction pair (the expression its NEE
// branch evaluates, pathtrace.cu:660-790).
//
// Estimator: with probability a (the network's selection output, in
// [0.1, 0.9]) draw w from the mixture, otherwise from the BSDF sampler;
// weight f(w)|cos| / (a p_guide(w) + (1 - a) p_bsdf(w)).
// Both densities are exact, so the estimator is unbiased for any network
// output. The BSDF sampler picks its lobe with probabilities that depend on
// wo only (Schlick Fresnel at wo instead of at the sampled microfacet), which
// is what makes its mixture density closed-form.
#pragma once
#include "bxdf.hpp"

namespace rbh {
namespace dev {

constexpr float kGuideMinRough = 0.05f;    // smoother vertices keep the reference sampler
constexpr int kMixFloats = 41;              // 8 lobes x (mean xyz, sharpness, weight), world space,
                                            // then the selection probability of the lobes
constexpr uint32_t QD_GUIDE_DIR = 11;       // quasi-random dimension of the mixture sample

// --------------------------------------------------------------- mixture
__device__ __forceinline__ float vmf_mix_pdf(const float* __restrict__ mix, V3 w) {
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float* q = mix + 5 * j;
        const float k = q[3];
        const float c = q[0] * w.x + q[1] * w.y + q[2] * w.z;
        s += q[4] * (k / (6.283185307179586f * -expm1f(-2.0f * k))) * expf(k * (c - 1.0f));
    }
    return s;
}

__device__ __forceinline__ V3 vmf_mix_sample(const float* __restrict__ mix, V3 u) {
    // first lobe whose cumulative weight exceeds u.x (the last one otherwise)
    float c = 0.0f;
    int j = 7;
    for (int i = 0; i < 7; ++i) {
        c += mix[5 * i + 4];
        if (u.x < c) { j = i; break; }
    }
    const float* q = mix + 5 * j;
    const float k = q[3];
    float cz = 1.0f + logf(u.y + (1.0f - u.y) * expf(-2.0f * k)) / k;
    cz = clampf(cz, -1.0f, 1.0f);
    const float sz = sqrtf(fmaxf(0.0f, 1.0f - cz * cz));
    float sp, cp;
    sincosf(6.283185307179586f * u.z, &sp, &cp);
    const Frame fr = frame_from_normal(V3{q[0], q[1], q[2]});
    return nrm(fr.t * (sz * cp) + fr.b * (sz * sp) + fr.n * cz);
}

// --------------------------------------------------------- complex surface
struct CxCtx {
    Rough a;
    DirCtx co;
    float eta, f0, eCompDiel, ePresDiff, albCond, A, B, sheen, base, metal, transl;
    float pSheen, pRefl, pTrans, pDiff;   // lobe selection of the closed-form sampler
};

__device__ __forceinline__ CxCtx cx_setup(const Luts& lut, const Surface& s, V3 wo) {
    CxCtx c;
    const float rough = s.roughness;
    const float alpha = rough * rough;
    c.a = Rough{alpha, alpha};
    c.co = dir_ctx(wo);
    c.eta = wo.z > 0.0f ? s.ior / 1.0f : 1.0f / s.ior;
    const float f0sr = (1.0f - c.eta) / (1.0f + c.eta);
    c.f0 = f0sr * f0sr;
    const float aco = fabsf(wo.z);
    const float zc = sqrtf(fabsf(f0sr));
    c.eCompDiel = 1.0f / lut3(lut.t + (c.eta > 1.0f ? kLutDielOut : kLutDielIn), aco, rough, zc);
    c.ePresDiff = 1.0f - lut3(lut.t + (c.eta > 1.0f ? kLutReflOut : kLutReflIn), aco, rough, zc);
    c.albCond = lut2(lut.t + kLutCond, aco, rough);
    float albSheen = 0.0f;
    c.A = c.B = 0.0f;
    if (s.sheen) {
        const float4 t = lut2x4(lut.t + kLutSheen, aco, s.sheenAlpha);
        c.A = t.x;
        c.B = t.y;
        albSheen = t.z;
        if (albSheen < 0.0001f) albSheen = 0.0f;
    }
    c.sheen = albSheen;
    c.base = 1.0f - albSheen;
    c.metal = s.metalness;
    c.transl = s.transmission;
    const float Fo = schlick(c.f0, aco);
    c.pSheen = c.sheen;
    c.pRefl = c.base * (c.metal + (1.0f - c.metal) * Fo);
    c.pTrans = c.base * (1.0f - c.metal) * c.transl * (1.0f - Fo);
    c.pDiff = c.base * (1.0f - c.metal) * (1.0f - c.transl) * (1.0f - Fo);
    return c;
}

// For the local direction pair (wo, wi), wo.z > 0:
//  fNee  the lobe sum the reference's NEE branch evaluates (pathtrace.cu:660-790);
//  fSamp the BSDF its sampling branch integrates (pathtrace.cu:791-893): the
//        same lobes, except that the diffuse lobe carries no (1 - F) factor,
//        the transmission lobe carries it once, and the sheen lobe only
//        reaches wo's side. Guided directions are weighted with fSamp so that
//        a guided render converges to the plain render's image.
//  pdf   the closed-form density of cx_sample.
// None of them includes the cosine.
__device__ __forceinline__ void cx_eval(const CxCtx& c, const ColorTables& ct, const Surface& s, const Wl& wl, V3 wo,
                                        V3 wi, S4& fNee, S4& fSamp, float& pdf) {
    fNee = s4(0.0f);
    fSamp = s4(0.0f);
    pdf = 0.0f;
    if (wi.z == 0.0f) return;
    const DirCtx ci = dir_ctx(wi);
    const float cti = wi.z, cto = wo.z;
    const bool refl = cti * cto > 0.0f;
    const float etaR = refl ? 1.0f : c.eta;
    V3 wm = nrm(wi * etaR + wo);
    if (wm.z < 0.0f) wm = -wm;
    const MCtx cm = m_ctx(wm);
    const float dmo = dot(wm, wo), dmi = dot(wm, wi);
    const float awm = fabsf(dmo);
    const float Fd = schlick(c.f0, awm);
    const float Ds = ggx_D(cm, c.a);
    const float Gs = smith_G(ci, c.co, c.a);
    const float wSheen = c.sheen, wC = c.base * c.metal, wDS = c.base * (1.0f - c.metal) * Fd;
    const float wT = c.base * (1.0f - c.metal) * c.transl;
    const float wD = c.base * (1.0f - c.metal) * (1.0f - c.transl);
    S4 flSheen = s4(0.0f);
    float pSheen = 0.0f;
    if (c.sheen != 0.0f) {
        float fns = 0.0f, pl = 0.0f;
        ltc_eval(wo, wi, c.A, c.B, fns, pl);
        flSheen = ct.rgb(s.sheenColor, wl) * fns;
        if (refl) pSheen = pl;   // the LTC sampler only returns directions on wo's side
    }
    const float vpdf = vndf_cap_pdf(wo, awm, c.co, cm, c.a);
    if (refl) {
        const S4 flC = Ds * Gs * ct.rgb(conductor_color(s.base, awm, c.albCond), wl) / fabsf(4.0f * wo.z * wi.z);
        const float flD = Ds * Gs * c.eCompDiel / fabsf(4.0f * wo.z * wi.z);
        const S4 flDiff = c.ePresDiff * ct.rgb(s.base, wl) / 3.14159265358979f;
        const S4 common = wSheen * flSheen + wC * flC + s4(wDS * flD);
        fNee = common + (wD * (1.0f - Fd)) * flDiff;
        fSamp = common + wD * flDiff;
        pdf = c.pSheen * pSheen + c.pRefl * (dmo > 0.0f ? vpdf / (4.0f * awm) : 0.0f) +
              c.pDiff * (wi.z / 3.14159265358979f);
    } else {
        fNee = wSheen * flSheen;
        if (c.transl != 0.0f && !(dmo * cto <= 0.0f || dmi * cti <= 0.0f)) {
            const float t = dmi + dmo / c.eta;
            const float den = t * t;
            const S4 flT = ct.rgb(s.base, wl) * Ds * (1.0f - Fd) * Gs * c.eCompDiel *
                           fabsf(dmo * dmi / (cto * cti * den)) / (c.eta * c.eta);
            fNee += (wT * (1.0f - Fd)) * flT;
            fSamp = wT * flT;
            pdf = c.pTrans * vpdf * (fabsf(dmi) / den);
        }
    }
}

// Draws a direction from the closed-form BSDF mixture; false when the drawn
// microfacet yields no direction (below the surface or total internal
// reflection), which the density of cx_eval accounts for.
__device__ __forceinline__ bool cx_sample(const CxCtx& c, V3 wo, V3 u, V2 u2, V3& wi) {
    if (u.z < c.pSheen) {
        wi = ltc_sample(wo, c.A, c.B, V2{u.x, u.y});
        return wi.z > 0.0f;
    }
    const float r = u.z - c.pSheen;
    if (r < c.pRefl + c.pTrans) {
        const V3 wm = vndf_cap(wo, c.a, V2{u.x, u.y});
        if (r < c.pRefl) {
            wi = nrm(reflect_about(wo, wm));
            return wi.z > 0.0f;
        }
        wi = refract_through(wo, wm, c.eta);
        return wi.z < 0.0f;
    }
    wi = cosine_hemisphere(u2);
    return wi.z > 0.0f;
}

}  // namespace dev
}  // namespace rbh

namespace rbh {
namespace dev {

// Guided ComplexSurface vertex (wo on the front side of the shading frame):
// NEE with the MIS weight of the combined density, then the one-sample MIS
// direction. Returns the combined density of the chosen direction (0: the
// path ends) and, in pbOut, the BSDF sampler's density of it. `mix` is the
// vertex's mixture in world space followed by its selection probability.
__device__ __forceinline__ float complex_guided(const ColorTables& ct, const Luts& lut, const Light& dl, const Qrng& q,
                                                const Frame& fr, const Surface& s, PathIO& p, S4& Lnee,
                                                const float* __restrict__ mix, V3& wiWorld, float& pbOut) {
    const V3 wo = fr.to_local(-p.rd);
    const CxCtx c = cx_setup(lut, s, wo);
    const float alpha = mix[kMixFloats - 1];
    if (!dl.occluded) {
        const V3 li = fr.to_local(dl.dir);
        S4 fNee, fSamp;
        float pb;
        cx_eval(c, ct, s, p.wl, wo, li, fNee, fSamp, pb);
        const float pmix = alpha * vmf_mix_pdf(mix, dl.dir) + (1.0f - alpha) * pb;
        const float mis = power_heuristic(dl.pdf, pmix);
        Lnee += dl.spectrum * p.T * fabsf(li.z) * mis * fNee / dl.pdf;
    }
    V3 wi;
    bool ok;
    if (q.s1(QD_GUIDE) < alpha) {
        wiWorld = vmf_mix_sample(mix, q.s3(QD_GUIDE_DIR));
        wi = fr.to_local(wiWorld);
        ok = true;
    } else {
        ok = cx_sample(c, wo, q.s3(QD_BXDF0), q.s2(QD_BXDF1), wi);
        wiWorld = fr.to_world(wi);
    }
    if (!ok) return 0.0f;
    S4 fNee, fSamp;
    float pb;
    cx_eval(c, ct, s, p.wl, wo, wi, fNee, fSamp, pb);
    const float pmix = alpha * vmf_mix_pdf(mix, wiWorld) + (1.0f - alpha) * pb;
    pbOut = pb;
    if (!(pmix > 0.0f) || isinf(pmix)) return 0.0f;
    p.T *= fSamp * (fabsf(wi.z) / pmix);
    p.bxdfPdf = pmix;
    if (wi.z < 0.0f) {
        p.refr *= c.eta;
        p.flags |= PF_RAY_REFRACTED;
    }
    p.rd = nrm(wiWorld);
    return pmix;
}

}  // namespace dev
}  // namespace rbh
