// TEST INFRASTRUCTURE — C entry points of the CPU oracle (rb_oracle.h).
#include "render.hpp"
#include "rb_bvh.hpp"
#include <vector>

namespace orc {
int cdf_build_f32(const float*, int, int, int, float*, float*, float*);
int cdf_build_f64(const float*, int, int, int, double*, double*, double*);
void cdf_invert(const float*, const float*, int, int, uint16_t*, uint16_t*);
int cdf_block_size(int);
int light_tree_build(const float*, const float*, uint32_t, std::vector<PackedNode>&, std::vector<uint32_t>&, std::vector<uint64_t>&);
int light_tree_build(const float*, const float*, uint32_t, const orc_shape_light*, const float*, uint32_t,
                     std::vector<PackedNode>&, std::vector<uint32_t>&, std::vector<uint64_t>&);
int render(const orc_scene*, const orc_params*, double*, int, orc_render_stats*);
}  // namespace orc

using namespace orc;

extern "C" {

// in: n x {sample, hash, offset, dim, kind(1|2|3)}; out: n x 3
int orc_sobol(const uint32_t* in, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = in + 5 * (size_t)i;
        QState q{r[0], r[2], r[1]};
        float* o = out + 3 * (size_t)i;
        o[0] = o[1] = o[2] = 0.0f;
        if (r[4] == 1) o[0] = sobol1(q, r[3]);
        else if (r[4] == 2) { V2 v = sobol2(q, r[3]); o[0] = v.x; o[1] = v.y; }
        else { V3 v = sobol3(q, r[3]); o[0] = v.x; o[1] = v.y; o[2] = v.z; }
    }
    return 0;
}

int orc_sg_erf(const float* x, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) out[i] = sg_erf(x[i]);
    return 0;
}

int orc_sobol_table(uint32_t* out) {
    const SobolTable& T = sobol_table();
    std::memcpy(out, T.t, sizeof(T.t));
    return 0;
}

int orc_pixel_hash(const uint32_t* xy, uint32_t* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) out[i] = pixel_hash(xy[2 * i], xy[2 * i + 1]);
    return 0;
}

int orc_wavelengths(const float* u, float* lambda, float* pdf, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        Wavelengths w = sample_visible(u[i]);
        for (int k = 0; k < NL; ++k) { lambda[4 * i + k] = w.lambda.v[k]; pdf[4 * i + k] = w.pdf.v[k]; }
    }
    return 0;
}

int orc_cdf_build(const float* t, int w, int h, int ch, float* cx, float* cy, float* integral) {
    return cdf_build_f32(t, w, h, ch, cx, cy, integral);
}
int orc_cdf_build_f64(const float* t, int w, int h, int ch, double* cx, double* cy, double* integral) {
    return cdf_build_f64(t, w, h, ch, cx, cy, integral);
}
void orc_cdf_invert(const float* cx, const float* cy, int w, int h, uint16_t* cond, uint16_t* marg) {
    cdf_invert(cx, cy, w, h, cond, marg);
}
int orc_cdf_block_size(int w) { return cdf_block_size(w); }

int orc_light_tree_build(const float* tv, const float* flux, uint32_t n, void* nodes, uint32_t* lp, uint64_t* bm) {
    std::vector<PackedNode> N;
    std::vector<uint32_t> L;
    std::vector<uint64_t> B;
    int r = light_tree_build(tv, flux, n, N, L, B);
    if (r < 0) return r;
    std::memcpy(nodes, N.data(), N.size() * sizeof(PackedNode));
    std::memcpy(lp, L.data(), L.size() * 4);
    std::memcpy(bm, B.data(), B.size() * 8);
    return r;
}

int orc_light_tree_build_shapes(const float* tv, const float* flux, uint32_t n, const orc_shape_light* shapes,
                                const float* shape_flux, uint32_t m, void* nodes, uint32_t* lp, uint64_t* bm) {
    std::vector<PackedNode> N;
    std::vector<uint32_t> L;
    std::vector<uint64_t> B;
    int r = light_tree_build(tv, flux, n, shapes, shape_flux, m, N, L, B);
    if (r < 0) return r;
    std::memcpy(nodes, N.data(), N.size() * sizeof(PackedNode));
    std::memcpy(lp, L.data(), L.size() * 4);
    std::memcpy(bm, B.data(), B.size() * 8);
    return r;
}

int orc_light_tree_pdf(const void* nodes, const uint32_t* lptrs, const uint64_t* masks, int env, const uint32_t* lights,
                       const float* pn, float* pdf, uint32_t n) {
    LightTreeView t{(const PackedNode*)nodes, lptrs, masks, env != 0};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = pn + 6 * (size_t)i;
        pdf[i] = light_tree_pdf(t, LT_TRIANGLE, lights[i], {r[0], r[1], r[2]}, {r[3], r[4], r[5]});
    }
    return 0;
}

int orc_light_tree_sample(const void* nodes, const uint32_t* lptrs, const uint64_t* masks, int env, const float* rpn,
                          uint32_t* ti, float* pdf, uint32_t n) {
    LightTreeView t{(const PackedNode*)nodes, lptrs, masks, env != 0};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rpn + 7 * (size_t)i;
        uint32_t type = 0, idx = 0;
        float p = 0.0f;
        light_tree_sample(t, r[0], {r[1], r[2], r[3]}, {r[4], r[5], r[6]}, type, idx, p);
        ti[2 * i] = type; ti[2 * i + 1] = idx; pdf[i] = p;
    }
    return 0;
}

int orc_oct_encode_u32(const float* v, uint32_t* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) out[i] = oct_encode_u32({v[3 * i], v[3 * i + 1], v[3 * i + 2]});
    return 0;
}
int orc_oct_decode_u32(const uint32_t* e, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) { V3 d = oct_decode_u32(e[i]); out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z; }
    return 0;
}

// Brute-force closest / any hit over all triangles (checker for the LBVH).
int orc_trace_brute(const float* verts, const uint32_t* idx, uint32_t nt, const float* rays, float* hits, uint32_t n, int anyhit) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        V3 o{r[0], r[1], r[2]}, d{r[4], r[5], r[6]};
        float tmin = r[3], best = r[7];
        uint32_t prim = 0xFFFFFFFFu;
        float bb1 = 0, bb2 = 0;
        for (uint32_t t = 0; t < nt; ++t) {
            const float* a = verts + 3 * (size_t)idx[3 * (size_t)t];
            const float* b = verts + 3 * (size_t)idx[3 * (size_t)t + 1];
            const float* c = verts + 3 * (size_t)idx[3 * (size_t)t + 2];
            float tt, b1, b2;
            if (ray_triangle(o, d, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, tmin, best, tt, b1, b2)) {
                best = tt; prim = t; bb1 = b1; bb2 = b2;
                if (anyhit) break;
            }
        }
        float* h = hits + 4 * (size_t)i;
        h[0] = prim == 0xFFFFFFFFu ? -1.0f : best;
        h[1] = u2f(prim);
        h[2] = bb1; h[3] = bb2;
    }
    return 0;
}

// Per-sample BxDF evaluation. Input record (64 x 32-bit words, floats unless
// noted): [0] material index (u32) [1] path flags (u32) [2..4] ray direction
// [5..13] shading frame columns T, B, N [14..17] light spectrum [18..20]
// light direction [21] light pdf [22] occluded (u32) [23] sample (u32)
// [24] dimension offset (u32) [25] pixel hash (u32) [26..29] wavelengths
// [30..33] wavelength pdfs [34] surface layers (u32) [35..37] base colour
// [38] metalness [39] ior [40] roughness [41] transmission [42..44] sheen
// colour [45] sheen alpha [46..49] throughput [50] refraction scale
// [51..54] L in.
// Output record (16 words): [0..3] L [4..7] throughput [8] bxdf pdf
// [9..11] new direction [12] flags (u32) [13] refraction scale.
int orc_bxdf_unit(const orc_scene* sc, const float* in, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = in + 64 * (size_t)i;
        auto U = [&](int k) { return f2u(r[k]); };
        const orc_material& m = sc->materials[U(0)];
        PathIO p;
        p.flags = U(1);
        p.rD = {r[2], r[3], r[4]};
        Local local;
        local.f = M3{{{r[5], r[6], r[7]}, {r[8], r[9], r[10]}, {r[11], r[12], r[13]}}};
        DirectLight dl;
        for (int k = 0; k < NL; ++k) dl.spectrum.v[k] = r[14 + k];
        dl.dir = {r[18], r[19], r[20]};
        dl.pdf = r[21];
        dl.occluded = U(22) != 0;
        QState q{U(23), U(24), U(25)};
        for (int k = 0; k < NL; ++k) { p.wl.lambda.v[k] = r[26 + k]; p.wl.pdf.v[k] = r[30 + k]; }
        p.wl.active = 4.0f;
        Surface s;
        s.layers = U(34);
        s.baseColor = {r[35], r[36], r[37]};
        s.metalness = r[38]; s.ior = r[39]; s.roughness = r[40]; s.transmission = r[41];
        s.sheenColor = {r[42], r[43], r[44]};
        s.sheenAlpha = r[45];
        for (int k = 0; k < NL; ++k) { p.T.v[k] = r[46 + k]; p.L.v[k] = r[51 + k]; }
        p.refrScale = r[50];
        p.bxdfPDF = 1.0f;
        if (m.bxdf == 0) bxdf_conductor(sc, m, dl, q, local, p);
        else if (m.bxdf == 2) bxdf_complex(sc, dl, q, local, s, p);
        float* o = out + 16 * (size_t)i;
        for (int k = 0; k < NL; ++k) { o[k] = p.L.v[k]; o[4 + k] = p.T.v[k]; }
        o[8] = p.bxdfPDF;
        o[9] = p.rD.x; o[10] = p.rD.y; o[11] = p.rD.z;
        o[12] = u2f(p.flags);
        o[13] = p.refrScale;
        o[14] = 0.0f; o[15] = 0.0f;
    }
    return 0;
}

// Single-function evaluations of the shading and light-tree math for the
// oracle's own pins (quadrature, closed forms, identities). Records are
// 8 floats in, 4 floats out; `fn` selects the function:
//  0 ggx_D(m[0..2], alpha x/y [3,4])                       -> D
//  1 ggx_G1(w[0..2], alpha [3,4])                          -> G1
//  2 vndf_sample_cap(wo[0..2], alpha [3,4], u [5,6])       -> m
//  3 vndf_pdf_cap(wo[0..2], alpha [3,4], m from in[5..7])  -> pdf of m
//  4 vndf_sample_bounded (as 2)                            -> m
//  5 vndf_pdf_bounded (as 3)                               -> pdf of m
//  6 fresnel_complex1(cos [0], eta [1], k [2])             -> R
//  7 f82(f0 [0..2], mu [3]) and schlick3                   -> f82 rgb, max |f82 - schlick|
//  8 ltc_sample(wo[0..2], A [3], B [4], u [5,6])           -> wi
//  9 ltc_eval(wo[0..2], A [3], B [4], wi from in[5..7])    -> f, pdf
// 10 sg_integral(s [0]), sg_expm1(x [1]), sg_expm1_over_x(x [1])
// 11 clamped cosine product integral over pi (cos [0], s [1]), upper [2], lower [3]
// 12 sg_product(axis1 [0..2] from oct, s1 [3], axis2 from in[4..5] oct, s2 [6]) -> axis, sharp ... (see below)
// 13 vmf_axis_length_to_sharpness(l [0]), vmf_sharpness_to_axis_length(s [1])
// 14 gaussian_filter_sample(u [0,1])                      -> offset
// 15 refract(v [0..2], n = +z, eta [3])                   -> t
// 16 sample_visible(u [0]) hero lambda/pdf                 -> lambda0, pdf0
int orc_math_unit(int fn, const float* in, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* a = in + 8 * (size_t)i;
        float* o = out + 4 * (size_t)i;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        const V3 w{a[0], a[1], a[2]};
        const Alpha al{a[3], a[4]};
        switch (fn) {
            case 0: o[0] = ggx_D(ctx_m(w), al); break;
            case 1: o[0] = ggx_G1(ctx_w(w), al); break;
            case 2: case 4: {
                const V3 m = fn == 2 ? vndf_sample_cap(w, al, {a[5], a[6]}) : vndf_sample_bounded(w, al, {a[5], a[6]});
                o[0] = m.x; o[1] = m.y; o[2] = m.z;
                break;
            }
            case 3: case 5: {
                // m is passed as (theta, phi) in a[5], a[6]
                const float st = std::sin(a[5]), ct = std::cos(a[5]);
                const V3 m{st * std::cos(a[6]), st * std::sin(a[6]), ct};
                const float aw = std::fabs(dot(w, m));
                o[0] = fn == 3 ? vndf_pdf_cap(w, aw, ctx_w(w), ctx_m(m), al) : vndf_pdf_bounded(w, aw, ctx_w(w), ctx_m(m), al);
                if (dot(w, m) <= 0.0f) o[0] = 0.0f;   // the samplers never return back-facing normals
                break;
            }
            case 6: o[0] = fresnel_complex1(a[0], a[1], a[2]); break;
            case 7: {
                const V3 f0{a[0], a[1], a[2]};
                const V3 f = f82(f0, a[3]), s = schlick3(f0, a[3]);
                o[0] = f.x; o[1] = f.y; o[2] = f.z;
                o[3] = std::max(std::fabs(f.x - s.x), std::max(std::fabs(f.y - s.y), std::fabs(f.z - s.z)));
                break;
            }
            case 8: {
                const V3 r = ltc_sample(w, a[3], a[4], {a[5], a[6]});
                o[0] = r.x; o[1] = r.y; o[2] = r.z;
                break;
            }
            case 9: {
                const float st = std::sin(a[5]), ct = std::cos(a[5]);
                const V3 wi{st * std::cos(a[6]), st * std::sin(a[6]), ct};
                ltc_eval(w, wi, a[3], a[4], o[0], o[1]);
                break;
            }
            case 10: o[0] = sg_integral(a[0]); o[1] = sg_expm1(a[1]); o[2] = sg_expm1_over_x(a[1]); break;
            case 11:
                o[0] = sg_clamped_cosine_product_integral_over_pi_2024(a[0], a[1]);
                o[1] = sg_upper_cc_over_2pi(a[1]);
                o[2] = sg_lower_cc_over_2pi(a[1]);
                break;
            case 12: {
                const V3 a1 = normalize(V3{a[0], a[1], a[2]});
                const float th = a[4], ph = a[5];
                const V3 a2{std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
                const SGLobe l = sg_product(a1, a[3], a2, a[6]);
                // out: sharpness, logAmplitude, axis . a1, axis . a2
                o[0] = l.sharpness; o[1] = l.logAmplitude; o[2] = dot(l.axis, a1); o[3] = dot(l.axis, a2);
                break;
            }
            case 13: o[0] = vmf_axis_length_to_sharpness(a[0]); o[1] = vmf_sharpness_to_axis_length(a[1]); break;
            case 14: {
                const V2 g = gaussian_filter_sample({a[0], a[1]});
                o[0] = g.x; o[1] = g.y;
                break;
            }
            case 15: {
                const V3 t = refract(w, V3{0.0f, 0.0f, 1.0f}, a[3]);
                o[0] = t.x; o[1] = t.y; o[2] = t.z;
                break;
            }
            case 17: o[0] = power_heuristic(a[0], a[1]); break;
            case 18: o[0] = pow5(a[0]); o[1] = schlick(a[1], a[0]); break;
            case 16: {
                const Wavelengths wl = sample_visible(a[0]);
                o[0] = wl.lambda.v[0]; o[1] = wl.pdf.v[0]; o[2] = wl.lambda.v[1]; o[3] = wl.pdf.v[1];
                break;
            }
            default: return -1;
        }
    }
    return 0;
}

int orc_render(const orc_scene* sc, const orc_params* p, double* accum, int threads, orc_render_stats* stats) {
    return render(sc, p, accum, threads, stats);
}

// src/kernels/image.cu:16-121 (AgX branch). The AgX matrices and contrast
// polynomial are the published AgX constants used by that file.
int orc_resolve(const double* accum, uint32_t w, uint32_t h, const float* m3, float expo, uint8_t* rgba) {
    const M3 agx{{{0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f},
                  {0.0784335999999992f, 0.878468636469772f, 0.0784336f},
                  {0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f}}};
    const M3 iagx{{{1.1968790051201738155f, -0.052896851757456180321f, -0.052971635514443794537f},
                   {-0.098020881140136776078f, 1.1519031299041727435f, -0.098043450117124120312f},
                   {-0.099029744079720471434f, -0.098961176844843346553f, 1.1510736726411610622f}}};
    const float minEv = -12.47393f, maxEv = 4.026069f;
    const M3 cs{{{m3[0], m3[1], m3[2]}, {m3[3], m3[4], m3[5]}, {m3[6], m3[7], m3[8]}}};
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const double* d = accum + 4 * i;
        double nv = 1.0 / d[3];
        V3 c{(float)(d[0] * nv), (float)(d[1] * nv), (float)(d[2] * nv)};
        c = V3{std::max(c.x, 0.0f), std::max(c.y, 0.0f), std::max(c.z, 0.0f)} * expo;
        c = mul(cs, c);
        c = mul(agx, c);
        // a channel that the matrices leave negative has no logarithm: the
        // clamp maps its NaN to minEv (fmax / fmin semantics), as the product does
        c = {std::fmin(std::fmax(std::log2(c.x), minEv), maxEv), std::fmin(std::fmax(std::log2(c.y), minEv), maxEv),
             std::fmin(std::fmax(std::log2(c.z), minEv), maxEv)};
        c = (c - v3s(minEv)) / (maxEv - minEv);
        auto poly = [](float x) {
            float x2 = x * x, x4 = x2 * x2;
            return -17.86f * x4 * x2 * x + 78.01f * x4 * x2 - 126.7f * x4 * x + 92.06f * x4 - 28.72f * x2 * x + 4.361f * x2 - 0.1718f * x + 0.002857f;
        };
        c = {poly(c.x), poly(c.y), poly(c.z)};
        c = mul(iagx, c);
        c = {clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f)};
        uint8_t* o = rgba + 4 * i;
        o[0] = (uint8_t)(c.x * 255.99999f); o[1] = (uint8_t)(c.y * 255.99999f); o[2] = (uint8_t)(c.z * 255.99999f); o[3] = 255;
    }
    return 0;
}

}  // extern "C"
