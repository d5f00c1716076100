/* A C caller of include/ratbite_hip.h that sizes every buffer exactly as the
 * header documents it, puts a guard band after each one and checks the band
 * after every call that writes into caller memory. It calls every entry
 * point of the header on a tiny scene (a floor quad under an emissive quad,
 * 16-bit index buffer, 8x4 env map, 8x8 film).
 *
 *   abi_check host   entry points that need no device (sizes and light tree)
 *   abi_check gpu    everything, on device 0
 * Prints "abi_check: ok" and exits 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ratbite_hip.h"

#define GUARD 256
static int failures = 0;

/* n bytes for the caller plus a guard band of GUARD bytes of 0xA5 */
static void* guarded(size_t n) {
    unsigned char* p = (unsigned char*)malloc(n + GUARD);
    memset(p, 0, n);
    memset(p + n, 0xA5, GUARD);
    return p;
}
static void check_guard(const void* p, size_t n, const char* what) {
    const unsigned char* b = (const unsigned char*)p + n;
    for (size_t i = 0; i < GUARD; ++i)
        if (b[i] != 0xA5) {
            fprintf(stderr, "abi_check: %s wrote past its documented %zu bytes\n", what, n);
            ++failures;
            return;
        }
}
#define EXPECT(cond, ...)                        \
    do {                                         \
        if (!(cond)) {                           \
            fprintf(stderr, "abi_check: ");      \
            fprintf(stderr, __VA_ARGS__);        \
            fprintf(stderr, "\n");               \
            ++failures;                          \
        }                                        \
    } while (0)
#define OK(ctx, call)                                                                               \
    do {                                                                                            \
        int rc_ = (call);                                                                           \
        EXPECT(rc_ == RBH_OK, "%s -> %d (%s)", #call, rc_, (ctx) ? rbh_last_error(ctx) : "no context"); \
    } while (0)

enum { NL = 2 };   /* emissive triangles */

static void light_records(rbh_emissive_tri* t) {
    /* the emissive quad at y = 1.9 facing down, camera-relative */
    const float v[4][3] = {{-0.3f, 1.9f, -3.3f}, {0.3f, 1.9f, -3.3f}, {0.3f, 1.9f, -2.7f}, {-0.3f, 1.9f, -2.7f}};
    const int tri[NL][3] = {{0, 2, 1}, {0, 3, 2}};
    for (int i = 0; i < NL; ++i) {
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) t[i].vertices[3 * k + a] = v[tri[i][k]][a];
        for (int k = 0; k < 6; ++k) t[i].uvs[k] = 0.0f;
        t[i].materialIndex = 1;
        t[i].primitiveDataIndex = (uint32_t)(2 + i);
    }
}

enum { NS = 2 };   /* shape lights */

static void shape_records(rbh_shape_light* s) {
    memset(s, 0, NS * sizeof(rbh_shape_light));
    /* a disk facing -z (quaternion: 180 degrees about x) and a sphere */
    s[0].position[0] = 0.6f; s[0].position[1] = 1.8f; s[0].position[2] = -3.0f;
    s[0].powerScale = 2.0f;
    s[0].frame[0] = 1.0f;
    s[0].radius = 0.2f;
    s[0].materialIndex = 1;
    s[0].type = RBH_SHAPE_DISK;
    s[1].position[0] = -0.6f; s[1].position[1] = 1.0f; s[1].position[2] = -3.0f;
    s[1].powerScale = 1.0f;
    s[1].frame[3] = 1.0f;
    s[1].radius = 0.15f;
    s[1].materialIndex = 1;
    s[1].type = RBH_SHAPE_SPHERE;
}

static void host_part(void) {
    EXPECT(sizeof(rbh_material) == 152 && sizeof(rbh_texture_desc) == 48 && sizeof(rbh_emissive_tri) == 68 &&
               sizeof(rbh_packed_node) == 32 && sizeof(rbh_shape_light) == 48,
           "record sizes");
    EXPECT(RBH_GUIDE_RECORD == 16 && RBH_GUIDE_OUTPUTS == 41 && RBH_GUIDE_PARAMS == 20329, "guide sizes");
    rbh_emissive_tri t[NL];
    light_records(t);
    const float flux[NL] = {1.0f, 1.0f};
    for (uint32_t flags = 0; flags < 2; ++flags) {
        rbh_packed_node* nodes = (rbh_packed_node*)guarded(2 * NL * sizeof(rbh_packed_node));
        uint32_t* ptrs = (uint32_t*)guarded(NL * 4);
        uint64_t* masks = (uint64_t*)guarded(NL * 8);
        const int nn = flags ? rbh_light_tree_build_host_ex(t, flux, NL, nodes, ptrs, masks, RBH_LIGHT_TREE_FP64)
                             : rbh_light_tree_build_host(t, flux, NL, nodes, ptrs, masks);
        EXPECT(nn >= 1 && nn <= 2 * NL, "light_tree_build_host returned %d", nn);
        check_guard(nodes, 2 * NL * sizeof(rbh_packed_node), "rbh_light_tree_build_host nodes");
        check_guard(ptrs, NL * 4, "rbh_light_tree_build_host lightPointers");
        check_guard(masks, NL * 8, "rbh_light_tree_build_host bitmasks");
        free(nodes); free(ptrs); free(masks);
    }
    {
        rbh_shape_light s[NS];
        shape_records(s);
        const float sflux[NS] = {2.0f, 3.0f};
        rbh_packed_node* nodes = (rbh_packed_node*)guarded(2 * (NL + NS) * sizeof(rbh_packed_node));
        uint32_t* ptrs = (uint32_t*)guarded((NL + NS) * 4);
        uint64_t* masks = (uint64_t*)guarded((NL + NS) * 8);
        const int nn = rbh_light_tree_build_host_shapes(t, flux, NL, s, sflux, NS, nodes, ptrs, masks, 0);
        EXPECT(nn >= 1 && nn <= 2 * (NL + NS), "light_tree_build_host_shapes returned %d", nn);
        check_guard(nodes, 2 * (NL + NS) * sizeof(rbh_packed_node), "rbh_light_tree_build_host_shapes nodes");
        check_guard(ptrs, (NL + NS) * 4, "rbh_light_tree_build_host_shapes lightPointers");
        check_guard(masks, (NL + NS) * 8, "rbh_light_tree_build_host_shapes bitmasks");
        int shapes = 0;
        for (int i = 0; i < NL + NS; ++i) shapes += (ptrs[i] >> 30) == 3u;
        EXPECT(shapes == NS, "%d shape-light pointers", shapes);
        free(nodes); free(ptrs); free(masks);
    }
}

static void gpu_part(void) {
    rbh_ctx c = NULL;
    OK(NULL, rbh_create(0, &c));
    if (!c) return;
    /* tables */
    float* dense = (float*)calloc(471, 4);
    for (int i = 0; i < 471; ++i) dense[i] = 1.0f;
    OK(c, rbh_upload_spectra(c, dense, 1));
    float* xyz = (float*)calloc(3 * 471, 4);
    float* basis = (float*)calloc(3 * 471, 4);
    for (int i = 0; i < 3 * 471; ++i) { xyz[i] = 0.3f; basis[i] = 0.5f; }
    OK(c, rbh_upload_color_tables(c, xyz, basis));
    float* lut2 = (float*)calloc(32 * 32, 4);
    float* lut3 = (float*)calloc(32 * 32 * 32, 4);
    float* sheen = (float*)calloc(32 * 32 * 4, 4);
    for (int i = 0; i < 32 * 32; ++i) lut2[i] = 0.9f;
    for (int i = 0; i < 32 * 32 * 32; ++i) lut3[i] = 0.9f;
    OK(c, rbh_upload_luts(c, lut2, lut3, lut3, lut3, lut3, sheen));
    OK(c, rbh_upload_textures(c, NULL, 0, NULL, 0));
    rbh_material m[2];
    memset(m, 0, sizeof(m));
    for (int i = 0; i < 2; ++i) {
        m[i].bxdf = 2;
        m[i].factors = 1 | 2 | 4;   /* base colour, metalness, roughness */
        m[i].baseColor[0] = m[i].baseColor[1] = m[i].baseColor[2] = 0.7f;
        m[i].baseColor[3] = 1.0f;
        m[i].roughness = 0.6f;
        m[i].ior = 1.5f;
        for (int k = 0; k < 7; ++k) m[i].tex[k] = -1;
    }
    m[1].emissive[0] = m[1].emissive[1] = m[1].emissive[2] = 5.0f;
    OK(c, rbh_upload_materials(c, m, 2));
    /* floor quad (triangles 0, 1) and the emissive quad (2, 3) */
    const float verts[8][3] = {{-2, 0, -5}, {2, 0, -5}, {2, 0, -1}, {-2, 0, -1},
                               {-0.3f, 1.9f, -3.3f}, {0.3f, 1.9f, -3.3f}, {0.3f, 1.9f, -2.7f}, {-0.3f, 1.9f, -2.7f}};
    const uint16_t idx16[12] = {0, 2, 1, 0, 3, 2, 4, 6, 5, 4, 7, 6};
    const uint32_t idx32[12] = {0, 2, 1, 0, 3, 2, 4, 6, 5, 4, 7, 6};
    const uint32_t triMat[4] = {0, 0, 1, 1};
    const uint32_t triLight[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0, 1};
    OK(c, rbh_upload_geometry(c, &verts[0][0], 8, idx32, 4, NULL, 0, triMat, triLight));
    OK(c, rbh_upload_geometry_ex(c, &verts[0][0], 8, idx16, RBH_INDEX_U16, 4, NULL, 0, triMat, triLight));
    OK(c, rbh_build_accel(c));
    rbh_emissive_tri lt[NL];
    light_records(lt);
    const float flux[NL] = {1.0f, 1.0f};
    OK(c, rbh_build_light_tree(c, lt, flux, NL));
    OK(c, rbh_build_light_tree_ex(c, lt, flux, NL, RBH_LIGHT_TREE_FP64));
    {
        rbh_packed_node* nodes = (rbh_packed_node*)guarded(2 * NL * sizeof(rbh_packed_node));
        uint32_t* ptrs = (uint32_t*)guarded(NL * 4);
        uint64_t* masks = (uint64_t*)guarded(NL * 8);
        uint32_t nn = 0;
        OK(c, rbh_get_light_tree(c, nodes, &nn, ptrs, masks));
        check_guard(nodes, 2 * NL * sizeof(rbh_packed_node), "rbh_get_light_tree nodes");
        check_guard(ptrs, NL * 4, "rbh_get_light_tree lightPointers");
        check_guard(masks, NL * 8, "rbh_get_light_tree bitmasks");
        OK(c, rbh_set_light_tree(c, nodes, nn, ptrs, masks, lt, NL));
        /* a tree whose inner node points back at itself is refused */
        if (nn > 1) {
            nodes[0].coreData = 0u;
            EXPECT(rbh_set_light_tree(c, nodes, nn, ptrs, masks, lt, NL) == RBH_ERR_INVALID, "cyclic tree accepted");
        }
        OK(c, rbh_build_light_tree(c, lt, flux, NL));
        free(nodes); free(ptrs); free(masks);
    }
    /* env map */
    const int ew = 8, eh = 4;
    float* env = (float*)calloc((size_t)ew * eh * 3, 4);
    for (int i = 0; i < ew * eh * 3; ++i) env[i] = 0.2f + 0.01f * (float)i;
    OK(c, rbh_envmap_build(c, env, ew, eh, 3, 0));
    OK(c, rbh_envmap_rebuild(c, RBH_CDF_FP64, 2));
    {
        uint16_t* cond = (uint16_t*)guarded((size_t)ew * eh * 2);
        uint16_t* marg = (uint16_t*)guarded((size_t)eh * 2);
        float* cdfY = (float*)guarded((size_t)eh * 4);
        double integral = 0.0;
        OK(c, rbh_envmap_read(c, cond, marg, cdfY, &integral));
        check_guard(cond, (size_t)ew * eh * 2, "rbh_envmap_read cond");
        check_guard(marg, (size_t)eh * 2, "rbh_envmap_read marg");
        check_guard(cdfY, (size_t)eh * 4, "rbh_envmap_read cdfY");
        EXPECT(integral > 0.0, "env integral %g", integral);
        free(cond); free(marg); free(cdfY);
    }
    /* render */
    const uint32_t W = 8, H = 8;
    rbh_launch_params lp;
    memset(&lp, 0, sizeof(lp));
    lp.width = W; lp.height = H;
    lp.maxPathDepth = lp.maxReflectedPathDepth = lp.maxTransmittedPathDepth = 4;
    lp.sampleOffset = 0; lp.sampleCount = 2;
    lp.camU[0] = 1.0f; lp.camV[2] = -1.0f; lp.camW[1] = 1.0f;
    lp.perspectiveScaleW = lp.perspectiveScaleH = 0.6f;
    OK(c, rbh_set_path_slots(c, 0));
    OK(c, rbh_render(c, &lp));
    EXPECT(rbh_query(c) == 1, "rbh_query after a finished render");
    OK(c, rbh_synchronize(c));
    rbh_stats st;
    OK(c, rbh_stats_get(c, &st));
    EXPECT(st.samples == (uint64_t)W * H * 2, "stats samples %llu", (unsigned long long)st.samples);
    double* acc = (double*)guarded((size_t)W * H * 32);
    OK(c, rbh_read_accum(c, acc));
    check_guard(acc, (size_t)W * H * 32, "rbh_read_accum");
    double wsum = 0.0;
    for (uint32_t i = 0; i < W * H; ++i) wsum += acc[4 * i + 3];
    EXPECT(wsum == 2.0 * W * H, "sample weights %g", wsum);
    OK(c, rbh_write_accum(c, acc, W, H));
    uint8_t* rgba = (uint8_t*)guarded((size_t)W * H * 4);
    const float cs[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    OK(c, rbh_resolve(c, cs, 1.0f, rgba));
    check_guard(rgba, (size_t)W * H * 4, "rbh_resolve");
    /* shape lights: a render is refused until the light tree is rebuilt */
    {
        rbh_shape_light s[NS];
        shape_records(s);
        const float sflux[NS] = {2.0f, 3.0f};
        OK(c, rbh_upload_shape_lights(c, s, sflux, NS));
        EXPECT(rbh_render(c, &lp) == RBH_ERR_STATE, "render with a stale light tree accepted");
        OK(c, rbh_build_light_tree(c, lt, flux, NL));
        rbh_packed_node* nodes = (rbh_packed_node*)guarded(2 * (NL + NS) * sizeof(rbh_packed_node));
        uint32_t* ptrs = (uint32_t*)guarded((NL + NS) * 4);
        uint64_t* masks = (uint64_t*)guarded((NL + NS) * 8);
        uint32_t nn = 0;
        OK(c, rbh_get_light_tree(c, nodes, &nn, ptrs, masks));
        check_guard(nodes, 2 * (NL + NS) * sizeof(rbh_packed_node), "rbh_get_light_tree nodes (shapes)");
        check_guard(ptrs, (NL + NS) * 4, "rbh_get_light_tree lightPointers (shapes)");
        check_guard(masks, (NL + NS) * 8, "rbh_get_light_tree bitmasks (shapes)");
        OK(c, rbh_set_light_tree(c, nodes, nn, ptrs, masks, lt, NL));
        OK(c, rbh_render(c, &lp));
        s[1].radius = -1.0f;
        EXPECT(rbh_upload_shape_lights(c, s, sflux, NS) == RBH_ERR_INVALID, "negative radius accepted");
        OK(c, rbh_upload_shape_lights(c, NULL, NULL, 0));
        OK(c, rbh_build_light_tree(c, lt, flux, NL));
        free(nodes); free(ptrs); free(masks);
    }
    rbh_tuning tu;
    OK(c, rbh_get_tuning(c, &tu));
    OK(c, rbh_set_tuning(c, &tu));
    OK(c, rbh_autotune(c, &lp, 1, &tu));
    /* two scenes in one batch: the context twice */
    rbh_ctx both[2] = {c, c};
    rbh_launch_params lps[2] = {lp, lp};
    double* done = (double*)guarded(2 * sizeof(double));
    OK(c, rbh_render_batch(c, both, lps, 2, done));
    check_guard(done, 2 * sizeof(double), "rbh_render_batch completionMs");
    /* the no-allocation build takes device buffers only; null ones are refused */
    EXPECT(rbh_envmap_build_noalloc(c, NULL, ew, eh, 3, 0, NULL, NULL, NULL, NULL) == RBH_ERR_INVALID,
           "rbh_envmap_build_noalloc accepted null buffers");
    OK(c, rbh_envmap_disable(c));
    /* guiding */
    rbh_guide_config gc;
    memset(&gc, 0, sizeof(gc));
    gc.lr = 1e-3f; gc.beta1 = 0.9f; gc.beta2 = 0.999f; gc.eps = 1e-8f;
    gc.sceneMin[0] = gc.sceneMin[1] = gc.sceneMin[2] = -5.0f;
    gc.sceneMax[0] = gc.sceneMax[1] = gc.sceneMax[2] = 5.0f;
    gc.seed = 7;
    gc.recordCapacity = 4096;
    OK(c, rbh_guide_configure(c, &gc));
    float* params = (float*)guarded(RBH_GUIDE_PARAMS * 4);
    OK(c, rbh_guide_get_params(c, params, RBH_GUIDE_PARAMS));
    check_guard(params, RBH_GUIDE_PARAMS * 4, "rbh_guide_get_params");
    OK(c, rbh_guide_set_params(c, params, RBH_GUIDE_PARAMS));
    const uint32_t n = 100;
    float* rec = (float*)calloc((size_t)n * RBH_GUIDE_RECORD, 4);
    for (uint32_t i = 0; i < n; ++i) {
        float* r = rec + (size_t)i * RBH_GUIDE_RECORD;
        for (int k = 0; k < 3; ++k) r[k] = 0.01f * (float)((i * (k + 3)) % 97);
        r[5] = 1.0f; r[8] = 1.0f; r[9] = 0.5f; r[10] = 0.3f;
        r[12] = 0.6f; r[13] = 0.8f; r[14] = 1.0f; r[15] = 0.3f;
    }
    float* grad = (float*)guarded(RBH_GUIDE_PARAMS * 4);
    float loss = 0.0f;
    OK(c, rbh_guide_gradient(c, rec, n, grad, &loss));
    check_guard(grad, RBH_GUIDE_PARAMS * 4, "rbh_guide_gradient");
    EXPECT(isfinite(loss), "loss %g", loss);
    float* out = (float*)guarded((size_t)n * RBH_GUIDE_OUTPUTS * 4);
    OK(c, rbh_guide_infer(c, rec, n, out));
    check_guard(out, (size_t)n * RBH_GUIDE_OUTPUTS * 4, "rbh_guide_infer");
    OK(c, rbh_guide_infer_fp8(c, rec, n, out));
    check_guard(out, (size_t)n * RBH_GUIDE_OUTPUTS * 4, "rbh_guide_infer_fp8");
    OK(c, rbh_guide_train(c, rec, n, 64, 2, &loss));
    lp.flags = RBH_RENDER_GUIDED | RBH_RENDER_COLLECT;
    OK(c, rbh_render(c, &lp));
    uint32_t nrec = 0;
    OK(c, rbh_guide_records(c, NULL, 0, &nrec));
    const uint32_t keep = nrec < 64 ? nrec : 64;
    float* recs = (float*)guarded((size_t)(keep ? keep : 1) * RBH_GUIDE_RECORD * 4);
    OK(c, rbh_guide_records(c, recs, keep, &nrec));
    check_guard(recs, (size_t)(keep ? keep : 1) * RBH_GUIDE_RECORD * 4, "rbh_guide_records");
    if (nrec) OK(c, rbh_guide_train_collected(c, 64, 1, &loss));
    OK(c, rbh_destroy(c));
    free(dense); free(xyz); free(basis); free(lut2); free(lut3); free(sheen); free(env); free(acc); free(rgba);
    free(done); free(params); free(rec); free(grad); free(out); free(recs);
}

int main(int argc, char** argv) {
    const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    host_part();
    if (gpu) gpu_part();
    if (failures) {
        fprintf(stderr, "abi_check: %d failure(s)\n", failures);
        return 1;
    }
    printf("abi_check: ok\n");
    return 0;
}
