// C ABI entry points (include/ratbite_hip.h): context lifetime, uploads and
// the environment-map tables. Rendering entry points live in render_api.cpp.
#include "context.hpp"
#include "guide.hpp"
#include <cstring>
#include <new>
#include <vector>
#include <algorithm>

namespace rbh {
hipError_t envmap_cdf_launch(rbh_ctx c, bool fp64);
hipError_t srgb_lut_fill(rbh_ctx c);
size_t envmap_scratch_bytes(int h);
hipError_t envmap_noalloc_init(int device);
hipError_t refresh_tri_classes(rbh_ctx c);
hipError_t envmap_cdf_noalloc(int device, hipStream_t s, bool fp64, const float* texels, int w, int h, int ch,
                              uint16_t* cond, uint16_t* marg, float* cdfY, double* normIntegral);
}  // namespace rbh

using namespace rbh;

static_assert(sizeof(rbh_material) == sizeof(Material), "rbh_material layout");
static_assert(sizeof(rbh_texture_desc) == sizeof(TextureDesc), "rbh_texture_desc layout");
static_assert(sizeof(rbh_emissive_tri) == sizeof(EmissiveTri), "rbh_emissive_tri layout");
static_assert(sizeof(rbh_packed_node) == sizeof(PackedNode), "rbh_packed_node layout");

#define CHECK_CTX(c) \
    do { if (!(c)) return RBH_ERR_INVALID; } while (0)

extern "C" {

int rbh_create(int device, rbh_ctx* out) {
    if (!out) return RBH_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RBH_ERR_HIP;
    if (device < 0 || device >= n) return RBH_ERR_INVALID;
    rbh_context* c = new (std::nothrow) rbh_context();
    if (!c) return RBH_ERR_NOMEM;
    c->device = device;
    // envmap_noalloc_init: this device's address of the no-allocation CDF
    // build's static scratch array, looked up once per device
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        envmap_noalloc_init(device) != hipSuccess) {
        delete c;
        return RBH_ERR_HIP;
    }
    *out = c;
    return RBH_OK;
}

int rbh_destroy(rbh_ctx c) {
    CHECK_CTX(c);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->guide) {
        rbh::guide_release(*static_cast<rbh::GuideState*>(c->guide));
        delete static_cast<rbh::GuideState*>(c->guide);
        c->guide = nullptr;
    }
    hipStream_t s = c->stream;
    delete c;  // DevBuf destructors free device memory
    if (s) (void)hipStreamDestroy(s);
    return RBH_OK;
}

const char* rbh_last_error(rbh_ctx c) { return c ? c->err.c_str() : "null context"; }

int rbh_query(rbh_ctx c) {
    CHECK_CTX(c);
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(c, RBH_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int rbh_synchronize(rbh_ctx c) {
    CHECK_CTX(c);
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

int rbh_upload_spectra(rbh_ctx c, const float* dense, uint32_t n) {
    CHECK_CTX(c);
    if (n && !dense) return fail(c, RBH_ERR_INVALID, "spectra pointer is null");
    if (n > 65535) return fail(c, RBH_ERR_RANGE, "at most 65535 spectra (16-bit indices)");
    RBH_ALLOC(c, c->spectra, (size_t)(n ? n : 1) * kDense * sizeof(float));
    if (n) RBH_HIP(c, hipMemcpyAsync(c->spectra.p, dense, (size_t)n * kDense * sizeof(float), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->nspectra = n;
    return RBH_OK;
}

int rbh_upload_color_tables(rbh_ctx c, const float* xyz, const float* basis) {
    CHECK_CTX(c);
    if (!xyz || !basis) return fail(c, RBH_ERR_INVALID, "colour tables are null");
    RBH_ALLOC(c, c->colorTables, 6 * kDense * sizeof(float));
    RBH_HIP(c, hipMemcpyAsync(c->colorTables.p, xyz, 3 * kDense * sizeof(float), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(c->colorTables.as<float>() + 3 * kDense, basis, 3 * kDense * sizeof(float), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->haveColor = true;
    return RBH_OK;
}

int rbh_upload_luts(rbh_ctx c, const float* cond, const float* dOut, const float* dIn, const float* rOut, const float* rIn,
                    const float* sheen) {
    CHECK_CTX(c);
    if (!cond || !dOut || !dIn || !rOut || !rIn || !sheen) return fail(c, RBH_ERR_INVALID, "a LUT pointer is null");
    RBH_ALLOC(c, c->luts, kLutTotal * sizeof(float));
    float* d = c->luts.as<float>();
    struct { const float* src; size_t off, n; } parts[] = {
        {cond, kLutCond, 32 * 32}, {dOut, kLutDielOut, 32 * 32 * 32}, {dIn, kLutDielIn, 32 * 32 * 32},
        {rOut, kLutReflOut, 32 * 32 * 32}, {rIn, kLutReflIn, 32 * 32 * 32}, {sheen, kLutSheen, 32 * 32 * 4}};
    for (auto& p : parts) RBH_HIP(c, hipMemcpyAsync(d + p.off, p.src, p.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->haveLuts = true;
    return RBH_OK;
}

int rbh_upload_textures(rbh_ctx c, const rbh_texture_desc* descs, uint32_t n, const void* blob, uint64_t blobBytes) {
    CHECK_CTX(c);
    if (n && (!descs || !blob)) return fail(c, RBH_ERR_INVALID, "texture arrays are null");
    static const uint32_t bpp[5] = {4, 1, 16, 12, 4};
    for (uint32_t i = 0; i < n; ++i) {
        const rbh_texture_desc& t = descs[i];
        if (t.format > 4 || t.width == 0 || t.height == 0) return fail(c, RBH_ERR_INVALID, "texture %u: bad format/size", i);
        uint64_t end = t.offset + (uint64_t)t.width * t.height * (t.depth ? t.depth : 1) * bpp[t.format];
        if (end > blobBytes) return fail(c, RBH_ERR_INVALID, "texture %u extends past the texel blob", i);
        if ((t.format >= 2) && (t.offset % 4)) return fail(c, RBH_ERR_INVALID, "texture %u: float texels must be 4-byte aligned", i);
    }
    RBH_ALLOC(c, c->textures, (size_t)(n ? n : 1) * sizeof(TextureDesc));
    RBH_ALLOC(c, c->texels, blobBytes ? blobBytes : 16);
    if (n) RBH_HIP(c, hipMemcpyAsync(c->textures.p, descs, (size_t)n * sizeof(TextureDesc), hipMemcpyHostToDevice, c->stream));
    if (blobBytes) RBH_HIP(c, hipMemcpyAsync(c->texels.p, blob, blobBytes, hipMemcpyHostToDevice, c->stream));
    RBH_ALLOC(c, c->srgbLut, 256 * sizeof(float));
    RBH_HIP(c, rbh::srgb_lut_fill(c));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    c->ntextures = n;
    return RBH_OK;
}

int rbh_upload_materials(rbh_ctx c, const rbh_material* m, uint32_t n) {
    CHECK_CTX(c);
    if (n == 0 || !m) return fail(c, RBH_ERR_INVALID, "no materials");
    for (uint32_t i = 0; i < n; ++i) {
        if (m[i].bxdf > 2) return fail(c, RBH_ERR_INVALID, "material %u: bxdf index %u", i, m[i].bxdf);
        for (int k = 0; k < 7; ++k)
            if ((m[i].textures >> k) & 1u)
                if (m[i].tex[k] < 0 || (uint32_t)m[i].tex[k] >= c->ntextures)
                    return fail(c, RBH_ERR_INVALID, "material %u: texture slot %d references %d of %u", i, k, m[i].tex[k], c->ntextures);
        if (m[i].bxdf == 0 && (m[i].iorSpectrum >= c->nspectra || m[i].acSpectrum >= c->nspectra))
            return fail(c, RBH_ERR_INVALID, "material %u: conductor spectra out of range", i);
        if (m[i].bxdf == 1 && m[i].iorSpectrum >= c->nspectra)
            return fail(c, RBH_ERR_INVALID, "material %u: dielectric ior spectrum out of range", i);
        if ((m[i].flags & 4u) && m[i].emissionSpectrum >= c->nspectra)
            return fail(c, RBH_ERR_INVALID, "material %u: emission spectrum out of range", i);
    }
    RBH_ALLOC(c, c->materials, (size_t)n * sizeof(Material));
    RBH_HIP(c, hipMemcpyAsync(c->materials.p, m, (size_t)n * sizeof(Material), hipMemcpyHostToDevice, c->stream));
    c->nmaterials = n;
    RBH_HIP(c, refresh_tri_classes(c));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    return RBH_OK;
}

int rbh_upload_geometry_ex(rbh_ctx c, const float* verts, uint32_t nverts, const void* indexData, uint32_t indexFormat,
                           uint32_t ntris, const void* attribs, uint64_t attribBytes, const uint32_t* triMaterial,
                           const uint32_t* triLight) {
    CHECK_CTX(c);
    if (indexFormat != RBH_INDEX_U32 && indexFormat != RBH_INDEX_U16)
        return fail(c, RBH_ERR_INVALID, "unknown index format %u", indexFormat);
    if (!verts || !indexData || !triMaterial || !triLight || nverts == 0 || ntris == 0)
        return fail(c, RBH_ERR_INVALID, "geometry arrays are null or empty");
    if (ntris > (1u << 30)) return fail(c, RBH_ERR_RANGE, "at most 2^30 triangles");
    // 16-bit index buffers (pathtrace.cu:1295-1299) are widened once here; the
    // leaf-order triangle records the kernels read hold 32-bit indices
    std::vector<uint32_t> widened;
    const uint32_t* indices = static_cast<const uint32_t*>(indexData);
    if (indexFormat == RBH_INDEX_U16) {
        const uint16_t* i16 = static_cast<const uint16_t*>(indexData);
        widened.assign(i16, i16 + 3ull * ntris);
        indices = widened.data();
    }
    for (uint64_t i = 0; i < 3ull * ntris; ++i)
        if (indices[i] >= nverts) return fail(c, RBH_ERR_INVALID, "index %llu out of range", (unsigned long long)i);
    int64_t maxLight = -1;
    for (uint32_t i = 0; i < ntris; ++i) {
        if (triMaterial[i] >= c->nmaterials) return fail(c, RBH_ERR_INVALID, "triangle %u: material %u of %u", i, triMaterial[i], c->nmaterials);
        if (triLight[i] != 0xFFFFFFFFu) {
            if (triLight[i] >= (1u << 30)) return fail(c, RBH_ERR_RANGE, "triangle %u: light index above 2^30", i);
            maxLight = std::max<int64_t>(maxLight, triLight[i]);
        }
    }
    // every material's vertex attributes must lie inside the attribute buffer
    {
        std::vector<Material> mats(c->nmaterials);
        RBH_HIP(c, hipMemcpy(mats.data(), c->materials.p, mats.size() * sizeof(Material), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < c->nmaterials; ++i) {
            const Material& m = mats[i];
            struct { uint32_t bit, off, size; } at[] = {{1u, m.normalOffset, 8u}, {2u, m.frameOffset, 16u},
                                                        {8u, m.texCoord1Offset, 8u}, {16u, m.texCoord2Offset, 8u}};
            for (auto& a : at) {
                if (!(m.attributes & a.bit)) continue;
                const uint64_t end = (uint64_t)(nverts - 1) * m.attributeStride + a.off + a.size;
                if (end > attribBytes || (a.off % 4) || (m.attributeStride % 4))
                    return fail(c, RBH_ERR_INVALID, "material %u: attribute (bit %u) at offset %u, stride %u exceeds %llu attribute bytes",
                                i, a.bit, a.off, m.attributeStride, (unsigned long long)attribBytes);
            }
        }
    }
    Geometry& g = c->geom;
    RBH_ALLOC(c, g.verts, (size_t)nverts * 12);
    RBH_ALLOC(c, g.indices, (size_t)ntris * 12);
    RBH_ALLOC(c, g.attribs, attribBytes ? attribBytes : 16);
    RBH_ALLOC(c, g.triMaterial, (size_t)ntris * 4);
    RBH_ALLOC(c, g.triLight, (size_t)ntris * 4);
    RBH_HIP(c, hipMemcpyAsync(g.verts.p, verts, (size_t)nverts * 12, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(g.indices.p, indices, (size_t)ntris * 12, hipMemcpyHostToDevice, c->stream));
    if (attribBytes) RBH_HIP(c, hipMemcpyAsync(g.attribs.p, attribs, attribBytes, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(g.triMaterial.p, triMaterial, (size_t)ntris * 4, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemcpyAsync(g.triLight.p, triLight, (size_t)ntris * 4, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    g.nverts = nverts;
    g.ntris = ntris;
    g.maxLight = maxLight;
    c->accel.built = false;
    return RBH_OK;
}

int rbh_upload_geometry(rbh_ctx c, const float* verts, uint32_t nverts, const uint32_t* indices, uint32_t ntris,
                        const void* attribs, uint64_t attribBytes, const uint32_t* triMaterial, const uint32_t* triLight) {
    return rbh_upload_geometry_ex(c, verts, nverts, indices, RBH_INDEX_U32, ntris, attribs, attribBytes, triMaterial,
                                  triLight);
}

// ------------------------------------------------------------ env map
static int envmap_run(rbh_ctx c, uint32_t flags) {
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    RBH_HIP(c, envmap_cdf_launch(c, (flags & RBH_CDF_FP64) != 0));
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    return RBH_OK;
}

static int envmap_finish(rbh_ctx c, uint32_t flags) {
    // The reference reads the integral right after an async copy
    // (rendering_interface.cpp:1404-1405); here the stream is synchronised.
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    RBH_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.cdfMs = ms;
    if (flags & RBH_CDF_FP64) {
        // the FP64 total divided in double; the renderer rounds it to float once
        double raw = 0.0;
        RBH_HIP(c, hipMemcpy(&raw, c->env.integralDev.p, sizeof(double), hipMemcpyDeviceToHost));
        c->env.integral = raw / ((double)c->env.width * (double)c->env.height);
    } else {
        float f = 0.0f;
        RBH_HIP(c, hipMemcpy(&f, c->env.integralDev.p, sizeof(float), hipMemcpyDeviceToHost));
        c->env.integral = f / ((float)c->env.width * (float)c->env.height);
    }
    c->env.enabled = true;
    return RBH_OK;
}

int rbh_envmap_build(rbh_ctx c, const float* texels, int w, int h, int ch, uint32_t flags) {
    CHECK_CTX(c);
    if (!texels) return fail(c, RBH_ERR_INVALID, "texels are null");
    if (ch != 3 && ch != 4) return fail(c, RBH_ERR_INVALID, "channels must be 3 or 4");
    if (w < 4 || h < 4 || (w % 4) || (h % 4)) return fail(c, RBH_ERR_INVALID, "env-map dimensions must be multiples of 4 (got %dx%d)", w, h);
    if (w > 32768) return fail(c, RBH_ERR_RANGE, "env-map width above 32768");
    if (h > 32768) return fail(c, RBH_ERR_RANGE, "env-map height above 32768");
    EnvMap& e = c->env;
    e.enabled = false;
    RBH_ALLOC(c, e.texels, (size_t)w * h * ch * sizeof(float));
    RBH_ALLOC(c, e.cond, (size_t)w * h * sizeof(uint16_t));
    RBH_ALLOC(c, e.marg, (size_t)h * sizeof(uint16_t));
    RBH_ALLOC(c, e.cdfY, (size_t)h * sizeof(float));
    RBH_ALLOC(c, e.integralDev, envmap_scratch_bytes(h));
    e.width = w; e.height = h; e.channels = ch;
    RBH_HIP(c, hipMemcpyAsync(e.texels.p, texels, (size_t)w * h * ch * sizeof(float), hipMemcpyHostToDevice, c->stream));
    int r = envmap_run(c, flags);
    if (r) return r;
    return envmap_finish(c, flags);
}

int rbh_envmap_rebuild(rbh_ctx c, uint32_t flags, int iterations) {
    CHECK_CTX(c);
    if (!c->env.texels.p) return fail(c, RBH_ERR_STATE, "no environment map uploaded");
    if (iterations < 1) iterations = 1;
    RBH_HIP(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < iterations; ++i) RBH_HIP(c, envmap_cdf_launch(c, (flags & RBH_CDF_FP64) != 0));
    RBH_HIP(c, hipEventRecord(c->ev1, c->stream));
    int r = envmap_finish(c, flags);
    c->stats.cdfMs /= iterations;
    return r;
}

int rbh_envmap_build_noalloc(rbh_ctx c, const float* texels, int w, int h, int ch, uint32_t flags, uint16_t* cond,
                             uint16_t* marg, float* cdfY, double* integralOut) {
    CHECK_CTX(c);
    if (!texels || !cond || !marg || !cdfY || !integralOut) return fail(c, RBH_ERR_INVALID, "a device buffer is null");
    if (flags & ~(uint32_t)RBH_CDF_FP64) return fail(c, RBH_ERR_INVALID, "unknown CDF flags 0x%x", flags);
    if (ch != 3 && ch != 4) return fail(c, RBH_ERR_INVALID, "channels must be 3 or 4");
    if (w < 4 || h < 4 || (w % 4) || (h % 4)) return fail(c, RBH_ERR_INVALID, "env-map dimensions must be multiples of 4 (got %dx%d)", w, h);
    if (w > 32768) return fail(c, RBH_ERR_RANGE, "env-map width above 32768");
    if (h > RBH_ENVMAP_NOALLOC_MAX_ROWS)
        return fail(c, RBH_ERR_RANGE, "env-map height %d above the %d rows of the static scratch", h,
                    (int)RBH_ENVMAP_NOALLOC_MAX_ROWS);
    const hipError_t e = envmap_cdf_noalloc(c->device, c->stream, (flags & RBH_CDF_FP64) != 0, texels, w, h, ch, cond,
                                            marg, cdfY, integralOut);
    if (e != hipSuccess) return fail(c, RBH_ERR_HIP, "no-allocation CDF build: %s", hipGetErrorString(e));
    return RBH_OK;
}

int rbh_envmap_read(rbh_ctx c, uint16_t* cond, uint16_t* marg, float* cdfY, double* integral) {
    CHECK_CTX(c);
    if (!c->env.enabled) return fail(c, RBH_ERR_STATE, "no environment map built");
    const EnvMap& e = c->env;
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    if (cond) RBH_HIP(c, hipMemcpy(cond, e.cond.p, (size_t)e.width * e.height * 2, hipMemcpyDeviceToHost));
    if (marg) RBH_HIP(c, hipMemcpy(marg, e.marg.p, (size_t)e.height * 2, hipMemcpyDeviceToHost));
    if (cdfY) RBH_HIP(c, hipMemcpy(cdfY, e.cdfY.p, (size_t)e.height * 4, hipMemcpyDeviceToHost));
    if (integral) *integral = e.integral;
    return RBH_OK;
}

int rbh_envmap_disable(rbh_ctx c) {
    CHECK_CTX(c);
    c->env.enabled = false;
    return RBH_OK;
}

int rbh_stats_get(rbh_ctx c, rbh_stats* out) {
    CHECK_CTX(c);
    if (!out) return RBH_ERR_INVALID;
    *out = c->stats;
    return RBH_OK;
}

}  // extern "C"
