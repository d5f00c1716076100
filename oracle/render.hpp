// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header). Shading-side
// declarations shared by the oracle's path tracer and its unit entry points.
#pragma once
#include "rb_math.hpp"
#include "rb_oracle.h"

namespace orc {

// PathStateBitfield, src/kernels/pathtrace.cu:25-37
enum PathFlags : uint32_t {
    PF_PREVIOUS_HIT_SPECULAR = 1u << 0, PF_CURRENT_HIT_SPECULAR = 1u << 1, PF_REGULARIZED = 1u << 2,
    PF_RAY_REFRACTED = 1u << 3, PF_PATH_TERMINATED = 1u << 4, PF_TRIANGULAR_GEOMETRY = 1u << 5,
    PF_RIGHT_HANDED_FRAME = 1u << 6, PF_REFRACTION_HAPPENED = 1u << 7, PF_FINISHED = 1u << 8,
};
// MaterialData bitfields, src/core/material.h:19-57
enum AttrBits : uint32_t { AT_NORMAL = 1, AT_FRAME = 2, AT_COLOR = 4, AT_TEX1 = 8, AT_TEX2 = 16 };
enum TexBits : uint32_t { TX_BASE_COLOR = 1, TX_NORMAL = 2, TX_MET_ROUGH = 4, TX_TRANSMISSION = 8,
                          TX_SHEEN_COLOR = 16, TX_SHEEN_ROUGH = 32, TX_EMISSION = 64 };
enum FactorBits : uint32_t { FC_BASE_COLOR = 1, FC_METALNESS = 2, FC_ROUGHNESS = 4, FC_TRANSMISSION = 8,
                             FC_CUTOFF = 16, FC_BLEND = 32, FC_SHEEN_COLOR = 64, FC_SHEEN_ROUGH = 128, FC_EMISSION = 256 };
enum TexSlot { TS_BASE = 0, TS_NORMAL = 1, TS_MR = 2, TS_TRANS = 3, TS_SHC = 4, TS_SHR = 5, TS_EM = 6 };
enum MatFlags : uint32_t { MF_DOUBLE_SIDED = 1, MF_SHEEN = 2, MF_SPECTRAL_EMISSION = 4 };

// microsurface::Surface, src/device/surface.h:380-419
struct Surface {
    uint32_t layers = 1;  // BASE
    V3 baseColor{1.0f, 1.0f, 1.0f};
    float metalness = 1.0f;
    float ior = 1.5f, roughness = 1.0f;
    float transmission = 0.0f;
    V3 sheenColor{1.0f, 1.0f, 1.0f};
    float sheenAlpha = 0.0f;
};
// DirectLightSampleData, src/kernels/pathtrace.cu:77-83
struct DirectLight {
    Spec spectrum = specs(0.0f);
    V3 dir{0, 0, 0};
    float pdf = 0.0f;
    bool occluded = false;
};

struct PathIO {
    Spec L, T;
    Wavelengths wl;
    float bxdfPDF;
    V3 rD;
    uint32_t flags;
    float refrScale;
};

float tex_sample1(const orc_scene* sc, const float* lut, int nx, int ny, int nz, float u, float v, float w);
void tex_sample(const orc_scene* sc, int32_t tex, float u, float v, float out[4]);

void bxdf_conductor(const orc_scene* sc, const orc_material& m, const DirectLight& dl, const QState& q,
                    const Local& local, PathIO& p);
void bxdf_complex(const orc_scene* sc, const DirectLight& dl, const QState& q, const Local& local,
                  const Surface& s, PathIO& p);
void bxdf_dielectric(const orc_scene* sc, const orc_material& m, const DirectLight& dl, const QState& q,
                     const Local& local, PathIO& p);

}  // namespace orc
