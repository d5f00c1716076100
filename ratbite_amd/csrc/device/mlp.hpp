// Shared device code of the guiding MLP: LDS staging, input encoding, the
// block-tiled BF16 MFMA forward pass of the training kernel
// (v_mfma_f32_16x16x32_bf16; rows = records, columns = features) and the FP8
// activation scale. Inference runs the wave-resident form in mlp_wave.hpp.
#pragma once
#include <hip/hip_runtime.h>
#include "../guide.hpp"

namespace rbh {
namespace mlp {

using namespace guide;

typedef __bf16 bf16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kWaves = 4;
constexpr int kRows = 16;                  // records per wave per tile
constexpr int kTile = kWaves * kRows;      // records per block tile
constexpr int kSX0 = 96 + 8;               // LDS row strides in bf16 (16 B multiples, staggered banks)
constexpr int kSH = 64 + 8;
constexpr int kSO = kOut + 1;              // f32 output rows
constexpr int kGradTiles = 12 + 16 * 3 + 20;   // weight-gradient 16x16 tiles over all layers
constexpr int kTilesPerWave = kGradTiles / kWaves;
static_assert(kGradTiles % kWaves == 0, "gradient tiles split evenly");

struct Lds {
    bf16 x0[kTile * kSX0];
    bf16 h[4][kTile * kSH];      // inputs of layers 1..4 (post-ReLU)
    bf16 dy[2][kTile * kSH];     // output gradients, ping-pong
    float out[kTile * kSO];      // raw network outputs
    float red[kWaves];
};

// ------------------------------------------------------------- encoding
// feature f of a record (positions already normalised to [0,1]^3)
__device__ __forceinline__ float enc_feature(const float* __restrict__ rec, int f) {
    if (f < 72) {
        const int a = f / 24, k = (f % 24) >> 1;
        const float x = rec[a] * (float)(1 << k);          // exact power-of-two scaling
        const float r = x - 2.0f * floorf(x * 0.5f);       // exact reduction modulo 2
        return (f & 1) ? cospif(r) : sinpif(r);
    }
    if (f < 78) return rec[3 + (f - 72)];                  // normal, wo
    if (f == 78) return rec[9];                            // roughness
    if (f == 79) return rec[10];                           // wavelength
    return 0.0f;
}

__device__ __forceinline__ void encode_rows(bf16* x0, const float* __restrict__ recs, uint32_t n, uint32_t base,
                                            int row0) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    const uint32_t s = base + row0 + r;
    const float* rec = recs + (size_t)s * kRec;
    for (int f = g; f < 96; f += 4) {
        const float v = (s < n && f < kIn) ? enc_feature(rec, f) : 0.0f;
        x0[(row0 + r) * kSX0 + f] = (bf16)v;
    }
}

// ------------------------------------------------------------- forward
// Y[rows][N] = act(X[rows][K] W^T + b) for this wave's 16 rows
template <int K, int N, bool RELU, bool F32OUT>
__device__ __forceinline__ void fwd_layer(const bf16* X, int sx, const bf16* __restrict__ Wf,
                                          const float* __restrict__ bias, void* Y, int sy, int row0) {
    __syncthreads();   // X is complete (written by other lanes); every thread of the block calls this
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int ct = 0; ct < N / 16; ++ct) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < K / 32; ++kb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(X + (row0 + r) * sx + kb * 32 + 8 * g);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(Wf + (ct * 16 + r) * K + kb * 32 + 8 * g);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        }
        const int col = ct * 16 + r;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + 4 * g + i;
            float v = c[i] + bb;
            if (RELU) v = fmaxf(v, 0.0f);
            if (F32OUT) {
                if (col < kOut) reinterpret_cast<float*>(Y)[row * sy + col] = v;
            } else {
                reinterpret_cast<bf16*>(Y)[row * sy + col] = (bf16)v;
            }
        }
    }
}

// The five layers for this wave's 16 rows: x0 -> h0 -> h1 -> h2 -> h3 -> out.
// Training keeps every layer input (four distinct hidden buffers); inference
// may pass two buffers alternately.
__device__ __forceinline__ void forward_rows(const bf16* x0, bf16* h0, bf16* h1, bf16* h2, bf16* h3, float* out,
                                             const bf16* __restrict__ wf, const float* __restrict__ bias, int row0) {
    fwd_layer<96, 64, true, false>(x0, kSX0, wf + wf_off(0), bias + bp_off(0), h0, kSH, row0);
    fwd_layer<64, 64, true, false>(h0, kSH, wf + wf_off(1), bias + bp_off(1), h1, kSH, row0);
    fwd_layer<64, 64, true, false>(h1, kSH, wf + wf_off(2), bias + bp_off(2), h2, kSH, row0);
    fwd_layer<64, 64, true, false>(h2, kSH, wf + wf_off(3), bias + bp_off(3), h3, kSH, row0);
    fwd_layer<64, 64, false, true>(h3, kSH, wf + wf_off(4), bias + bp_off(4), out, kSO, row0);
}

__device__ __forceinline__ void forward_rows(Lds& L, const bf16* __restrict__ wf, const float* __restrict__ bias,
                                             int row0) {
    forward_rows(L.x0, L.h[0], L.h[1], L.h[2], L.h[3], L.out, wf, bias, row0);
}

// ----------------------------------------------------------------- FP8
// E4M3 inference with per-tensor scales: weights carry one scale per layer,
// activations one scale per layer and 16-record tile (the largest magnitude
// of the tile's rows maps to kFp8Max). kFp8Max stays within the finite range
// of both E4M3 encodings (240 and 448).
constexpr float kFp8Max = 240.0f;

__device__ __forceinline__ float wave_amax(const bf16* X, int sx, int K, int row0) {
    const int lane = __lane_id();
    float m = 0.0f;
    for (int q = lane; q < kRows * K; q += 64) m = fmaxf(m, fabsf((float)X[(row0 + q / K) * sx + q % K]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    return m;
}

}  // namespace mlp
}  // namespace rbh
