// Wave-resident inference of the guiding MLP. Each wave loads the B operands
// of all five layers into registers once (BF16: 168 VGPRs, FP8: 84) and then
// streams tiles of 16 records through its own LDS rows: encode, five
// v_mfma_f32_16x16x32 layers (A from LDS, B from registers, the A fragment of
// a layer loaded once for all its column tiles), outputs. Waves never wait
// for each other; the LDS hand-over between lanes of one wave is ordered by
// wavefront-scope fences. The arithmetic (bf16 activations, FP32
// accumulation, bias added after the product, per-tile FP8 activation scale)
// is that of the block-tiled forward pass in mlp.hpp.
#pragma once
#include <type_traits>

#include "mlp.hpp"

namespace rbh {
namespace mlp {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS rows of one wave: encoded inputs, hidden activations (ping-pong) and
// the FP32 outputs of the current 16-record tile
struct WaveLds {
    bf16 x0[kRows * kSX0];
    bf16 h[2][kRows * kSH];
    float out[kRows * kSO];
};

// ------------------------------------------------------------------ BF16
struct WaveWeightsBf16 {
    bf16x8 w0[4][3], w1[4][2], w2[4][2], w3[4][2], w4[3][2];
    float b0[4], b1[4], b2[4], b3[4], b4[3];
};

template <int NT, int KB>
__device__ __forceinline__ void load_layer(const bf16* __restrict__ wf, const float* __restrict__ bias, int l,
                                           bf16x8 (&w)[NT][KB], float (&b)[NT]) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            w[ct][kb] = *reinterpret_cast<const bf16x8*>(wf + wf_off(l) + (ct * 16 + r) * k_pad(l) + kb * 32 + 8 * g);
        b[ct] = bias[bp_off(l) + ct * 16 + r];
    }
}

__device__ __forceinline__ void load_weights(WaveWeightsBf16& W, const bf16* __restrict__ wf,
                                             const float* __restrict__ bias) {
    load_layer(wf, bias, 0, W.w0, W.b0);
    load_layer(wf, bias, 1, W.w1, W.b1);
    load_layer(wf, bias, 2, W.w2, W.b2);
    load_layer(wf, bias, 3, W.w3, W.b3);
    load_layer(wf, bias, 4, W.w4, W.b4);
}

// Y = act(X W^T + b) for the wave's 16 rows; X rows in LDS with stride sx
template <int NT, int KB, bool RELU, bool F32OUT>
__device__ __forceinline__ void wave_layer(const bf16* X, int sx, const bf16x8 (&w)[NT][KB], const float (&b)[NT],
                                           void* Y, int sy) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    bf16x8 a[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) a[kb] = *reinterpret_cast<const bf16x8*>(X + r * sx + kb * 32 + 8 * g);
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], w[ct][kb], c, 0, 0, 0);
        const int col = ct * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i;
            float v = c[i] + b[ct];
            if (RELU) v = fmaxf(v, 0.0f);
            if (F32OUT) {
                if (col < kOut) reinterpret_cast<float*>(Y)[row * sy + col] = v;
            } else {
                reinterpret_cast<bf16*>(Y)[row * sy + col] = (bf16)v;
            }
        }
    }
    wave_sync();
}

__device__ __forceinline__ void wave_forward(const WaveWeightsBf16& W, WaveLds& L) {
    wave_layer<4, 3, true, false>(L.x0, kSX0, W.w0, W.b0, L.h[0], kSH);
    wave_layer<4, 2, true, false>(L.h[0], kSH, W.w1, W.b1, L.h[1], kSH);
    wave_layer<4, 2, true, false>(L.h[1], kSH, W.w2, W.b2, L.h[0], kSH);
    wave_layer<4, 2, true, false>(L.h[0], kSH, W.w3, W.b3, L.h[1], kSH);
    wave_layer<3, 2, false, true>(L.h[1], kSH, W.w4, W.b4, L.out, kSO);
}

// ------------------------------------------------------------------- FP8
struct WaveWeightsFp8 {
    long w0[4][3], w1[4][2], w2[4][2], w3[4][2], w4[3][2];
    float b0[4], b1[4], b2[4], b3[4], b4[3];
    float s[kLayers];
};

template <int NT, int KB>
__device__ __forceinline__ void load_layer(const uint8_t* __restrict__ wq, const float* __restrict__ bias, int l,
                                           long (&w)[NT][KB], float (&b)[NT]) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            w[ct][kb] = *reinterpret_cast<const long*>(wq + wf_off(l) + (ct * 16 + r) * k_pad(l) + kb * 32 + 8 * g);
        b[ct] = bias[bp_off(l) + ct * 16 + r];
    }
}

__device__ __forceinline__ void load_weights(WaveWeightsFp8& W, const uint8_t* __restrict__ wq,
                                             const float* __restrict__ ws, const float* __restrict__ bias) {
    load_layer(wq, bias, 0, W.w0, W.b0);
    load_layer(wq, bias, 1, W.w1, W.b1);
    load_layer(wq, bias, 2, W.w2, W.b2);
    load_layer(wq, bias, 3, W.w3, W.b3);
    load_layer(wq, bias, 4, W.w4, W.b4);
#pragma unroll
    for (int l = 0; l < kLayers; ++l) W.s[l] = ws[l];
}

template <int NT, int KB, bool RELU, bool F32OUT>
__device__ __forceinline__ void wave_layer_fp8(const bf16* X, int sx, const long (&w)[NT][KB], float wscale,
                                               const float (&b)[NT], void* Y, int sy) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    // activation scale of this 16-row tile: its largest magnitude maps to kFp8Max
    const float amax = wave_amax(X, sx, KB * 32, 0);
    const float xs = amax > 0.0f ? amax / kFp8Max : 1.0f;
    const float inv = 1.0f / xs;
    const float outScale = xs * wscale;
    long a[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const bf16x8 a16 = *reinterpret_cast<const bf16x8*>(X + r * sx + kb * 32 + 8 * g);
        int lo = __builtin_amdgcn_cvt_pk_fp8_f32((float)a16[0] * inv, (float)a16[1] * inv, 0, false);
        lo = __builtin_amdgcn_cvt_pk_fp8_f32((float)a16[2] * inv, (float)a16[3] * inv, lo, true);
        int hi = __builtin_amdgcn_cvt_pk_fp8_f32((float)a16[4] * inv, (float)a16[5] * inv, 0, false);
        hi = __builtin_amdgcn_cvt_pk_fp8_f32((float)a16[6] * inv, (float)a16[7] * inv, hi, true);
        a[kb] = (long)(uint32_t)lo | ((long)(uint32_t)hi << 32);
    }
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) c = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[kb], w[ct][kb], c, 0, 0, 0);
        const int col = ct * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * g + i;
            float v = c[i] * outScale + b[ct];
            if (RELU) v = fmaxf(v, 0.0f);
            if (F32OUT) {
                if (col < kOut) reinterpret_cast<float*>(Y)[row * sy + col] = v;
            } else {
                reinterpret_cast<bf16*>(Y)[row * sy + col] = (bf16)v;
            }
        }
    }
    wave_sync();
}

__device__ __forceinline__ void wave_forward(const WaveWeightsFp8& W, WaveLds& L) {
    wave_layer_fp8<4, 3, true, false>(L.x0, kSX0, W.w0, W.s[0], W.b0, L.h[0], kSH);
    wave_layer_fp8<4, 2, true, false>(L.h[0], kSH, W.w1, W.s[1], W.b1, L.h[1], kSH);
    wave_layer_fp8<4, 2, true, false>(L.h[1], kSH, W.w2, W.s[2], W.b2, L.h[0], kSH);
    wave_layer_fp8<4, 2, true, false>(L.h[0], kSH, W.w3, W.s[3], W.b3, L.h[1], kSH);
    wave_layer_fp8<3, 2, false, true>(L.h[1], kSH, W.w4, W.s[4], W.b4, L.out, kSO);
}

// encodes records [base, base + 16) of recs (rows past n are zero) into x0
__device__ __forceinline__ void wave_encode(bf16* x0, const float* __restrict__ recs, uint32_t n, uint32_t base) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    const uint32_t s = base + r;
    const float* rec = recs + (size_t)s * kRec;
    for (int f = g; f < 96; f += 4) {
        const float v = (s < n && f < kIn) ? enc_feature(rec, f) : 0.0f;
        x0[r * kSX0 + f] = (bf16)v;
    }
    wave_sync();
}

// raw outputs -> mixture (normalised means, clamped sharpness, softmax
// weights) followed by the selection probability of the lobes
__device__ __forceinline__ void to_mixture(const float* __restrict__ o, float* __restrict__ mix) {
    float lmax = -3.4e38f;
    for (int j = 0; j < kLobes; ++j) lmax = fmaxf(lmax, o[5 * j + 4]);
    float ws = 0.0f;
    for (int j = 0; j < kLobes; ++j) ws += expf(o[5 * j + 4] - lmax);
    for (int j = 0; j < kLobes; ++j) {
        const float* q = o + 5 * j;
        const float n = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]), 1e-6f);
        mix[5 * j + 0] = q[0] / n;
        mix[5 * j + 1] = q[1] / n;
        mix[5 * j + 2] = q[2] / n;
        mix[5 * j + 3] = expf(fminf(fmaxf(q[3], kLogKappaMin), kLogKappaMax));
        mix[5 * j + 4] = expf(q[4] - lmax) / ws;
    }
    mix[kMix] = kSelMin + (kSelMax - kSelMin) / (1.0f + expf(-o[kMix]));
}

// Inference over records [0, n) (n read from *nDev when given): MIX writes
// the 41-float mixture per record, otherwise the raw 41 outputs.
template <bool FP8, bool MIX>
__device__ __forceinline__ void wave_infer(WaveLds& L, const float* __restrict__ recs, uint32_t n,
                                           const bf16* __restrict__ wf, const uint8_t* __restrict__ wq,
                                           const float* __restrict__ ws, const float* __restrict__ bias,
                                           float* __restrict__ out) {
    typename std::conditional<FP8, WaveWeightsFp8, WaveWeightsBf16>::type W;
    if constexpr (FP8) load_weights(W, wq, ws, bias);
    else load_weights(W, wf, bias);
    const int lane = __lane_id();
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t t = wave; (size_t)t * kRows < n; t += waves) {
        const uint32_t base = t * kRows;
        wave_encode(L.x0, recs, n, base);
        wave_forward(W, L);
        if (MIX) {
            if (lane < kRows && base + lane < n) to_mixture(L.out + lane * kSO, out + (size_t)(base + lane) * kOut);
        } else {
            for (int q = lane; q < kRows * kOut; q += 64) {
                const int rr = q / kOut, c = q % kOut;
                if (base + rr < n) out[(size_t)(base + rr) * kOut + c] = L.out[rr * kSO + c];
            }
        }
        wave_sync();   // the tile's LDS rows are read before the next tile overwrites them
    }
}

}  // namespace mlp
}  // namespace rbh
