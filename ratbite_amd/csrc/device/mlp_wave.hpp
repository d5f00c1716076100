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
    const float* bias;   // padded per-layer biases (read at use, L1-resident)
};

template <int NT, int KB>
__device__ __forceinline__ void load_layer(const bf16* __restrict__ wf, int l, bf16x8 (&w)[NT][KB]) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            w[ct][kb] = *reinterpret_cast<const bf16x8*>(wf + wf_off(l) + (ct * 16 + r) * k_pad(l) + kb * 32 + 8 * g);
    }
}

__device__ __forceinline__ void load_weights(WaveWeightsBf16& W, const bf16* __restrict__ wf,
                                             const float* __restrict__ bias) {
    load_layer(wf, 0, W.w0);
    load_layer(wf, 1, W.w1);
    load_layer(wf, 2, W.w2);
    load_layer(wf, 3, W.w3);
    load_layer(wf, 4, W.w4);
    W.bias = bias;
}

// Y = act(X W^T + b) for the wave's 16 rows; X rows in LDS with stride sx
template <int NT, int KB, bool RELU, bool F32OUT>
__device__ __forceinline__ void wave_layer(const bf16* X, int sx, const bf16x8 (&w)[NT][KB],
                                           const float* __restrict__ bl, void* Y, int sy) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    float b[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) b[ct] = bl[ct * 16 + r];
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
    wave_layer<4, 3, true, false>(L.x0, kSX0, W.w0, W.bias + bp_off(0), L.h[0], kSH);
    wave_layer<4, 2, true, false>(L.h[0], kSH, W.w1, W.bias + bp_off(1), L.h[1], kSH);
    wave_layer<4, 2, true, false>(L.h[1], kSH, W.w2, W.bias + bp_off(2), L.h[0], kSH);
    wave_layer<4, 2, true, false>(L.h[0], kSH, W.w3, W.bias + bp_off(3), L.h[1], kSH);
    wave_layer<3, 2, false, true>(L.h[1], kSH, W.w4, W.bias + bp_off(4), L.out, kSO);
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

// The record of this lane's row (lane & 15) in registers, loaded with four
// 16-byte reads issued together (zeros past n).
struct WaveRec {
    float4 q[4];
};
__device__ __forceinline__ WaveRec load_rec(const float* __restrict__ recs, uint32_t n, uint32_t base) {
    WaveRec w;
    const uint32_t s = base + (__lane_id() & 15);
    const float4* p = reinterpret_cast<const float4*>(recs + (size_t)s * kRec);
#pragma unroll
    for (int k = 0; k < 4; ++k) w.q[k] = s < n ? p[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return w;
}

// encodes the wave's 16 records into x0: the features of enc_feature, each
// sin/cos pair from one sincospif; lane (r, g) writes pairs g, g + 4, ... of
// record r and features 72 + g, 76 + g, ...
__device__ __forceinline__ void wave_encode(bf16* x0, const WaveRec& w, bool live) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
    const float pos[3] = {w.q[0].x, w.q[0].y, w.q[0].z};
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int p = g + 4 * j;          // pair index: axis j / 3, frequency p - 12 (j / 3)
        const int k = p - 12 * (j / 3);
        float sn = 0.0f, cs = 0.0f;
        if (live) {
            const float x = pos[j / 3] * (float)(1 << k);   // exact power-of-two scaling
            const float rr = x - 2.0f * floorf(x * 0.5f);   // exact reduction modulo 2
            sincospif(rr, &sn, &cs);
        }
        x0[r * kSX0 + 2 * p] = (bf16)sn;
        x0[r * kSX0 + 2 * p + 1] = (bf16)cs;
    }
    // features 72..79: normal (rec 3..5), wo (6..8), roughness (9), wavelength (10)
    const float d[8] = {w.q[0].w, w.q[1].x, w.q[1].y, w.q[1].z, w.q[1].w, w.q[2].x, w.q[2].y, w.q[2].z};
    const float d0 = g == 0 ? d[0] : g == 1 ? d[1] : g == 2 ? d[2] : d[3];
    const float d1 = g == 0 ? d[4] : g == 1 ? d[5] : g == 2 ? d[6] : d[7];
    x0[r * kSX0 + 72 + g] = (bf16)(live ? d0 : 0.0f);
    x0[r * kSX0 + 76 + g] = (bf16)(live ? d1 : 0.0f);
#pragma unroll
    for (int f = 80; f < 96; f += 4) x0[r * kSX0 + f + g] = (bf16)0.0f;
    wave_sync();
}

// The mixture of the wave's 16 records from L.out, all 64 lanes busy: lane
// (r, g) converts lobes 2g and 2g + 1 of record r; the softmax maximum and
// sum run over the record's four lanes (shuffles at distance 16 and 32).
__device__ __forceinline__ void wave_mixture(const float* __restrict__ o, float* __restrict__ mix, bool live) {
    const int g = __lane_id() >> 4;
    float lg[2], m = -3.4e38f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        lg[h] = o[5 * (2 * g + h) + 4];
        m = fmaxf(m, lg[h]);
    }
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    float e[2], sum = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        e[h] = expf(lg[h] - m);
        sum += e[h];
    }
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    if (!live) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int j = 2 * g + h;
        const float* q = o + 5 * j;
        const float n = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]), 1e-6f);
        mix[5 * j + 0] = q[0] / n;
        mix[5 * j + 1] = q[1] / n;
        mix[5 * j + 2] = q[2] / n;
        mix[5 * j + 3] = expf(fminf(fmaxf(q[3], kLogKappaMin), kLogKappaMax));
        mix[5 * j + 4] = e[h] / sum;
    }
    if (g == 0) mix[kMix] = kSelMin + (kSelMax - kSelMin) / (1.0f + expf(-o[kMix]));
}

// Inference over records [0, n) (n read from *nDev when given): MIX writes
// the 41-float mixture per record, otherwise the raw 41 outputs.
template <bool FP8, bool MIX>
__device__ __forceinline__ void wave_infer(WaveLds& L, const float* __restrict__ recs, uint32_t n,
                                           const bf16* __restrict__ wf, const uint8_t* __restrict__ wq,
                                           const float* __restrict__ ws, const float* __restrict__ bias,
                                           float* __restrict__ out) {
    const int lane = __lane_id();
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    if ((size_t)wave * kRows >= n) return;   // no tile for this wave: skip the weight loads
    typename std::conditional<FP8, WaveWeightsFp8, WaveWeightsBf16>::type W;
    if constexpr (FP8) load_weights(W, wq, ws, bias);
    else load_weights(W, wf, bias);
    WaveRec next = load_rec(recs, n, wave * kRows);
    for (uint32_t t = wave; (size_t)t * kRows < n; t += waves) {
        const uint32_t base = t * kRows;
        const WaveRec cur = next;
        next = load_rec(recs, n, base + waves * kRows);   // prefetch the wave's next tile
        wave_encode(L.x0, cur, base + (lane & 15) < n);
        wave_forward(W, L);
        if (MIX) {
            const int r = lane & 15;
            wave_mixture(L.out + r * kSO, out + (size_t)(base + r) * kOut, base + r < n);
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
