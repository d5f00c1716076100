// Path-guiding network on CDNA4 matrix cores: BF16 MFMA forward, fused
// loss + backward with per-block weight-gradient accumulation, reduction +
// Adam, and inference (BF16 or FP8 E4M3 with per-tensor scales).
//
// Work layout: a block is 4 waves and handles tiles of 64 training records;
// each wave owns 16 rows (records) of the tile through the forward pass and
// the backward data pass (v_mfma_f32_16x16x32_bf16, rows = records, columns =
// features, contraction over features), so those passes need no block-wide
// synchronisation. The weight gradients contract over the 64 records of the
// tile (rows = output features, columns = input features); each wave owns a
// fixed fifth of the 80 gradient tiles and keeps them in registers across all
// tiles the block processes, then writes one partial per block. Activations
// are staged in LDS as bf16 [record][feature] rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "radix.hpp"

#include "../context.hpp"
#include "../guide.hpp"
#include "../device/mlp_wave.hpp"

// the sizes the public header documents
static_assert(RBH_GUIDE_RECORD == rbh::guide::kRec, "record size");
static_assert(RBH_GUIDE_OUTPUTS == rbh::guide::kOut, "output row size");
static_assert(RBH_GUIDE_PARAMS == rbh::guide::kParams, "parameter count");

namespace rbh {

using namespace guide;
using namespace mlp;

namespace {

// ------------------------------------------------------------- mixture
// log of kappa / (2 pi (1 - exp(-2 kappa)))
__device__ __forceinline__ float vmf_log_norm(float kappa) {
    return logf(kappa) - 1.8378770664093453f - logf(-expm1f(-2.0f * kappa));
}

struct LobeView {
    float mx, my, mz, nrm, lkraw, kappa;
};
__device__ __forceinline__ LobeView lobe_at(const float* o, int j) {
    const float* q = o + 5 * j;
    LobeView v;
    v.nrm = fmaxf(sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]), 1e-6f);
    v.mx = q[0] / v.nrm;
    v.my = q[1] / v.nrm;
    v.mz = q[2] / v.nrm;
    v.lkraw = q[3];
    v.kappa = expf(fminf(fmaxf(q[3], kLogKappaMin), kLogKappaMax));
    return v;
}

// Loss of one record, weight * -log(a q(d) + (1 - a) p_bsdf(d)) with q the
// lobe mixture and a the selection probability; writes
// d(loss)/d(outputs) * weight / weightSum as bf16 into dst[0..63] (zeros past
// the 41 outputs). The lobes are re-derived from the outputs in each pass to
// keep few values live.
__device__ float record_grad(const float* __restrict__ o, const float* __restrict__ rec, float invWsum, bf16* dst,
                             bool fixedSel) {
    const float dx = rec[11], dy = rec[12], dz = rec[13], wt = rec[14], pb = rec[15];
    float lmax = -3.4e38f;
    for (int j = 0; j < kLobes; ++j) lmax = fmaxf(lmax, o[5 * j + 4]);
    float wsum = 0.0f;
    for (int j = 0; j < kLobes; ++j) wsum += expf(o[5 * j + 4] - lmax);
    const float lw0 = lmax + logf(wsum);            // log of the softmax denominator
    float cmax = -3.4e38f;
    for (int j = 0; j < kLobes; ++j) {
        const LobeView v = lobe_at(o, j);
        const float lc = (o[5 * j + 4] - lw0) + vmf_log_norm(v.kappa) + v.kappa * (v.mx * dx + v.my * dy + v.mz * dz - 1.0f);
        cmax = fmaxf(cmax, lc);
    }
    float se = 0.0f;
    for (int j = 0; j < kLobes; ++j) {
        const LobeView v = lobe_at(o, j);
        const float lc = (o[5 * j + 4] - lw0) + vmf_log_norm(v.kappa) + v.kappa * (v.mx * dx + v.my * dy + v.mz * dz - 1.0f);
        se += expf(lc - cmax);
    }
    const float lse = cmax + logf(se);              // log q
    const float q = expf(lse);
    const float sig = 1.0f / (1.0f + expf(-o[kMix]));
    const float a = kSelMin + (kSelMax - kSelMin) * sig;
    const float P = fmaxf(a * q + (1.0f - a) * pb, 1e-30f);
    const float s0 = wt * invWsum;
    const float s = s0 * (a * q / P);               // the lobes' share of the combined density
    for (int j = 0; j < kLobes; ++j) {
        const LobeView v = lobe_at(o, j);
        const float dot = v.mx * dx + v.my * dy + v.mz * dz;
        const float w = expf(o[5 * j + 4] - lw0);
        const float lc = (o[5 * j + 4] - lw0) + vmf_log_norm(v.kappa) + v.kappa * (dot - 1.0f);
        const float gj = expf(lc - lse);
        const float dlnorm = 1.0f / v.kappa - 2.0f / expm1f(2.0f * v.kappa);
        const float dk = -s * gj * (dlnorm + dot - 1.0f);
        const bool inside = v.lkraw > kLogKappaMin && v.lkraw < kLogKappaMax;
        const float c = -s * gj * v.kappa;
        const float gx = c * dx, gy = c * dy, gz = c * dz;
        const float pr = gx * v.mx + gy * v.my + gz * v.mz;
        dst[5 * j + 0] = (bf16)((gx - v.mx * pr) / v.nrm);
        dst[5 * j + 1] = (bf16)((gy - v.my * pr) / v.nrm);
        dst[5 * j + 2] = (bf16)((gz - v.mz * pr) / v.nrm);
        dst[5 * j + 3] = (bf16)(inside ? dk * v.kappa : 0.0f);
        dst[5 * j + 4] = (bf16)(-s * (gj - w));
    }
    // selection: d(-log P)/da = -(q - p_bsdf) / P, da/dc = (kSelMax - kSelMin) sig (1 - sig)
    // (with a fixed selection the selection output stays zero, a = 0.5)
    dst[kMix] = fixedSel ? (bf16)0.0f : (bf16)(-s0 * (q - pb) / P * (kSelMax - kSelMin) * sig * (1.0f - sig));
    for (int c2 = kOut; c2 < 64; ++c2) dst[c2] = (bf16)0.0f;
    return wt * -logf(P);
}

// ------------------------------------------------------------- backward
// dX[rows][NIN] = (dY[rows][64] W) * (H > 0) for this wave's 16 rows
template <int NIN>
__device__ __forceinline__ void bwd_data(const bf16* dY, const bf16* __restrict__ Wt, const bf16* H, bf16* dX,
                                         int row0) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int ct = 0; ct < NIN / 16; ++ct) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(dY + (row0 + r) * kSH + kb * 32 + 8 * g);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(Wt + (ct * 16 + r) * 64 + kb * 32 + 8 * g);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
        }
        const int col = ct * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + 4 * g + i;
            const float v = (float)H[row * kSH + col] > 0.0f ? c[i] : 0.0f;
            dX[row * kSH + col] = (bf16)v;
        }
    }
}

// gradient tile t -> (layer, row tile over outputs, column tile over inputs)
__device__ __forceinline__ void grad_tile(int t, int& l, int& rt, int& ct) {
    if (t < 12) { l = 4; rt = t / 4; ct = t % 4; return; }
    t -= 12;
    if (t < 48) { l = 3 - t / 16; rt = (t % 16) / 4; ct = t % 4; return; }
    t -= 48;
    l = 0; rt = t / 5; ct = t % 5;
}

// one 16x16 tile of dW_l = dY_l^T X_l over the 64 records of the tile
__device__ __forceinline__ f32x4 grad_tile_mfma(const bf16* dY, const bf16* X, int sx, int rt, int ct, f32x4 acc) {
    const int lane = __lane_id(), r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < kTile / 32; ++kb) {
        bf16x8 a, b;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int s = kb * 32 + 8 * g + q;
            a[q] = dY[s * kSH + rt * 16 + r];
            b[q] = X[s * sx + ct * 16 + r];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_guide_train(const float* __restrict__ recs, uint32_t n,
                                                     const float* __restrict__ scalars, const bf16* __restrict__ wf,
                                                     const bf16* __restrict__ wt, const float* __restrict__ bias,
                                                     float* __restrict__ partials, float* __restrict__ lossOut,
                                                     uint32_t fixedSel) {
    __shared__ Lds L;
    const int wave = threadIdx.x >> 6, lane = __lane_id(), r = lane & 15, g = lane >> 4;
    const int row0 = wave * kRows;
    const float invW = scalars[0] > 0.0f ? 1.0f / scalars[0] : 0.0f;
    f32x4 acc[kTilesPerWave];
#pragma unroll
    for (int j = 0; j < kTilesPerWave; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float accb[kLayers] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float lossAcc = 0.0f;

    for (uint32_t tile = blockIdx.x; (size_t)tile * kTile < n; tile += gridDim.x) {
        const uint32_t base = tile * kTile;
        __syncthreads();   // the previous tile's gradient reads are done
        encode_rows(L.x0, recs, n, base, row0);
        forward_rows(L, wf, bias, row0);
        __syncthreads();
        // output gradient of this wave's rows (lanes 0..15, one record each)
        if (g == 0) {
            const uint32_t s = base + row0 + r;
            bf16* dst = L.dy[0] + (row0 + r) * kSH;
            if (s < n) {
                lossAcc += record_grad(L.out + (row0 + r) * kSO, recs + (size_t)s * kRec, invW, dst, fixedSel != 0u);
            } else {
                for (int c = 0; c < 64; ++c) dst[c] = (bf16)0.0f;
            }
        }
        // layers 4..0: weight/bias gradients from dY_l, then dY_{l-1}
        int cur = 0;
#pragma unroll 1
        for (int l = kLayers - 1; l >= 0; --l) {
            __syncthreads();   // dY_l of all rows is in L.dy[cur]
            const bf16* X = l == 0 ? L.x0 : L.h[l - 1];
            const int sx = l == 0 ? kSX0 : kSH;
#pragma unroll
            for (int j = 0; j < kTilesPerWave; ++j) {
                int tl, rt, ct;
                grad_tile(j * kWaves + wave, tl, rt, ct);
                if (tl == l) acc[j] = grad_tile_mfma(L.dy[cur], X, sx, rt, ct, acc[j]);
            }
            if ((int)threadIdx.x < fan_out(l)) {
                float sb = 0.0f;
                for (int s = 0; s < kTile; ++s) sb += (float)L.dy[cur][s * kSH + threadIdx.x];
#pragma unroll
                for (int q = 0; q < kLayers; ++q)
                    if (q == l) accb[q] += sb;
            }
            if (l > 0) bwd_data<64>(L.dy[cur], wt + wf_off(l), L.h[l - 1], L.dy[cur ^ 1], row0);
            cur ^= 1;
        }
    }

    // partial gradient of this block in the master layout
    float* part = partials + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int j = 0; j < kTilesPerWave; ++j) {
        int l, rt, ct;
        grad_tile(j * kWaves + wave, l, rt, ct);
        const int in = ct * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int out = rt * 16 + 4 * g + i;
            if (out < fan_out(l) && in < fan_in(l)) part[w_off(l) + out * fan_in(l) + in] = acc[j][i];
        }
    }
    for (int l = 0; l < kLayers; ++l)
        if (threadIdx.x < fan_out(l)) part[b_off(l) + threadIdx.x] = accb[l];
    // block loss
    float v = lossAcc;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) L.red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(lossOut, (L.red[0] + L.red[1] + L.red[2] + L.red[3]) * invW);
}

__global__ void k_zero_f32(float* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0f;
}

__global__ __launch_bounds__(256) void k_weight_sum(const float* __restrict__ recs, uint32_t n, float* out) {
    float s = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += recs[(size_t)i * kRec + 14];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0) atomicAdd(out, s);
}

// ------------------------------------------------- collected training set
struct PositiveWeight {
    const float* recs;
    __device__ __forceinline__ bool operator()(uint32_t i) const { return recs[(size_t)i * kRec + 14] > 0.0f; }
};

// out[j] = recs[idx[(j * stride) mod m]]: a fixed permutation (stride is
// coprime with m) so that consecutive batches mix all regions of the image
__global__ __launch_bounds__(256) void k_gather_records(const float* __restrict__ recs, const uint32_t* __restrict__ idx,
                                                        uint32_t m, uint32_t stride, float* __restrict__ out) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        const uint32_t s = idx[(uint32_t)(((uint64_t)j * stride) % m)];
        const float4* src = reinterpret_cast<const float4*>(recs + (size_t)s * kRec);
        float4* dst = reinterpret_cast<float4*>(out + (size_t)j * kRec);
#pragma unroll
        for (int k = 0; k < kRec / 4; ++k) dst[k] = src[k];
    }
}

// w = min(w, mult * mean w), the mean read from *sum
__global__ __launch_bounds__(256) void k_clamp_weights(float* __restrict__ recs, uint32_t m, const float* __restrict__ sum,
                                                       float mult) {
    const float cap = mult * (*sum / (float)m);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x)
        recs[(size_t)j * kRec + 14] = fminf(recs[(size_t)j * kRec + 14], cap);
}

// parameter p of the master layout -> (layer, out, in) or bias (in = -1)
__device__ __forceinline__ void param_index(int p, int& l, int& out, int& in) {
    l = 0;
    while (l + 1 < kLayers && p >= w_off(l + 1)) ++l;
    const int q = p - w_off(l);
    if (q < fan_in(l) * fan_out(l)) { out = q / fan_in(l); in = q % fan_in(l); }
    else { out = q - fan_in(l) * fan_out(l); in = -1; }
}

__device__ __forceinline__ void publish(int p, float w, bf16* wf, bf16* wt, float* bias) {
    int l, out, in;
    param_index(p, l, out, in);
    if (in < 0) { bias[bp_off(l) + out] = w; return; }
    wf[wf_off(l) + out * k_pad(l) + in] = (bf16)w;
    wt[wf_off(l) + in * 64 + out] = (bf16)w;
}

// sum the block partials; mode 0: Adam step and publish, mode 1: write the gradient
__global__ __launch_bounds__(256) void k_guide_reduce(const float* __restrict__ partials, int nblocks, int mode,
                                                      float* __restrict__ master, float* __restrict__ m,
                                                      float* __restrict__ v, float lr, float b1, float b2, float eps,
                                                      float bc1, float bc2, bf16* wf, bf16* wt, float* bias,
                                                      float* __restrict__ gradOut) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float gsum = 0.0f;
    for (int b = 0; b < nblocks; ++b) gsum += partials[(size_t)b * kParams + p];
    if (mode == 1) { gradOut[p] = gsum; return; }
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    master[p] = w;
    publish(p, w, wf, wt, bias);
}

__global__ __launch_bounds__(256) void k_guide_publish(const float* __restrict__ master, bf16* wf, bf16* wt, float* bias) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams) publish(p, master[p], wf, wt, bias);
}

// inference of staged records (raw outputs), the renderer's wave-resident
// kernel body; FP8: E4M3 weights with per-layer scales
template <bool FP8>
__global__ __launch_bounds__(256, 2) void k_guide_infer(const float* __restrict__ recs, uint32_t n,
                                                        const bf16* __restrict__ wf, const uint8_t* __restrict__ wq,
                                                        const float* __restrict__ wqs, const float* __restrict__ bias,
                                                        float* __restrict__ out) {
    __shared__ WaveLds L[4];
    wave_infer<FP8, false>(L[threadIdx.x >> 6], recs, n, wf, wq, wqs, bias, out);
}

// FP8 copy of the weights: one block per layer finds the largest magnitude,
// then writes E4M3 bytes of W / scale into the [n_pad][k_pad] layout.
__global__ __launch_bounds__(256) void k_fp8_quant(const float* __restrict__ master, uint8_t* __restrict__ wq,
                                                   float* __restrict__ scales) {
    __shared__ float red[4];
    const int l = blockIdx.x;
    const float* W = master + w_off(l);
    const int nw = fan_in(l) * fan_out(l);
    float m = 0.0f;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) m = fmaxf(m, fabsf(W[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (__lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    const float amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    const float sc = amax > 0.0f ? amax / kFp8Max : 1.0f;
    if (threadIdx.x == 0) scales[l] = sc;
    const float inv = 1.0f / sc;
    uint8_t* q = wq + wf_off(l);
    for (int i = threadIdx.x; i < n_pad(l) * k_pad(l); i += blockDim.x) {
        const int out = i / k_pad(l), in = i % k_pad(l);
        const float v = (out < fan_out(l) && in < fan_in(l)) ? W[out * fan_in(l) + in] * inv : 0.0f;
        q[i] = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.0f, 0, false) & 0xFF);
    }
}

int train_blocks(rbh_ctx c) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

template <typename T> T* dptr(void* p) { return static_cast<T*>(p); }

}  // namespace

// ------------------------------------------------------------------ host
void guide_release(GuideState& g) {
    void* ptrs[] = {g.master, g.adamM, g.adamV, g.wf, g.wt, g.bias, g.partials, g.scalars, g.records, g.trainRecs, g.recCount, g.fp8w, g.fp8s};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    g = GuideState{};
}

}  // namespace rbh

using namespace rbh;

namespace {

GuideState& gs(rbh_ctx c) {
    if (!c->guide) c->guide = new GuideState();
    return *static_cast<GuideState*>(c->guide);
}

hipError_t alloc(void** p, size_t bytes) {
    if (*p) return hipSuccess;
    return hipMalloc(p, bytes);
}

// deterministic He-uniform initialisation (SplitMix64 stream)
void init_params(uint32_t seed, std::vector<float>& p) {
    p.assign(kParams, 0.0f);
    uint64_t st = 0x9E3779B97F4A7C15ull * (seed + 1);
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < kLayers; ++l) {
        const double lim = std::sqrt(6.0 / fan_in(l));
        for (int i = 0; i < fan_in(l) * fan_out(l); ++i) p[w_off(l) + i] = (float)((2.0 * next() - 1.0) * lim);
    }
}

// The selection output's weights and bias (output kMix of the last layer):
// zero and never trained when the selection is fixed at 0.5, which makes the
// network the 40-output mixture network.
void clear_selection_row(std::vector<float>& p) {
    const int l = kLayers - 1;
    for (int i = 0; i < fan_in(l); ++i) p[w_off(l) + kMix * fan_in(l) + i] = 0.0f;
    p[b_off(l) + kMix] = 0.0f;
}

int upload_master(rbh_ctx c, const float* params) {
    GuideState& g = gs(c);
    std::vector<float> fixedParams;
    if (g.fixedSel) {
        fixedParams.assign(params, params + kParams);
        clear_selection_row(fixedParams);
        params = fixedParams.data();
    }
    RBH_HIP(c, hipMemcpyAsync(g.master, params, kParams * 4, hipMemcpyHostToDevice, c->stream));
    RBH_HIP(c, hipMemsetAsync(g.adamM, 0, kParams * 4, c->stream));
    RBH_HIP(c, hipMemsetAsync(g.adamV, 0, kParams * 4, c->stream));
    hipLaunchKernelGGL(k_guide_publish, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, dptr<float>(g.master),
                       dptr<bf16>(g.wf), dptr<bf16>(g.wt), dptr<float>(g.bias));
    hipLaunchKernelGGL(k_fp8_quant, dim3(kLayers), dim3(256), 0, c->stream, dptr<float>(g.master), dptr<uint8_t>(g.fp8w),
                       dptr<float>(g.fp8s));
    RBH_HIP(c, hipGetLastError());
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    g.step = 0;
    return RBH_OK;
}

// gradient pass over n device records; mode 0 = Adam step, 1 = gradient to gradOut
int run_train(rbh_ctx c, const float* drecs, uint32_t n, int mode, float* gradOut, float* lossOut) {
    GuideState& g = gs(c);
    const int blocks = (int)std::min<size_t>((size_t)train_blocks(c), (n + kTile - 1) / kTile);
    if (blocks <= 0) return fail(c, RBH_ERR_INVALID, "no training records");
    if ((size_t)blocks > g.partialBlocks) {
        if (g.partials) (void)hipFree(g.partials);
        g.partials = nullptr;
        RBH_HIP(c, hipMalloc(&g.partials, (size_t)blocks * kParams * 4));
        g.partialBlocks = blocks;
    }
    float* sc = dptr<float>(g.scalars);
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, c->stream, sc, 4);
    hipLaunchKernelGGL(k_weight_sum, dim3(256), dim3(256), 0, c->stream, drecs, n, sc);
    hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(256), 0, c->stream, drecs, n, sc, dptr<bf16>(g.wf),
                       dptr<bf16>(g.wt), dptr<float>(g.bias), dptr<float>(g.partials), sc + 1, g.fixedSel ? 1u : 0u);
    if (mode == 0) ++g.step;
    const float bc1 = 1.0f - std::pow(g.beta1, (float)std::max<uint32_t>(1, g.step));
    const float bc2 = 1.0f - std::pow(g.beta2, (float)std::max<uint32_t>(1, g.step));
    hipLaunchKernelGGL(k_guide_reduce, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, dptr<float>(g.partials),
                       blocks, mode, dptr<float>(g.master), dptr<float>(g.adamM), dptr<float>(g.adamV), g.lr, g.beta1,
                       g.beta2, g.eps, bc1, bc2, dptr<bf16>(g.wf), dptr<bf16>(g.wt), dptr<float>(g.bias),
                       gradOut);
    if (mode == 0)
        hipLaunchKernelGGL(k_fp8_quant, dim3(kLayers), dim3(256), 0, c->stream, dptr<float>(g.master),
                           dptr<uint8_t>(g.fp8w), dptr<float>(g.fp8s));
    RBH_HIP(c, hipGetLastError());
    if (lossOut) {
        float tmp[4];
        RBH_HIP(c, hipMemcpyAsync(tmp, sc, 16, hipMemcpyDeviceToHost, c->stream));
        RBH_HIP(c, hipStreamSynchronize(c->stream));
        *lossOut = tmp[1];
    }
    return RBH_OK;
}

}  // namespace

extern "C" {

int rbh_guide_configure(rbh_ctx c, const rbh_guide_config* cfg) {
    if (!c || !cfg) return RBH_ERR_INVALID;
    if (!(cfg->lr > 0.0f) || !(cfg->beta1 >= 0.0f && cfg->beta1 < 1.0f) || !(cfg->beta2 >= 0.0f && cfg->beta2 < 1.0f) ||
        !(cfg->eps > 0.0f))
        return fail(c, RBH_ERR_INVALID, "Adam parameters out of range");
    if (cfg->fixedSelection > 1u) return fail(c, RBH_ERR_INVALID, "fixedSelection must be 0 or 1");
    GuideState& g = gs(c);
    RBH_HIP(c, hipSetDevice(c->device));
    RBH_HIP(c, alloc(&g.master, kParams * 4));
    RBH_HIP(c, alloc(&g.adamM, kParams * 4));
    RBH_HIP(c, alloc(&g.adamV, kParams * 4));
    RBH_HIP(c, alloc(&g.wf, kWfTotal * 2));
    RBH_HIP(c, alloc(&g.wt, kWfTotal * 2));
    RBH_HIP(c, alloc(&g.bias, kBiasPad * 4));
    RBH_HIP(c, alloc(&g.scalars, 16));
    RBH_HIP(c, alloc(&g.fp8w, kWfTotal));
    RBH_HIP(c, alloc(&g.fp8s, 64));
    RBH_HIP(c, hipMemsetAsync(g.fp8w, 0, kWfTotal, c->stream));
    RBH_HIP(c, hipMemsetAsync(g.wf, 0, kWfTotal * 2, c->stream));
    RBH_HIP(c, hipMemsetAsync(g.wt, 0, kWfTotal * 2, c->stream));
    RBH_HIP(c, hipMemsetAsync(g.bias, 0, kBiasPad * 4, c->stream));
    g.recCapacity = cfg->recordCapacity ? cfg->recordCapacity : (1u << 22);
    if (g.records) { (void)hipFree(g.records); g.records = nullptr; }
    if (g.trainRecs) { (void)hipFree(g.trainRecs); g.trainRecs = nullptr; }
    RBH_HIP(c, hipMalloc(&g.records, (size_t)g.recCapacity * 24 * 4));
    RBH_HIP(c, hipMalloc(&g.trainRecs, (size_t)g.recCapacity * kRec * 4));
    RBH_HIP(c, alloc(&g.recCount, 16));
    RBH_HIP(c, hipMemsetAsync(g.recCount, 0, 16, c->stream));
    g.lr = cfg->lr;
    g.beta1 = cfg->beta1;
    g.beta2 = cfg->beta2;
    g.eps = cfg->eps;
    g.fixedSel = cfg->fixedSelection != 0u;
    for (int a = 0; a < 3; ++a) {
        const float e = cfg->sceneMax[a] - cfg->sceneMin[a];
        if (!(e > 0.0f)) return fail(c, RBH_ERR_INVALID, "scene bounds must have positive extent");
        g.sceneMin[a] = cfg->sceneMin[a];
        g.sceneInvExtent[a] = 1.0f / e;
    }
    std::vector<float> p;
    init_params(cfg->seed, p);
    const int rc = upload_master(c, p.data());
    if (rc) return rc;
    g.configured = true;
    return RBH_OK;
}

int rbh_guide_set_params(rbh_ctx c, const float* params, uint32_t n) {
    if (!c || !params) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    if (n != (uint32_t)kParams) return fail(c, RBH_ERR_INVALID, "expected %d parameters, got %u", kParams, n);
    return upload_master(c, params);
}

int rbh_guide_get_params(rbh_ctx c, float* params, uint32_t n) {
    if (!c || !params) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    if (n != (uint32_t)kParams) return fail(c, RBH_ERR_INVALID, "expected %d parameters, got %u", kParams, n);
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    RBH_HIP(c, hipMemcpy(params, gs(c).master, kParams * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

static int stage_records(rbh_ctx c, const float* recs, uint32_t n, DevBuf& buf) {
    RBH_ALLOC(c, buf, (size_t)std::max<uint32_t>(n, 1) * kRec * 4);
    RBH_HIP(c, hipMemcpyAsync(buf.p, recs, (size_t)n * kRec * 4, hipMemcpyHostToDevice, c->stream));
    return RBH_OK;
}

int rbh_guide_gradient(rbh_ctx c, const float* recs, uint32_t n, float* grad, float* loss) {
    if (!c || !recs || !grad || n == 0) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    DevBuf d, gbuf;
    int rc = stage_records(c, recs, n, d);
    if (rc) return rc;
    RBH_ALLOC(c, gbuf, kParams * 4);
    float l = 0.0f;
    rc = run_train(c, d.as<float>(), n, 1, gbuf.as<float>(), &l);
    if (rc) return rc;
    RBH_HIP(c, hipMemcpy(grad, gbuf.p, kParams * 4, hipMemcpyDeviceToHost));
    if (loss) *loss = l;
    return RBH_OK;
}

int rbh_guide_train(rbh_ctx c, const float* recs, uint32_t n, uint32_t batch, uint32_t steps, float* lastLoss) {
    if (!c || !recs || n == 0 || batch == 0) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    DevBuf d;
    int rc = stage_records(c, recs, n, d);
    if (rc) return rc;
    float l = 0.0f;
    for (uint32_t s = 0; s < steps; ++s) {
        // consecutive batches of the staged records, wrapping around
        const uint32_t off = (uint32_t)(((uint64_t)s * batch) % n);
        const uint32_t len = std::min(batch, n - off);
        rc = run_train(c, d.as<float>() + (size_t)off * kRec, len, 0, nullptr, s + 1 == steps ? &l : nullptr);
        if (rc) return rc;
    }
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    if (lastLoss) *lastLoss = l;
    return RBH_OK;
}

static int collected(rbh_ctx c, uint32_t* n) {
    GuideState& g = gs(c);
    uint32_t cnt = 0;
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    RBH_HIP(c, hipMemcpy(&cnt, g.recCount, 4, hipMemcpyDeviceToHost));
    *n = std::min(cnt, g.recCapacity);
    return RBH_OK;
}

int rbh_guide_records(rbh_ctx c, float* out, uint32_t max, uint32_t* n) {
    if (!c || !n) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    uint32_t k = 0;
    int rc = collected(c, &k);
    if (rc) return rc;
    *n = k;
    if (out && max) RBH_HIP(c, hipMemcpy(out, gs(c).trainRecs, (size_t)std::min(k, max) * kRec * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

// Training set of a collecting render, in the pending-record buffer (free
// between renders): the records with a positive weight, in a fixed
// pseudo-random order, weights clamped twice at kClampMean times their mean.
// Clamping only changes what the network is fitted to, never the render's
// estimator; it keeps a few one-sample radiance outliers from dominating
// every batch.
constexpr float kClampMean = 10.0f;

static int prepare_collected(rbh_ctx c, uint32_t n, uint32_t* m) {
    GuideState& g = gs(c);
    const float* recs = static_cast<const float*>(g.trainRecs);
    DevBuf idx, tmp;
    RBH_ALLOC(c, idx, (size_t)n * 4 + 16);
    uint32_t* cnt = idx.as<uint32_t>() + n + (4 - n % 4) % 4;
    RBH_ALLOC(c, tmp, radix::select_scratch_bytes(n));
    RBH_HIP(c, radix::select_indices(tmp.p, n, PositiveWeight{recs}, idx.as<uint32_t>(), cnt, c->stream));
    uint32_t k = 0;
    RBH_HIP(c, hipMemcpyAsync(&k, cnt, 4, hipMemcpyDeviceToHost, c->stream));
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    if (k == 0) return fail(c, RBH_ERR_STATE, "no collected training record has a positive weight");
    static const uint32_t primes[] = {1000003u, 999983u, 104729u, 7919u};
    uint32_t stride = 1u;
    for (uint32_t p : primes)
        if (k % p != 0u) { stride = p % k == 0u ? 1u : p; break; }
    float* out = static_cast<float*>(g.records);
    const int blocks = (int)std::min<uint32_t>(4096u, (k + 255u) / 256u);
    hipLaunchKernelGGL(k_gather_records, dim3(blocks), dim3(256), 0, c->stream, recs, idx.as<uint32_t>(), k, stride, out);
    float* sum = dptr<float>(g.scalars) + 2;
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, c->stream, sum, 1);
        hipLaunchKernelGGL(k_weight_sum, dim3(256), dim3(256), 0, c->stream, out, k, sum);
        hipLaunchKernelGGL(k_clamp_weights, dim3(blocks), dim3(256), 0, c->stream, out, k, sum, kClampMean);
    }
    RBH_HIP(c, hipGetLastError());
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    *m = k;
    return RBH_OK;
}

int rbh_guide_train_collected(rbh_ctx c, uint32_t batch, uint32_t steps, float* lastLoss) {
    if (!c || batch == 0) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    uint32_t n0 = 0, n = 0;
    int rc = collected(c, &n0);
    if (rc) return rc;
    if (n0 == 0) return fail(c, RBH_ERR_STATE, "no training records collected");
    rc = prepare_collected(c, n0, &n);
    if (rc) return rc;
    float l = 0.0f;
    const float* d = static_cast<const float*>(gs(c).records);
    for (uint32_t s = 0; s < steps; ++s) {
        const uint32_t off = (uint32_t)(((uint64_t)s * batch) % n);
        const uint32_t len = std::min(batch, n - off);
        rc = run_train(c, d + (size_t)off * kRec, len, 0, nullptr, s + 1 == steps ? &l : nullptr);
        if (rc) return rc;
    }
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    if (lastLoss) *lastLoss = l;
    return RBH_OK;
}

static int infer(rbh_ctx c, const float* recs, uint32_t n, float* out, bool fp8) {
    if (!c || !recs || !out) return RBH_ERR_INVALID;
    if (!gs(c).configured) return fail(c, RBH_ERR_STATE, "rbh_guide_configure first");
    if (n == 0) return RBH_OK;
    GuideState& g = gs(c);
    DevBuf d, o;
    int rc = stage_records(c, recs, n, d);
    if (rc) return rc;
    RBH_ALLOC(c, o, (size_t)n * kOut * 4);
    // four waves of 16 records per block
    const int blocks = (int)std::min<size_t>((size_t)train_blocks(c) * 2, (n + kTile - 1) / kTile);
    if (fp8)
        hipLaunchKernelGGL((k_guide_infer<true>), dim3(blocks), dim3(256), 0, c->stream, d.as<float>(), n,
                           dptr<bf16>(g.wf), dptr<uint8_t>(g.fp8w), dptr<float>(g.fp8s), dptr<float>(g.bias), o.as<float>());
    else
        hipLaunchKernelGGL((k_guide_infer<false>), dim3(blocks), dim3(256), 0, c->stream, d.as<float>(), n,
                           dptr<bf16>(g.wf), nullptr, dptr<float>(g.fp8s), dptr<float>(g.bias), o.as<float>());
    RBH_HIP(c, hipGetLastError());
    RBH_HIP(c, hipStreamSynchronize(c->stream));
    RBH_HIP(c, hipMemcpy(out, o.p, (size_t)n * kOut * 4, hipMemcpyDeviceToHost));
    return RBH_OK;
}

int rbh_guide_infer(rbh_ctx c, const float* recs, uint32_t n, float* out) { return infer(c, recs, n, out, false); }

int rbh_guide_infer_fp8(rbh_ctx c, const float* recs, uint32_t n, float* out) { return infer(c, recs, n, out, true); }

}  // extern "C"
