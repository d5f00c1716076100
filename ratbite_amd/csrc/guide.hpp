// Path-guiding network: sizes, parameter layout and device state.
//
// An MLP 80 -> 64 -> 64 -> 64 -> 64 -> 41 (ReLU on the hidden layers) maps a
// shading point (12-frequency sin/cos encoding of the scene-normalised
// position, shading normal, outgoing direction, roughness, normalised
// wavelength) to 8 von Mises-Fisher lobes (mean direction, log sharpness,
// mixture logit) and a selection logit c: a renderer draws from the lobes
// with probability kSelMin + (kSelMax - kSelMin) sigmoid(c), otherwise from
// the BSDF. It has no counterpart in the reference; it extends the
// BSDF sampling of src/kernels/pathtrace.cu:548-905 through one-sample MIS.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rbh {
namespace guide {

constexpr int kIn = 80, kWidth = 64, kLobes = 8, kMix = 5 * kLobes, kOut = kMix + 1, kRec = 16, kLayers = 5, kFreq = 12;
constexpr float kLogKappaMin = -3.0f, kLogKappaMax = 7.0f;
constexpr float kSelMin = 0.1f, kSelMax = 0.9f;   // range of the learned selection probability

// per layer: true fan-in / fan-out and the MFMA-padded sizes
__host__ __device__ constexpr int fan_in(int l) { return l == 0 ? kIn : kWidth; }
__host__ __device__ constexpr int fan_out(int l) { return l == kLayers - 1 ? kOut : kWidth; }
__host__ __device__ constexpr int k_pad(int l) { return l == 0 ? 96 : 64; }     // multiple of 32
__host__ __device__ constexpr int n_pad(int) { return 64; }                      // multiple of 32

// FP32 master parameters, flattened per layer as W[out][in] then b[out]
__host__ __device__ constexpr int w_off(int l) {
    return l == 0 ? 0 : w_off(l - 1) + fan_in(l - 1) * fan_out(l - 1) + fan_out(l - 1);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + fan_in(l) * fan_out(l); }
constexpr int kParams = b_off(kLayers - 1) + kOut;   // 20329
static_assert(kParams == 20329, "parameter count");

// BF16 device copies: forward [n_pad][k_pad] and transposed [k_pad][n_pad]
__host__ __device__ constexpr int wf_off(int l) { return l == 0 ? 0 : wf_off(l - 1) + n_pad(l - 1) * k_pad(l - 1); }
constexpr int kWfTotal = wf_off(kLayers);
// padded FP32 biases, n_pad per layer
__host__ __device__ constexpr int bp_off(int l) { return l * 64; }
constexpr int kBiasPad = kLayers * 64;

}  // namespace guide

// Device-side state of the guiding network, owned by the context.
struct GuideState {
    bool configured = false;
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float sceneMin[3] = {0, 0, 0}, sceneInvExtent[3] = {1, 1, 1};
    uint32_t step = 0;                 // Adam steps taken
    bool fixedSel = false;             // selection fixed at 0.5 (the selection output held at zero)
    void* master = nullptr;            // float[kParams]
    void* adamM = nullptr;             // float[kParams]
    void* adamV = nullptr;             // float[kParams]
    void* wf = nullptr;                // bf16 forward weights (kWfTotal)
    void* wt = nullptr;                // bf16 transposed weights (kWfTotal)
    void* bias = nullptr;              // float[kBiasPad]
    void* partials = nullptr;          // float[blocks][kParams + kBiasPad]
    size_t partialBlocks = 0;
    void* scalars = nullptr;           // float[4]: weight sum, loss sum
    void* records = nullptr;           // pending records of a collecting render (float[cap][24])
    void* trainRecs = nullptr;         // finished training records (float[cap][kRec])
    void* recCount = nullptr;          // uint32 counter
    uint32_t recCapacity = 0;
    void* fp8w = nullptr;              // FP8 E4M3 forward weights (kWfTotal bytes, forward layout)
    void* fp8s = nullptr;              // float[kLayers] weight scales
};

void guide_release(GuideState& g);

}  // namespace rbh
