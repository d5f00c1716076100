// TEST INFRASTRUCTURE — CPU oracle (see rb_math.hpp header).
//
// Environment-map 2D CDF build and inversion, restating
// src/kernels/cdf.cu:61-535 and the host launch in
// src/core/rendering_interface.cpp:1364-1419.
//
// The FP32 row/marginal prefix sums follow the kernel's association order
// exactly (per-thread float4 component scan, then a Hillis-Steele scan over
// the .w lanes of a block of B threads, then the running total of earlier
// chunks), so the resulting CDF values — and hence the uint16 inversion
// indices — are reproducible bit for bit. B is 256/512/1024 as chosen at
// rendering_interface.cpp:1380-1382.
//
// Row weights sin(pi*(i+k/4+1/8)/(H/4)) (cdf.cu:112-132) are evaluated as the
// correctly rounded float of a double-precision sine of the same float
// argument.
#include "rb_math.hpp"
#include <vector>

namespace orc {

struct F4 { float x, y, z, w; };

static inline float texel_power(const float* t, int ch) {
    // src/kernels/cdf.cu:136-165: average of r,g,b, clamped at 0
    return std::max(0.0f, (t[0] + t[1] + t[2]) / 3.0f);
}

int cdf_block_size(int w) {
    int numElements = (w + 3) / 4;
    return numElements <= 256 ? 256 : (numElements <= 512 ? 512 : 1024);
}

float cdf_row_weight(int index, int k, int numElements) {
    static const float off[4] = {0.125f, 0.375f, 0.625f, 0.875f};
    float arg = PI * ((float)index + off[k]) / (float)numElements;
    return (float)std::sin((double)arg);
}

// One block-wide inclusive scan (src/kernels/cdf.cu:256-300), applied to B
// lanes of float4, carrying `total` between calls.
static void prefix_sum_block(std::vector<F4>& val, int B, float& total, std::vector<float>& smem) {
    for (int t = 0; t < B; ++t) {
        val[t].y += val[t].x;
        val[t].z += val[t].y;
        val[t].w += val[t].z;
    }
    for (int i = 1; i < B; i <<= 1) {
        for (int t = 0; t < B; ++t) smem[t] = val[t].w;
        for (int t = i; t < B; ++t) {
            float o = smem[t - i];
            val[t].x += o; val[t].y += o; val[t].z += o; val[t].w += o;
        }
    }
    for (int t = 0; t < B; ++t) {
        val[t].x += total; val[t].y += total; val[t].z += total; val[t].w += total;
    }
    total = val[B - 1].w;
}

// Returns 0 on success. cdf_x: w*h floats, cdf_y: h floats, integral: the
// kernel's raw total (the host divides by w*h afterwards).
int cdf_build_f32(const float* texels, int w, int h, int channels, float* cdf_x, float* cdf_y, float* integral) {
    if (w % 4 || h % 4 || w <= 0 || h <= 0) return -1;
    const int numElements = w / 4;
    const int B = cdf_block_size(w);
    const int nChunks = (numElements + B - 1) / B;
    if (nChunks > 8) return -2;
    std::vector<F4> val(B);
    std::vector<float> smem(B);
    std::vector<F4> rowv((size_t)nChunks * B);
    for (int y = 0; y < h; ++y) {
        float total = 0.0f;
        for (int c = 0; c < nChunks; ++c) {
            for (int t = 0; t < B; ++t) {
                int idx = t + B * c;
                if (idx < numElements) {
                    const float* p = texels + ((size_t)y * w + (size_t)idx * 4) * channels;
                    val[t] = {texel_power(p, channels), texel_power(p + channels, channels),
                              texel_power(p + 2 * channels, channels), texel_power(p + 3 * channels, channels)};
                } else {
                    val[t] = {0, 0, 0, 0};
                }
            }
            prefix_sum_block(val, B, total, smem);
            for (int t = 0; t < B; ++t) rowv[(size_t)c * B + t] = val[t];
        }
        float norm = total != 0.0f ? 1.0f / total : 0.0f;
        for (int c = 0; c < nChunks; ++c)
            for (int t = 0; t < B; ++t) {
                int idx = t + B * c;
                if (idx >= numElements) continue;
                const F4& v = rowv[(size_t)c * B + t];
                float* o = cdf_x + (size_t)y * w + (size_t)idx * 4;
                o[0] = v.x * norm; o[1] = v.y * norm; o[2] = v.z * norm; o[3] = v.w * norm;
            }
        cdf_y[y] = total;
    }
    // Marginal pass (src/kernels/cdf.cu:418-463), same block size B.
    const int numElY = h / 4;
    const int nBlocksY = (numElY + B - 1) / B;
    float total = 0.0f;
    for (int c = 0; c < nBlocksY; ++c) {
        for (int t = 0; t < B; ++t) {
            int idx = t + B * c;
            if (idx < numElY) {
                float* p = cdf_y + (size_t)idx * 4;
                val[t] = {p[0] * cdf_row_weight(idx, 0, numElY), p[1] * cdf_row_weight(idx, 1, numElY),
                          p[2] * cdf_row_weight(idx, 2, numElY), p[3] * cdf_row_weight(idx, 3, numElY)};
            } else {
                val[t] = {0, 0, 0, 0};
            }
        }
        prefix_sum_block(val, B, total, smem);
        for (int t = 0; t < B; ++t) {
            int idx = t + B * c;
            if (idx < numElY) {
                float* p = cdf_y + (size_t)idx * 4;
                p[0] = val[t].x; p[1] = val[t].y; p[2] = val[t].z; p[3] = val[t].w;
            }
        }
    }
    float norm = total != 0.0f ? 1.0f / total : 0.0f;
    for (int i = 0; i < h; ++i) cdf_y[i] *= norm;
    *integral = total;
    return 0;
}

// FP64 reference of the same distribution (sequential sums in double, the
// row weight sin(pi (y + 1/2) / H) of cdf.cu:112-132 evaluated in double),
// used to pin the FP64 mode and to state the FP32 build's error. cdf values
// are returned as double.
int cdf_build_f64(const float* texels, int w, int h, int channels, double* cdf_x, double* cdf_y, double* integral) {
    if (w % 4 || h % 4) return -1;
    for (int y = 0; y < h; ++y) {
        double s = 0.0;
        for (int x = 0; x < w; ++x) {
            s += (double)texel_power(texels + ((size_t)y * w + x) * channels, channels);
            cdf_x[(size_t)y * w + x] = s;
        }
        double n = s != 0.0 ? 1.0 / s : 0.0;
        for (int x = 0; x < w; ++x) cdf_x[(size_t)y * w + x] *= n;
        cdf_y[y] = s;
    }
    double s = 0.0;
    for (int y = 0; y < h; ++y) {
        s += cdf_y[y] * std::sin(3.14159265358979323846 * ((double)y + 0.5) / (double)h);
        cdf_y[y] = s;
    }
    double n = s != 0.0 ? 1.0 / s : 0.0;
    for (int y = 0; y < h; ++y) cdf_y[y] *= n;
    *integral = s;
    return 0;
}

// src/kernels/cdf.cu:493-535
static uint16_t invert_one(const float* cdf, int n, int j) {
    int low = 0, high = n - 1;
    float p = (float)j / (float)(n - 1);
    int mid = (high - low) / 2;
    do {
        if (p > cdf[mid]) low = mid; else high = mid;
        mid = low + (high - low) / 2;
    } while (low + 1 != high);
    return (uint16_t)high;
}

void cdf_invert(const float* cdf_x, const float* cdf_y, int w, int h, uint16_t* cond, uint16_t* marg) {
    for (int y = 0; y < h; ++y)
        for (int j = 0; j < w; ++j) cond[(size_t)y * w + j] = invert_one(cdf_x + (size_t)y * w, w, j);
    for (int j = 0; j < h; ++j) marg[j] = invert_one(cdf_y, h, j);
}

}  // namespace orc
