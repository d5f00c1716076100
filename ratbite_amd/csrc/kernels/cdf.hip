// Environment-map importance-sampling tables on gfx950.
//
// Replaces buildCDFLatLongRGBUpTo{4,8,16,32}k + invertCDFToIndices
// (src/kernels/cdf.cu:287-535, launched at rendering_interface.cpp:1364-1419)
// with ONE fused kernel: each workgroup owns a row, streams its texels from
// HBM once, builds the row CDF in LDS, inverts it from LDS into uint16
// indices and writes only those (14 B per RGB32F texel of HBM traffic instead
// of the reference's texel read + FP32 CDF write + CDF re-read + index
// write). The last workgroup to finish (device-scope counter) builds the
// marginal CDF with the sin(theta) Jacobian, inverts it and publishes the
// integral, then re-arms the counter so the launch can be replayed without a
// memset.
//
// FP32 mode keeps the reference's association order exactly: per-thread
// float4 component scan, then a block-wide Hillis-Steele scan over the .w
// lanes with the reference's block size B (256/512/1024, chosen from W/4 as
// rendering_interface.cpp:1380-1382), then the running total of previous
// chunks. The CDF floats, and therefore the inverted indices, are
// reproducible bit for bit. FP64 mode carries the prefix sums and the
// normalisation in double and rounds each CDF value once.
#include "../context.hpp"
#include <cmath>

namespace rbh {

namespace {

template <typename T> struct V4 { T x, y, z, w; };
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ inline float texel_power(float r, float g, float b) {
    // src/kernels/cdf.cu:136-144, 150-169: mean of RGB, clamped at zero
    return fmaxf(0.0f, (r + g + b) / 3.0f);
}

__device__ inline V4<float> load4(const float* __restrict__ row, int idx, int channels) {
    V4<float> v;
    if (channels == 3) {
        const f4v* p = reinterpret_cast<const f4v*>(row + (size_t)idx * 12);
        f4v a = __builtin_nontemporal_load(p + 0);
        f4v b = __builtin_nontemporal_load(p + 1);
        f4v c = __builtin_nontemporal_load(p + 2);
        v.x = texel_power(a.x, a.y, a.z);
        v.y = texel_power(a.w, b.x, b.y);
        v.z = texel_power(b.z, b.w, c.x);
        v.w = texel_power(c.y, c.z, c.w);
    } else {
        const f4v* p = reinterpret_cast<const f4v*>(row + (size_t)idx * 16);
        f4v a = __builtin_nontemporal_load(p + 0);
        f4v b = __builtin_nontemporal_load(p + 1);
        f4v c = __builtin_nontemporal_load(p + 2);
        f4v d = __builtin_nontemporal_load(p + 3);
        v.x = texel_power(a.x, a.y, a.z);
        v.y = texel_power(b.x, b.y, b.z);
        v.z = texel_power(c.x, c.y, c.z);
        v.w = texel_power(d.x, d.y, d.z);
    }
    return v;
}

// sin of the reference's float argument pi*(i + k/4 + 1/8)/(H/4)
// (cdf.cu:112-132), as the correctly rounded float of a double sine.
__device__ inline float row_weight_f32(int index, int k, int numElements) {
    const float off = 0.125f + 0.25f * (float)k;
    float arg = 3.14159265358979323846f * ((float)index + off) / (float)numElements;
    return (float)sin((double)arg);
}
__device__ inline double row_weight_f64(int row, int h) {
    return sin(3.14159265358979323846 * ((double)row + 0.5) / (double)h);
}

// Block-wide inclusive Hillis-Steele scan over .w lanes with the same
// step/association as cdf.cu:264-282. Double-buffered LDS: one barrier per
// step is enough because step k+2 reuses a buffer only after every lane has
// passed the barrier of step k+1.
template <int B, typename T>
__device__ inline void block_scan(V4<T>& v, T* __restrict__ sbuf, T& total, T* __restrict__ stotal) {
    const int t = threadIdx.x;
    v.y += v.x; v.z += v.y; v.w += v.z;
    int k = 0;
#pragma unroll
    for (int i = 1; i < B; i <<= 1, ++k) {
        T* buf = sbuf + (k & 1) * B;
        buf[t] = v.w;
        __syncthreads();
        if (t >= i) {
            T o = buf[t - i];
            v.x += o; v.y += o; v.z += o; v.w += o;
        }
    }
    v.x += total; v.y += total; v.z += total; v.w += total;
    if (t == B - 1) *stotal = v.w;
    __syncthreads();
    total = *stotal;
}

// Binary search of cdf.cu:517-534 over an LDS-resident CDF.
__device__ inline uint16_t invert_one(const float* __restrict__ cdf, int n, int j) {
    int low = 0, high = n - 1;
    float p = (float)j / (float)(n - 1);
    int mid = (high - low) / 2;
    do {
        if (p > cdf[mid]) low = mid; else high = mid;
        mid = low + (high - low) / 2;
    } while (low + 1 != high);
    return (uint16_t)high;
}

template <int B, typename T>
__global__ __launch_bounds__(B) void k_envmap_cdf(const float* __restrict__ texels, int w, int h, int channels,
                                                   uint16_t* __restrict__ cond, uint16_t* __restrict__ marg,
                                                   float* __restrict__ cdfY, T* __restrict__ rowTotals,
                                                   T* __restrict__ integral, unsigned int* __restrict__ counter) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sbuf = reinterpret_cast<T*>(smem_raw);                    // 2*B scan buffers
    T* stotal = sbuf + 2 * B;                                    // broadcast slot
    const int maxwh = w > h ? w : h;
    T* raw = stotal + 2;                                         // max(w, h) un-normalised prefixes
    float* cdf = sizeof(T) == 4 ? reinterpret_cast<float*>(raw) : reinterpret_cast<float*>(raw + maxwh);
    __shared__ bool done;

    const int t = threadIdx.x;
    const int row = blockIdx.x;
    const int numEl = w >> 2;
    const int nChunks = (numEl + B - 1) / B;
    const float* rowp = texels + (size_t)row * w * channels;

    // ---- row CDF into LDS
    T total = (T)0;
    for (int c = 0; c < nChunks; ++c) {
        const int idx = t + B * c;
        V4<T> v{(T)0, (T)0, (T)0, (T)0};
        if (idx < numEl) {
            V4<float> f = load4(rowp, idx, channels);
            v = V4<T>{(T)f.x, (T)f.y, (T)f.z, (T)f.w};
        }
        block_scan<B, T>(v, sbuf, total, stotal);
        if (idx < numEl) {
            raw[4 * idx + 0] = v.x; raw[4 * idx + 1] = v.y; raw[4 * idx + 2] = v.z; raw[4 * idx + 3] = v.w;
        }
    }
    {
        // each lane normalises the entries it wrote (cdf.cu:380-392)
        const T norm = total != (T)0 ? (T)1 / total : (T)0;
        for (int c = 0; c < nChunks; ++c) {
            const int idx = t + B * c;
            if (idx < numEl)
                for (int k = 0; k < 4; ++k) cdf[4 * idx + k] = (float)(raw[4 * idx + k] * norm);
        }
    }
    if (t == 0) rowTotals[row] = total;
    __syncthreads();

    // ---- invert the row from LDS: two indices per thread, one 32-bit store
    uint32_t* cond32 = reinterpret_cast<uint32_t*>(cond + (size_t)row * w);
    for (int j2 = t; j2 < (w >> 1); j2 += B) {
        uint32_t a = invert_one(cdf, w, 2 * j2);
        uint32_t b = invert_one(cdf, w, 2 * j2 + 1);
        cond32[j2] = a | (b << 16);
    }

    // ---- last workgroup builds the marginal
    __threadfence();
    if (t == 0) done = atomicAdd(counter, 1u) == (unsigned)(h - 1);
    __syncthreads();
    if (!done) return;
    __threadfence();

    const int numElY = h >> 2;
    const int nBlocksY = (numElY + B - 1) / B;
    T mtotal = (T)0;
    for (int c = 0; c < nBlocksY; ++c) {
        const int idx = t + B * c;
        V4<T> v{(T)0, (T)0, (T)0, (T)0};
        if (idx < numElY) {
            const T* rt = rowTotals + (size_t)idx * 4;
            if constexpr (sizeof(T) == 4) {
                v = V4<T>{rt[0] * row_weight_f32(idx, 0, numElY), rt[1] * row_weight_f32(idx, 1, numElY),
                          rt[2] * row_weight_f32(idx, 2, numElY), rt[3] * row_weight_f32(idx, 3, numElY)};
            } else {
                v = V4<T>{rt[0] * row_weight_f64(4 * idx + 0, h), rt[1] * row_weight_f64(4 * idx + 1, h),
                          rt[2] * row_weight_f64(4 * idx + 2, h), rt[3] * row_weight_f64(4 * idx + 3, h)};
            }
        }
        block_scan<B, T>(v, sbuf, mtotal, stotal);
        if (idx < numElY) {
            T* rt = rowTotals + (size_t)idx * 4;  // reuse as the un-normalised marginal prefix
            rt[0] = v.x; rt[1] = v.y; rt[2] = v.z; rt[3] = v.w;
        }
    }
    __syncthreads();
    if constexpr (sizeof(T) == 4) {
        const float norm = mtotal != 0.0f ? 1.0f / mtotal : 0.0f;
        for (int c = 0; c < nBlocksY; ++c) {
            const int idx = t + B * c;
            if (idx < numElY)
                for (int k = 0; k < 4; ++k) cdf[4 * idx + k] = rowTotals[4 * idx + k] * norm;
        }
    } else {
        const double norm = mtotal != 0.0 ? 1.0 / mtotal : 0.0;
        for (int c = 0; c < nBlocksY; ++c) {
            const int idx = t + B * c;
            if (idx < numElY)
                for (int k = 0; k < 4; ++k) cdf[4 * idx + k] = (float)(rowTotals[4 * idx + k] * norm);
        }
    }
    __syncthreads();
    for (int i = t; i < h; i += B) cdfY[i] = cdf[i];
    for (int j = t; j < h; j += B) marg[j] = invert_one(cdf, h, j);
    if (t == 0) {
        *integral = mtotal;
        *counter = 0u;  // re-arm for the next build
    }
}

int block_size_for(int w) {
    int numElements = (w + 3) / 4;
    return numElements <= 256 ? 256 : (numElements <= 512 ? 512 : 1024);
}

template <int B, typename T>
hipError_t launch_cdf(rbh_ctx c, int w, int h, int ch, void* rowTotals, void* integral) {
    const size_t maxwh = (size_t)(w > h ? w : h);
    size_t lds = (size_t)(2 * B + 2) * sizeof(T) + maxwh * sizeof(T) + (sizeof(T) == 4 ? 0 : maxwh * sizeof(float));
    auto fn = k_envmap_cdf<B, T>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(h), dim3(B), lds, c->stream, c->env.texels.as<float>(), w, h, ch,
                       c->env.cond.as<uint16_t>(), c->env.marg.as<uint16_t>(), c->env.cdfY.as<float>(),
                       static_cast<T*>(rowTotals), static_cast<T*>(integral), c->env.counter.as<unsigned int>());
    return hipGetLastError();
}

}  // namespace

// Launches the fused build on the resident env map; returns a HIP error.
hipError_t envmap_cdf_launch(rbh_ctx c, bool fp64) {
    const int w = c->env.width, h = c->env.height, ch = c->env.channels;
    const int B = block_size_for(w);
    // rowTotals (h values of T) + integral (one T) live in integralDev
    char* base = c->env.integralDev.as<char>();
    void* integral = base;
    void* rowTotals = base + 16;
    if (!fp64) {
        if (B == 256) return launch_cdf<256, float>(c, w, h, ch, rowTotals, integral);
        if (B == 512) return launch_cdf<512, float>(c, w, h, ch, rowTotals, integral);
        return launch_cdf<1024, float>(c, w, h, ch, rowTotals, integral);
    }
    if (B == 256) return launch_cdf<256, double>(c, w, h, ch, rowTotals, integral);
    if (B == 512) return launch_cdf<512, double>(c, w, h, ch, rowTotals, integral);
    return launch_cdf<1024, double>(c, w, h, ch, rowTotals, integral);
}

size_t envmap_scratch_bytes(int h) { return 16 + (size_t)h * sizeof(double); }

}  // namespace rbh
