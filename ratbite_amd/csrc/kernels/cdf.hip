// Environment-map importance-sampling tables on gfx950.
//
// Replaces buildCDFLatLongRGBUpTo{4,8,16,32}k + invertCDFToIndices
// (src/kernels/cdf.cu:287-535, launched at rendering_interface.cpp:1364-1419)
// with ONE fused kernel. Each wave owns a row (up to four rows per
// workgroup, as many as fit 80 KB of LDS so that two workgroups share a CU):
// it streams the row's texels from HBM once, scans them in registers, keeps
// the normalised row CDF in its own LDS row, inverts it from LDS into uint16
// indices and writes only those (14 B per RGB32F texel of HBM traffic instead
// of the reference's texel read + FP32 CDF write + CDF re-read + index
// write). The last workgroup to finish (device-scope counter) builds the
// marginal CDF with the sin(theta) Jacobian the same way with its first wave,
// inverts it, publishes the integral and re-arms the counter, so the launch
// can be replayed without a memset.
//
// FP32 mode keeps the reference's association order exactly: per-thread
// float4 component scan, then the block-wide Hillis-Steele scan over the .w
// lanes with the reference's block size B (256/512/1024, chosen from W/4 as
// rendering_interface.cpp:1380-1382), then the running total of previous
// chunks. Here one wave runs all B virtual lanes with cross-lane reads and no
// barrier (wave_scan). The CDF floats, and therefore the inverted indices,
// are reproducible bit for bit. The inversion is linear in the row length
// for rows whose probed entries are non-decreasing (the usual case) and falls
// back to the reference's binary search otherwise. FP64 mode carries the
// prefix sums and the normalisation in double and rounds each CDF value once.
#include "../context.hpp"
#include <cmath>
#include <mutex>

namespace rbh {

namespace {

template <typename T> struct V4 { T x, y, z, w; };

// LDS floats per wave: the row (or marginal) CDF of max(w, h) floats, then
// as many 16-bit marks for the linear inversion (max(w, h) is a multiple of
// 4). Rows wider than 16384 leave the marks out (they would not fit 160 KB)
// and are inverted by the binary search.
__host__ __device__ inline bool use_marks(int maxwh) { return maxwh <= 16384; }
__host__ __device__ inline size_t wave_lds_floats(int maxwh) {
    return (size_t)maxwh + (use_marks(maxwh) ? (size_t)maxwh / 2 : 0);
}
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

// The reference's block-wide inclusive Hillis-Steele scan (cdf.cu:264-282:
// per-thread component prefix, then steps 1, 2, 4, ..., B/2 that add the
// .w value of lane t - i from the previous step to all four components of
// lane t >= i) run by ONE wave with no LDS and no barrier: virtual lane
// t = 64 k + lane of the B = 64 V lanes lives in register k of hardware lane
// `lane`. A step i < 64 reads lane (lane - i) mod 64 of register k (t >= i
// within the same register) or of register k - 1 (the source wrapped into
// the previous 64), one cross-lane read per register; a step i = 64 m reads
// register k - m of the same lane. Every virtual lane adds the same operands
// in the same order as the reference, so the sums are bit-identical.
template <int V, typename T>
__device__ __forceinline__ void wave_scan(T (&x)[V], T (&y)[V], T (&z)[V], T (&w)[V], T& total) {
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        y[k] += x[k];
        z[k] += y[k];
        w[k] += z[k];
    }
#pragma unroll
    for (int i = 1; i < 64; i <<= 1) {
        T s[V];
        const int src = (lane - i) & 63;
#pragma unroll
        for (int k = 0; k < V; ++k) s[k] = __shfl(w[k], src);
        const bool own = lane >= i;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (own || k > 0) {
                const T o = own ? s[k] : s[k > 0 ? k - 1 : 0];
                x[k] += o; y[k] += o; z[k] += o; w[k] += o;
            }
        }
    }
#pragma unroll
    for (int m = 1; m < V; m <<= 1) {
        // descending k: register k - m still holds the previous step's value
#pragma unroll
        for (int k = V - 1; k >= m; --k) {
            const T o = w[k - m];
            x[k] += o; y[k] += o; z[k] += o; w[k] += o;
        }
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
        x[k] += total; y[k] += total; z[k] += total; w[k] += total;
    }
    total = __shfl(w[V - 1], 63);
}

// Loads chunk c of a row's power prefixes: virtual lane t covers texels
// 4t .. 4t+3 (zero past the row).
template <int V, typename T>
__device__ __forceinline__ void load_chunk(const float* __restrict__ rowp, int numEl, int c, int channels, T (&x)[V],
                                           T (&y)[V], T (&z)[V], T (&w)[V]) {
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = c * 64 * V + 64 * k + lane;
        x[k] = y[k] = z[k] = w[k] = (T)0;
        if (idx < numEl) {
            const V4<float> f = load4(rowp, idx, channels);
            x[k] = (T)f.x; y[k] = (T)f.y; z[k] = (T)f.z; w[k] = (T)f.w;
        }
    }
}

// Chunk c of the weighted row totals (the marginal pass's input).
template <int V, typename T>
__device__ __forceinline__ void load_marginal(const T* __restrict__ rowTerms, int numElY, int c, T (&x)[V], T (&y)[V],
                                              T (&z)[V], T (&w)[V]) {
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = c * 64 * V + 64 * k + lane;
        x[k] = y[k] = z[k] = w[k] = (T)0;
        if (idx < numElY) {
            const T* rt = rowTerms + (size_t)idx * 4;
            x[k] = rt[0]; y[k] = rt[1]; z[k] = rt[2]; w[k] = rt[3];
        }
    }
}

template <int V, typename T>
__device__ __forceinline__ void store_chunk(T* __restrict__ dst, int numEl, int c, const T (&x)[V], const T (&y)[V],
                                            const T (&z)[V], const T (&w)[V]) {
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = c * 64 * V + 64 * k + lane;
        if (idx < numEl) {
            dst[4 * idx + 0] = x[k]; dst[4 * idx + 1] = y[k]; dst[4 * idx + 2] = z[k]; dst[4 * idx + 3] = w[k];
        }
    }
}

template <int V, typename T>
__device__ __forceinline__ void store_chunk_lds(float* __restrict__ cdf, int numEl, int c, const T (&x)[V],
                                                const T (&y)[V], const T (&z)[V], const T (&w)[V]) {
    const int lane = __lane_id();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int idx = c * 64 * V + 64 * k + lane;
        if (idx < numEl) {
            cdf[4 * idx + 0] = (float)x[k]; cdf[4 * idx + 1] = (float)y[k];
            cdf[4 * idx + 2] = (float)z[k]; cdf[4 * idx + 3] = (float)w[k];
        }
    }
}

// The binary search of cdf.cu:517-534 (invert_one) for the n outputs of an
// LDS-resident CDF, eight searches per lane interleaved so that their LDS
// reads overlap; outputs j, j+1 are stored as one 32-bit word.
__device__ __forceinline__ void invert_row_search(const float* __restrict__ cdf, int n, uint16_t* __restrict__ out) {
    constexpr int K = 8;   // 4 pairs
    const int lane = __lane_id();
    const int pairs = n >> 1;
    uint32_t* out32 = reinterpret_cast<uint32_t*>(out);
    const float inv = (float)(n - 1);
    for (int base = 0; base < pairs; base += 64 * (K / 2)) {
        int low[K], high[K], mid[K];
        float p[K];
        bool act[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int pr = base + 64 * (k >> 1) + lane;
            const int j = 2 * pr + (k & 1);
            act[k] = pr < pairs;
            p[k] = (float)j / inv;
            low[k] = 0;
            high[k] = n - 1;
            mid[k] = (high[k] - low[k]) / 2;
        }
        bool any = true;
        while (any) {
            float c[K];
#pragma unroll
            for (int k = 0; k < K; ++k) c[k] = cdf[mid[k]];
            any = false;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const bool gt = p[k] > c[k];
                const int nl = gt ? mid[k] : low[k], nh = gt ? high[k] : mid[k];
                low[k] = act[k] ? nl : low[k];
                high[k] = act[k] ? nh : high[k];
                mid[k] = low[k] + (high[k] - low[k]) / 2;
                act[k] = act[k] && low[k] + 1 != high[k];
                any = any || act[k];
            }
            any = __ballot(any) != 0ull;
        }
#pragma unroll
        for (int k = 0; k < K; k += 2) {
            const int pr = base + 64 * (k >> 1) + lane;
            if (pr < pairs) out32[pr] = (uint32_t)high[k] | ((uint32_t)high[k + 1] << 16);
        }
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    // this wave's LDS writes are visible to all its lanes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The first output j in [0, n-1] whose query q_j = RN(j / (n-1)) (a
// correctly rounded float quotient, as invert_one computes it) exceeds c >= 0,
// or n, without dividing: RN(x) > c exactly when x lies above the midpoint
// m of c and the next float, or on it when ties-to-even rounds m up (c's
// significand odd). m (n-1) is exact in double (25 + 15 significant bits).
__device__ __forceinline__ int first_above(float c, int n) {
    const uint32_t b = __float_as_uint(c);
    if (b >= 0x7f7fffffu) return n;                      // the largest float, inf: no query exceeds them
    const double t = 0.5 * ((double)c + (double)__uint_as_float(b + 1u)) * (double)(n - 1);
    const double f = floor(t);
    const double j = (f == t && (b & 1u)) ? f : f + 1.0;
    return j >= (double)n ? n : (int)j;
}

// The outputs of invert_one for all j = 0..n-1 in O(n). The search only
// probes entries 1..n-2, and while those are non-decreasing it returns
// 1 + #{m in [1, n-2] : cdf[m] < p_j} (at least 1, at most n-1). Entry m is
// below the queries from first_above(cdf[m]) on, so output j is the largest
// m + 1 whose first_above(cdf[m]) <= j: each m marks its first query (only
// the last of several entries sharing one first query writes it), and a
// running maximum over the marks gives every output. A row whose probed
// entries decrease somewhere (float rounding of the reference's prefix sums
// can make a later entry an ulp smaller) or hold a NaN takes the exact
// binary search instead, as do all rows when `marks` (n 16-bit words of
// this wave) is null.
__device__ __forceinline__ void invert_row(const float* __restrict__ cdf, uint16_t* __restrict__ marks, int n,
                                           uint16_t* __restrict__ out) {
    const int lane = __lane_id();
    if (marks == nullptr) {
        invert_row_search(cdf, n, out);
        return;
    }
    uint2* z64 = reinterpret_cast<uint2*>(marks);
    for (int q = lane; 4 * q < n; q += 64) z64[q] = make_uint2(0u, 0u);
    wave_sync_lds();
    // one pass over the probed entries: the order check and the marks (a
    // row that fails the check is redone by the binary search)
    bool bad = false;
    for (int m = 1 + lane; m <= n - 2; m += 64) {
        const float c = cdf[m];
        const bool last = m + 1 > n - 2;
        const float cn = last ? 0.0f : cdf[m + 1];
        bad = bad || !(c >= 0.0f) || (!last && !(cn >= c));
        const int jm = first_above(c, n);
        const int jn = last ? n : first_above(cn, n);
        if (jm < jn && jm < n) marks[jm] = (uint16_t)(m + 1);
    }
    if (__ballot(bad) != 0ull) {
        invert_row_search(cdf, n, out);
        return;
    }
    wave_sync_lds();
    // running maximum from 1, four consecutive outputs per lane (n is a
    // multiple of 4): in-lane maximum, then a wave-wide exclusive maximum
    uint32_t carry = 1u;
    const uint2* m64 = reinterpret_cast<const uint2*>(marks);
    uint2* o64 = reinterpret_cast<uint2*>(out);
    for (int base = 0; base < n; base += 256) {
        const int q = (base >> 2) + lane;          // group of four outputs
        const bool live = 4 * q < n;
        uint32_t v0 = 0u, v1 = 0u, v2 = 0u, v3 = 0u;
        if (live) {
            const uint2 g = m64[q];
            v0 = g.x & 0xFFFFu; v1 = g.x >> 16; v2 = g.y & 0xFFFFu; v3 = g.y >> 16;
        }
        v1 = max(v1, v0); v2 = max(v2, v1); v3 = max(v3, v2);
        uint32_t sc = v3;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(sc, o);
            if (lane >= o) sc = max(sc, t);
        }
        uint32_t ex = __shfl_up(sc, 1);
        if (lane == 0) ex = 0u;
        ex = max(ex, carry);
        if (live)
            o64[q] = make_uint2(max(ex, v0) | (max(ex, v1) << 16), max(ex, v2) | (max(ex, v3) << 16));
        carry = max(carry, __shfl(sc, 63));
    }
}

// One wave per row: the row's power prefixes in registers (chunks of the
// reference's B = 64 V lanes, running total between chunks), the normalised
// CDF in the wave's LDS row, the row's inverted indices from LDS. The last
// workgroup to finish (device-scope counter) builds the marginal with its
// first wave the same way, publishes the integral and re-arms the counter.
// FP32 keeps the un-normalised prefixes in LDS and normalises them in place;
// FP64 scans the row twice (the second time from L2) and rounds each
// normalised value to float once.
template <int V, typename T, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_envmap_cdf(const float* __restrict__ texels, int w, int h, int channels,
                                                            uint16_t* __restrict__ cond, uint16_t* __restrict__ marg,
                                                            float* __restrict__ cdfY, T* __restrict__ rowTotals,
                                                            T* __restrict__ integral, unsigned int* __restrict__ counter,
                                                            double* __restrict__ normIntegral) {
    extern __shared__ __align__(16) float lds[];
    __shared__ bool last;
    constexpr int B = 64 * V;
    constexpr bool kF32 = sizeof(T) == 4;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int maxwh = w > h ? w : h;
    float* cdf = lds + (size_t)wave * wave_lds_floats(maxwh);
    uint16_t* marks = use_marks(maxwh) ? reinterpret_cast<uint16_t*>(cdf + maxwh) : nullptr;
    const int row = blockIdx.x * WAVES + wave;     // h is a multiple of WAVES
    const int numEl = w >> 2;
    const int nChunks = (numEl + B - 1) / B;
    const float* rowp = texels + (size_t)row * w * channels;

    T total = (T)0;
    for (int c = 0; c < nChunks; ++c) {
        T x[V], y[V], z[V], q[V];
        load_chunk<V, T>(rowp, numEl, c, channels, x, y, z, q);
        wave_scan<V, T>(x, y, z, q, total);
        if constexpr (kF32) store_chunk_lds<V, T>(cdf, numEl, c, x, y, z, q);
    }
    const T norm = total != (T)0 ? (T)1 / total : (T)0;
    if constexpr (kF32) {
        // cdf.cu:380-392, in place (each lane rescales what it wrote)
        for (int c = 0; c < nChunks; ++c)
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int idx = c * B + 64 * k + lane;
                if (idx < numEl)
                    for (int e = 0; e < 4; ++e) cdf[4 * idx + e] = cdf[4 * idx + e] * norm;
            }
    } else {
        T t2 = (T)0;
        for (int c = 0; c < nChunks; ++c) {
            T x[V], y[V], z[V], q[V];
            load_chunk<V, T>(rowp, numEl, c, channels, x, y, z, q);
            wave_scan<V, T>(x, y, z, q, t2);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int idx = c * B + 64 * k + lane;
                if (idx < numEl) {
                    cdf[4 * idx + 0] = (float)(x[k] * norm); cdf[4 * idx + 1] = (float)(y[k] * norm);
                    cdf[4 * idx + 2] = (float)(z[k] * norm); cdf[4 * idx + 3] = (float)(q[k] * norm);
                }
            }
        }
    }
    // the row's marginal term: its total times the row weight (cdf.cu:112-132),
    // formed here so that the last workgroup only scans
    if (lane == 0) {
        if constexpr (kF32) rowTotals[row] = total * row_weight_f32(row >> 2, row & 3, h >> 2);
        else rowTotals[row] = total * row_weight_f64(row, h);
    }
    wave_sync_lds();
    invert_row(cdf, marks, w, cond + (size_t)row * w);

    // ---- the last workgroup builds the marginal
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1u;
    __syncthreads();
    if (!last || wave != 0) return;
    __threadfence();
    const int numElY = h >> 2;
    const int nChunksY = (numElY + B - 1) / B;
    T mtotal = (T)0;
    for (int c = 0; c < nChunksY; ++c) {
        T x[V], y[V], z[V], q[V];
        load_marginal<V, T>(rowTotals, numElY, c, x, y, z, q);
        wave_scan<V, T>(x, y, z, q, mtotal);
        store_chunk<V, T>(rowTotals, numElY, c, x, y, z, q);   // the un-normalised marginal prefix, in place
    }
    const T mnorm = mtotal != (T)0 ? (T)1 / mtotal : (T)0;
    for (int i = lane; i < h; i += 64) {
        const float v = (float)(rowTotals[i] * mnorm);
        cdf[i] = v;
        cdfY[i] = v;
    }
    wave_sync_lds();
    invert_row(cdf, marks, h, marg);
    if (lane == 0) {
        *integral = mtotal;
        // the integral divided by w h as the host does it after the build
        // (FP32: in float, FP64: in double), for callers that read no host value
        if (normIntegral) {
            if constexpr (kF32) *normIntegral = (double)(mtotal / ((float)w * (float)h));
            else *normIntegral = mtotal / ((double)w * (double)h);
        }
        *counter = 0u;   // re-arm for the next build
    }
}

int block_size_for(int w) {
    int numElements = (w + 3) / 4;
    return numElements <= 256 ? 256 : (numElements <= 512 ? 512 : 1024);
}

// Rows per workgroup: as many waves (up to four) as fit 80 KB of LDS,
// so that two workgroups share a CU; at least one.
int waves_for(int w, int h) {
    const size_t row = wave_lds_floats(w > h ? w : h) * sizeof(float);
    return row * 4 <= 81920 ? 4 : (row * 2 <= 81920 ? 2 : 1);
}

template <int V, typename T, int WAVES>
hipError_t launch_cdf_w(hipStream_t s, const float* texels, int w, int h, int ch, uint16_t* cond, uint16_t* marg,
                        float* cdfY, void* rowTotals, void* integral, unsigned int* counter, double* normIntegral) {
    const size_t lds = (size_t)WAVES * wave_lds_floats(w > h ? w : h) * sizeof(float);
    auto fn = k_envmap_cdf<V, T, WAVES>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(h / WAVES), dim3(64 * WAVES), lds, s, texels, w, h, ch, cond, marg, cdfY,
                       static_cast<T*>(rowTotals), static_cast<T*>(integral), counter, normIntegral);
    return hipGetLastError();
}

template <int V, typename T>
hipError_t launch_cdf_v(hipStream_t s, const float* texels, int w, int h, int ch, uint16_t* cond, uint16_t* marg,
                        float* cdfY, void* rowTotals, void* integral, unsigned int* counter, double* normIntegral) {
    const int waves = waves_for(w, h);
    if (waves == 4)
        return launch_cdf_w<V, T, 4>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
    if (waves == 2)
        return launch_cdf_w<V, T, 2>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
    return launch_cdf_w<V, T, 1>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
}

// The fused build over device buffers: texels (w*h*ch floats), outputs cond
// (w*h), marg (h), cdfY (h); scratch: rowTotals (h values of T, T = double
// in FP64 mode), integral (one T) and the last-workgroup counter (zero
// before the first launch; every launch leaves it at zero).
hipError_t envmap_cdf_run(hipStream_t s, bool fp64, const float* texels, int w, int h, int ch, uint16_t* cond,
                          uint16_t* marg, float* cdfY, void* rowTotals, void* integral, unsigned int* counter,
                          double* normIntegral = nullptr) {
    const int B = block_size_for(w);
    if (!fp64) {
        if (B == 256)
            return launch_cdf_v<4, float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
        if (B == 512)
            return launch_cdf_v<8, float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
        return launch_cdf_v<16, float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
    }
    if (B == 256)
        return launch_cdf_v<4, double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
    if (B == 512)
        return launch_cdf_v<8, double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
    return launch_cdf_v<16, double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, counter, normIntegral);
}

// ---- the build without allocations (rbh_envmap_build_noalloc)
//
// Its scratch (the counter, the raw integral, the h weighted row totals)
// lives in this fixed device array: every GPU the library runs on holds its
// own copy until the process ends. Host code never passes the array's name
// to a kernel (in host code it is only a stand-in for the symbol): it uses
// the address hipGetSymbolAddress returned for the device, kept per device.
constexpr size_t kNoAllocCounter = 0, kNoAllocIntegral = 256, kNoAllocRows = 512;
constexpr size_t kNoAllocBytes = kNoAllocRows + (size_t)RBH_ENVMAP_NOALLOC_MAX_ROWS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::once_flag g_scratchOnce[kMaxDevices];
static hipError_t g_scratchErr[kMaxDevices];

__global__ void k_zero_counter(unsigned int* __restrict__ c) {
    if (threadIdx.x == 0) *c = 0u;
}

}  // namespace

// Looks up this device's copy of the scratch array (once per device; the
// device must be current). Called when a context is created.
hipError_t envmap_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_scratchOnce[device], [device]() {
        void* p = nullptr;
        g_scratchErr[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[device] = g_scratchErr[device] == hipSuccess ? p : nullptr;
    });
    return g_scratchErr[device];
}

// The fused build over caller device buffers with the static scratch of
// `device` (looked up by envmap_noalloc_init): no allocation, kernel
// launches on `s` only (the counter is cleared by a kernel, not a memset, so
// the call can be captured into a graph).
hipError_t envmap_cdf_noalloc(int device, hipStream_t s, bool fp64, const float* texels, int w, int h, int ch,
                              uint16_t* cond, uint16_t* marg, float* cdfY, double* normIntegral) {
    if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return hipErrorNotInitialized;
    char* base = static_cast<char*>(g_scratch[device]);
    unsigned int* counter = reinterpret_cast<unsigned int*>(base + kNoAllocCounter);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, s, counter);
    return envmap_cdf_run(s, fp64, texels, w, h, ch, cond, marg, cdfY, base + kNoAllocRows, base + kNoAllocIntegral,
                          counter, normIntegral);
}

// Launches the fused build on the resident env map; returns a HIP error.
hipError_t envmap_cdf_launch(rbh_ctx c, bool fp64) {
    // rowTotals (h values of T) + integral (one T) live in integralDev
    char* base = c->env.integralDev.as<char>();
    return envmap_cdf_run(c->stream, fp64, c->env.texels.as<float>(), c->env.width, c->env.height, c->env.channels,
                          c->env.cond.as<uint16_t>(), c->env.marg.as<uint16_t>(), c->env.cdfY.as<float>(), base + 16,
                          base, c->env.counter.as<unsigned int>());
}

size_t envmap_scratch_bytes(int h) { return 16 + (size_t)h * sizeof(double); }

}  // namespace rbh
