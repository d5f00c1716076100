// Environment-map importance-sampling tables on gfx950.
//
// Replaces buildCDFLatLongRGBUpTo{4,8,16,32}k + invertCDFToIndices
// (src/kernels/cdf.cu:287-535, launched at rendering_interface.cpp:1364-1419)
// with two kernels. In the first a workgroup of up to eight waves owns a
// row: it streams the row's texels from HBM once, scans them in registers,
// keeps the normalised row CDF in its LDS row, inverts it from LDS into
// uint16 indices and writes only those and the row's marginal term (14 B per
// RGB32F texel of HBM traffic instead of the reference's texel read + FP32
// CDF write + CDF re-read + index write). The second, one workgroup, builds
// the marginal CDF with the sin(theta) Jacobian the same way, inverts it and
// publishes the integral. (A last-workgroup ticket that fused the two cost a
// device-scope fence and a same-address atomic per row, more than the launch
// boundary does.)
//
// FP32 mode keeps the reference's association order exactly: per-thread
// float4 component scan, then the block-wide Hillis-Steele scan over the .w
// lanes with the reference's block size B (256/512/1024, chosen from W/4 as
// rendering_interface.cpp:1380-1382), then the running total of previous
// chunks. Here the row's waves run the B virtual lanes with cross-lane reads
// and two barriers (group_scan). The CDF floats, and therefore the inverted indices,
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

// LDS floats per row: the row (or marginal) CDF of max(w, h) floats, then
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
// per-thread component prefix, then steps 1, 2, 4, ..., B/2 that add the .w
// value of lane t - i from the previous step to all four components of lane
// t >= i), run by the WPR waves of a workgroup with two barriers. Virtual
// lane t = 64 k + lane of the B = 64 V lanes lives in register k - p VW of
// hardware lane `lane` of wave p = k / VW (VW = V / WPR registers per wave).
//  - A step i < 64 reads lane (lane - i) mod 64 of register k (t >= i within
//    the register) or of register k - 1 (the source wrapped into the previous
//    64). For a wave's lowest register, register k - 1 is the top register
//    of the wave below. Each wave keeps a copy of it (the ghost), taken after
//    the component prefix and stepped with its own registers except for the
//    wrapped-in term, which would need the register below it. Step i reads
//    ghost lanes 64 - i .. 63, and those depend only on ghost lanes
//    >= 64 - (i + i/2 + ... + 1) = 65 - 2i >= 1 of earlier states, never on a
//    lane below the step that wrapped (lanes < i' at step i' <= i/2), so the
//    values read are exact.
//  - A step i = 64 m reads register k - m of the same lane. After the steps
//    below 64 each wave publishes its registers' .w values, and every wave
//    runs these steps on the whole column of V values (the additions their
//    owners would make, in the same order), adding each step's term to its
//    own registers; the column's last entry gives the chunk total.
// Every virtual lane adds the same operands in the same order as the
// reference, so the sums are bit-identical.
template <int V, int VW, int P, typename T>
__device__ __forceinline__ void column_steps(T (&col)[V], T (&x)[VW], T (&y)[VW], T (&z)[VW], T (&w)[VW]) {
#pragma unroll
    for (int m = 1; m < V; m <<= 1) {
        // descending k: entry k - m still holds the previous step's value
#pragma unroll
        for (int k = V - 1; k >= m; --k) {
            const T o = col[k - m];
            col[k] += o;
            if (k >= P * VW && k < (P + 1) * VW) {
                x[k - P * VW] += o; y[k - P * VW] += o; z[k - P * VW] += o; w[k - P * VW] += o;
            }
        }
    }
}
template <int V, int VW, int P, typename T>
__device__ __forceinline__ void column_dispatch(int p, T (&col)[V], T (&x)[VW], T (&y)[VW], T (&z)[VW], T (&w)[VW]) {
    if constexpr (P < V / VW) {
        if (p == P) column_steps<V, VW, P, T>(col, x, y, z, w);
        else column_dispatch<V, VW, P + 1, T>(p, col, x, y, z, w);
    }
}

template <int V, int WPR, typename T>
__device__ __forceinline__ void group_scan(T (&x)[V / WPR], T (&y)[V / WPR], T (&z)[V / WPR], T (&w)[V / WPR],
                                           T& total, T* __restrict__ xcol, T* __restrict__ xghost) {
    constexpr int VW = V / WPR;
    const int lane = __lane_id();
    const int p = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        y[k] += x[k];
        z[k] += y[k];
        w[k] += z[k];
    }
    xghost[p * 64 + lane] = w[VW - 1];
    __syncthreads();
    T g = p > 0 ? xghost[(p - 1) * 64 + lane] : (T)0;
#pragma unroll
    for (int i = 1; i < 64; i <<= 1) {
        T s[VW];
        const int src = (lane - i) & 63;
#pragma unroll
        for (int k = 0; k < VW; ++k) s[k] = __shfl(w[k], src);
        const T sg = __shfl(g, src);
        const bool own = lane >= i;
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            if (own || k > 0 || p > 0) {
                const T o = own ? s[k] : (k > 0 ? s[k > 0 ? k - 1 : 0] : sg);
                x[k] += o; y[k] += o; z[k] += o; w[k] += o;
            }
        }
        if (own) g += sg;
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) xcol[(p * VW + k) * 64 + lane] = w[k];
    __syncthreads();
    T col[V];
#pragma unroll
    for (int k = 0; k < V; ++k) col[k] = xcol[k * 64 + lane];
    column_dispatch<V, VW, 0, T>(p, col, x, y, z, w);
#pragma unroll
    for (int k = 0; k < VW; ++k) {
        x[k] += total; y[k] += total; z[k] += total; w[k] += total;
    }
    total = __shfl(col[V - 1] + total, 63);
}

// Element (virtual lane) of register k of this wave in chunk c.
template <int V, int WPR>
__device__ __forceinline__ int part_index(int c, int k) {
    return c * 64 * V + 64 * ((int)(threadIdx.x >> 6) * (V / WPR) + k) + __lane_id();
}

// This wave's part of chunk c of a row's power prefixes: virtual lane t
// covers texels 4t .. 4t+3 (zero past the row).
template <int V, int WPR, typename T>
__device__ __forceinline__ void load_part(const float* __restrict__ rowp, int numEl, int c, int channels,
                                          T (&x)[V / WPR], T (&y)[V / WPR], T (&z)[V / WPR], T (&w)[V / WPR]) {
#pragma unroll
    for (int k = 0; k < V / WPR; ++k) {
        const int idx = part_index<V, WPR>(c, k);
        x[k] = y[k] = z[k] = w[k] = (T)0;
        if (idx < numEl) {
            const V4<float> f = load4(rowp, idx, channels);
            x[k] = (T)f.x; y[k] = (T)f.y; z[k] = (T)f.z; w[k] = (T)f.w;
        }
    }
}

// This wave's part of chunk c of the weighted row totals (the marginal's input).
template <int V, int WPR, typename T>
__device__ __forceinline__ void load_marginal_part(const T* __restrict__ rowTerms, int numElY, int c,
                                                   T (&x)[V / WPR], T (&y)[V / WPR], T (&z)[V / WPR],
                                                   T (&w)[V / WPR]) {
#pragma unroll
    for (int k = 0; k < V / WPR; ++k) {
        const int idx = part_index<V, WPR>(c, k);
        x[k] = y[k] = z[k] = w[k] = (T)0;
        if (idx < numElY) {
            const T* rt = rowTerms + (size_t)idx * 4;
            x[k] = rt[0]; y[k] = rt[1]; z[k] = rt[2]; w[k] = rt[3];
        }
    }
}

template <int V, int WPR, typename T>
__device__ __forceinline__ void store_part(T* __restrict__ dst, int numEl, int c, const T (&x)[V / WPR],
                                           const T (&y)[V / WPR], const T (&z)[V / WPR], const T (&w)[V / WPR]) {
#pragma unroll
    for (int k = 0; k < V / WPR; ++k) {
        const int idx = part_index<V, WPR>(c, k);
        if (idx < numEl) {
            dst[4 * idx + 0] = x[k]; dst[4 * idx + 1] = y[k]; dst[4 * idx + 2] = z[k]; dst[4 * idx + 3] = w[k];
        }
    }
}

// CDF values of this wave's part of chunk c into the LDS row, each value
// times `scale` (no scaling for scale == 1, which the caller passes while the
// row total is still unknown).
template <int V, int WPR, typename T>
__device__ __forceinline__ void store_part_lds(float* __restrict__ cdf, int numEl, int c, const T (&x)[V / WPR],
                                               const T (&y)[V / WPR], const T (&z)[V / WPR], const T (&w)[V / WPR],
                                               bool scaled, T scale) {
#pragma unroll
    for (int k = 0; k < V / WPR; ++k) {
        const int idx = part_index<V, WPR>(c, k);
        if (idx < numEl) {
            f4v v;
            if (scaled) {
                v.x = (float)(x[k] * scale); v.y = (float)(y[k] * scale);
                v.z = (float)(z[k] * scale); v.w = (float)(w[k] * scale);
            } else {
                v.x = (float)x[k]; v.y = (float)y[k]; v.z = (float)z[k]; v.w = (float)w[k];
            }
            *reinterpret_cast<f4v*>(cdf + 4 * idx) = v;
        }
    }
}

// The binary search of cdf.cu:517-534 (invert_one) for outputs 2 pr, 2 pr + 1
// of an LDS-resident CDF of n entries, pr in [pbeg, pend), eight searches
// per lane interleaved so that their LDS reads overlap; the pair is stored
// as one 32-bit word.
__device__ __forceinline__ void invert_search(const float* __restrict__ cdf, int n, uint16_t* __restrict__ out,
                                              int pbeg, int pend) {
    constexpr int K = 8;   // 4 pairs
    const int lane = __lane_id();
    uint32_t* out32 = reinterpret_cast<uint32_t*>(out);
    const float inv = (float)(n - 1);
    for (int base = pbeg; base < pend; base += 64 * (K / 2)) {
        int low[K], high[K], mid[K];
        float p[K];
        bool act[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int pr = base + 64 * (k >> 1) + lane;
            const int j = 2 * pr + (k & 1);
            act[k] = pr < pend;
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
            if (pr < pend) out32[pr] = (uint32_t)high[k] | ((uint32_t)high[k + 1] << 16);
        }
    }
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

// The outputs of invert_one for all j = 0..n-1 in O(n), by the WPR waves of
// the workgroup. The search only probes entries 1..n-2, and while those are
// non-decreasing it returns 1 + #{m in [1, n-2] : cdf[m] < p_j} (at least 1,
// at most n-1). Entry m is below the queries from first_above(cdf[m]) on,
// so output j is the largest m + 1 whose first_above(cdf[m]) <= j: each m
// marks its first query (only the last of several entries sharing one first
// query writes it), and a running maximum over the marks gives every output.
// Each wave takes a contiguous quarter-aligned segment of the outputs; the
// running maximum enters it with the maxima of the segments below. A row
// whose probed entries decrease somewhere (float rounding of the reference's
// prefix sums can make a later entry an ulp smaller) or hold a NaN takes the
// exact binary search instead, as do all rows when `marks` is null.
// On entry the marks (n 16-bit words) are zero and *badFlag is 0, both
// visible to the workgroup; every wave of the workgroup calls this.
template <int WPR>
__device__ __forceinline__ void invert_group(const float* __restrict__ cdf, uint16_t* __restrict__ marks, int n,
                                             uint16_t* __restrict__ out, int* badFlag, uint32_t* segMax) {
    const int lane = __lane_id();
    const int p = (int)(threadIdx.x >> 6);
    const int pairs = n >> 1;
    const int pb = (int)((long long)p * pairs / WPR), pe = (int)((long long)(p + 1) * pairs / WPR);
    if (marks == nullptr) {
        invert_search(cdf, n, out, pb, pe);
        return;
    }
    // one pass over the probed entries: the order check and the marks. Each
    // lane takes G consecutive entries; the first query above the entry after
    // its last one comes from the next lane (lane 63 computes it), so
    // first_above runs G + 1 times per G entries instead of 2 G.
    constexpr int G = 4;
    bool bad = false;
    for (int base = 1 + G * (int)threadIdx.x; base - G * lane <= n - 2; base += G * 64 * WPR) {
        float c[G + 1];
        int j[G + 1];
#pragma unroll
        for (int g = 0; g <= G; ++g) c[g] = base + g <= n - 1 ? cdf[base + g] : 0.0f;
#pragma unroll
        for (int g = 0; g < G; ++g) j[g] = base + g <= n - 2 ? first_above(c[g], n) : n;
        j[G] = __shfl_down(j[0], 1);
        if (lane == 63) j[G] = base + G <= n - 2 ? first_above(c[G], n) : n;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int m = base + g;
            if (m <= n - 2) {
                const bool last = m + 1 > n - 2;
                bad = bad || !(c[g] >= 0.0f) || (!last && !(c[g + 1] >= c[g]));
                const int jm = j[g];
                const int jn = last ? n : j[g + 1];
                if (jm < jn && jm < n) marks[jm] = (uint16_t)(m + 1);
            }
        }
    }
    if (__ballot(bad) != 0ull && lane == 0) *badFlag = 1;
    __syncthreads();
    if (*badFlag) {
        invert_search(cdf, n, out, pb, pe);
        return;
    }
    // segment of groups of four outputs (n is a multiple of 4)
    const int nq = n >> 2;
    const int q0 = (int)((long long)p * nq / WPR), q1 = (int)((long long)(p + 1) * nq / WPR);
    const uint2* m64 = reinterpret_cast<const uint2*>(marks);
    uint32_t mx = 0u;
    for (int q = q0 + lane; q < q1; q += 64) {
        const uint2 g = m64[q];
        mx = max(mx, max(max(g.x & 0xFFFFu, g.x >> 16), max(g.y & 0xFFFFu, g.y >> 16)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o));
    if (lane == 0) segMax[p] = mx;
    __syncthreads();
    uint32_t carry = 1u;
    for (int pp = 0; pp < p; ++pp) carry = max(carry, segMax[pp]);
    // running maximum, four consecutive outputs per lane: in-lane maximum,
    // then a wave-wide exclusive maximum
    uint2* o64 = reinterpret_cast<uint2*>(out);
    for (int base = q0; base < q1; base += 64) {
        const int q = base + lane;
        const bool live = q < q1;
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

// Zeroes n 16-bit marks (n a multiple of 4) with the whole workgroup.
template <int WPR>
__device__ __forceinline__ void zero_marks(uint16_t* __restrict__ marks, int n) {
    if (marks == nullptr) return;
    uint2* z64 = reinterpret_cast<uint2*>(marks);
    for (int q = (int)threadIdx.x; 4 * q < n; q += 64 * WPR) z64[q] = make_uint2(0u, 0u);
}

// One workgroup of WPR waves per row: the row's power prefixes in
// registers (chunks of the reference's B = 64 V lanes, running total between
// chunks), the normalised CDF in the workgroup's LDS row, the row's inverted
// indices from LDS, and the row's marginal term for the marginal kernel.
// FP32 normalises the prefixes as it stores them when the row is one chunk,
// otherwise keeps them in LDS and normalises them in place; FP64 scans the
// row twice (the second time from L2) and rounds each normalised value to
// float once.
template <int V, int WPR, typename T>
__global__ __launch_bounds__(64 * WPR) void k_envmap_rows(const float* __restrict__ texels, int w, int h, int channels,
                                                           uint16_t* __restrict__ cond, T* __restrict__ rowTotals) {
    constexpr int VW = V / WPR;
    constexpr int B = 64 * V;
    constexpr bool kF32 = sizeof(T) == 4;
    extern __shared__ __align__(16) float lds[];
    __shared__ int badFlag;
    __shared__ uint32_t segMax[WPR];
    float* cdf = lds;
    uint16_t* marks = use_marks(w) ? reinterpret_cast<uint16_t*>(cdf + w) : nullptr;
    T* xcol = reinterpret_cast<T*>(lds + wave_lds_floats(w));
    T* xghost = xcol + V * 64;
    const int row = blockIdx.x;
    const int numEl = w >> 2;
    const int nChunks = (numEl + B - 1) / B;
    const float* rowp = texels + (size_t)row * w * channels;
    zero_marks<WPR>(marks, w);
    if (threadIdx.x == 0) badFlag = 0;
    // (the scan's barriers order these before the inversion)

    T total = (T)0;
    for (int c = 0; c < nChunks; ++c) {
        T x[VW], y[VW], z[VW], q[VW];
        load_part<V, WPR, T>(rowp, numEl, c, channels, x, y, z, q);
        group_scan<V, WPR, T>(x, y, z, q, total, xcol, xghost);
        if constexpr (kF32) {
            // cdf.cu:380-392: a one-chunk row is normalised as it is stored
            const T norm = total != (T)0 ? (T)1 / total : (T)0;
            store_part_lds<V, WPR, T>(cdf, numEl, c, x, y, z, q, nChunks == 1, norm);
        }
    }
    const T norm = total != (T)0 ? (T)1 / total : (T)0;
    if constexpr (kF32) {
        if (nChunks > 1) {
            // in place, each lane rescaling what it wrote
            for (int c = 0; c < nChunks; ++c)
#pragma unroll
                for (int k = 0; k < VW; ++k) {
                    const int idx = part_index<V, WPR>(c, k);
                    if (idx < numEl) {
                        f4v* e = reinterpret_cast<f4v*>(cdf + 4 * idx);
                        f4v v = *e;
                        v.x = v.x * norm; v.y = v.y * norm; v.z = v.z * norm; v.w = v.w * norm;
                        *e = v;
                    }
                }
        }
    } else {
        T t2 = (T)0;
        for (int c = 0; c < nChunks; ++c) {
            T x[VW], y[VW], z[VW], q[VW];
            load_part<V, WPR, T>(rowp, numEl, c, channels, x, y, z, q);
            group_scan<V, WPR, T>(x, y, z, q, t2, xcol, xghost);
            store_part_lds<V, WPR, T>(cdf, numEl, c, x, y, z, q, true, norm);
        }
    }
    // the row's marginal term: its total times the row weight (cdf.cu:112-132),
    // formed here so that the marginal kernel only scans
    if (threadIdx.x == 0) {
        if constexpr (kF32) rowTotals[row] = total * row_weight_f32(row >> 2, row & 3, h >> 2);
        else rowTotals[row] = total * row_weight_f64(row, h);
    }
    __syncthreads();
    invert_group<WPR>(cdf, marks, w, cond + (size_t)row * w, &badFlag, segMax);
}

// The marginal CDF over the h row terms in one workgroup, the same scan with
// the same block size B as the rows (cdf.cu:112-132, 287-392), its
// inversion, and the integral (the un-normalised total; also divided by w h
// into normIntegral when given).
template <int V, int WPR, typename T>
__global__ __launch_bounds__(64 * WPR) void k_envmap_marginal(int w, int h, uint16_t* __restrict__ marg,
                                                               float* __restrict__ cdfY, T* __restrict__ rowTotals,
                                                               T* __restrict__ integral, double* __restrict__ normIntegral) {
    constexpr int VW = V / WPR;
    constexpr int B = 64 * V;
    extern __shared__ __align__(16) float lds[];
    __shared__ int badFlag;
    __shared__ uint32_t segMax[WPR];
    float* cdf = lds;
    uint16_t* marks = use_marks(h) ? reinterpret_cast<uint16_t*>(cdf + h) : nullptr;
    T* xcol = reinterpret_cast<T*>(lds + wave_lds_floats(h));
    T* xghost = xcol + V * 64;
    zero_marks<WPR>(marks, h);
    if (threadIdx.x == 0) badFlag = 0;
    const int numElY = h >> 2;
    const int nChunksY = (numElY + B - 1) / B;
    T mtotal = (T)0;
    if (nChunksY == 1) {
        // the usual case: the prefixes stay in registers until normalised
        T x[VW], y[VW], z[VW], q[VW];
        load_marginal_part<V, WPR, T>(rowTotals, numElY, 0, x, y, z, q);
        group_scan<V, WPR, T>(x, y, z, q, mtotal, xcol, xghost);
        const T mnorm = mtotal != (T)0 ? (T)1 / mtotal : (T)0;
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            const int idx = part_index<V, WPR>(0, k);
            if (idx < numElY) {
                f4v v;
                v.x = (float)(x[k] * mnorm); v.y = (float)(y[k] * mnorm);
                v.z = (float)(z[k] * mnorm); v.w = (float)(q[k] * mnorm);
                *reinterpret_cast<f4v*>(cdf + 4 * idx) = v;
                cdfY[4 * idx + 0] = v.x; cdfY[4 * idx + 1] = v.y;   // a caller's buffer: 4-byte alignment only
                cdfY[4 * idx + 2] = v.z; cdfY[4 * idx + 3] = v.w;
            }
        }
    } else {
        for (int c = 0; c < nChunksY; ++c) {
            T x[VW], y[VW], z[VW], q[VW];
            load_marginal_part<V, WPR, T>(rowTotals, numElY, c, x, y, z, q);
            group_scan<V, WPR, T>(x, y, z, q, mtotal, xcol, xghost);
            store_part<V, WPR, T>(rowTotals, numElY, c, x, y, z, q);   // the un-normalised marginal prefix, in place
        }
        const T mnorm = mtotal != (T)0 ? (T)1 / mtotal : (T)0;
        // each lane normalises the prefixes it stored
        for (int c = 0; c < nChunksY; ++c)
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                const int idx = part_index<V, WPR>(c, k);
                if (idx < numElY)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float v = (float)(rowTotals[4 * idx + e] * mnorm);
                        cdf[4 * idx + e] = v;
                        cdfY[4 * idx + e] = v;
                    }
            }
    }
    __syncthreads();
    invert_group<WPR>(cdf, marks, h, marg, &badFlag, segMax);
    if (threadIdx.x == 0) {
        *integral = mtotal;
        // the integral divided by w h as the host does it after the build
        // (FP32: in float, FP64: in double), for callers that read no host value
        if (normIntegral) {
            if constexpr (sizeof(T) == 4) *normIntegral = (double)(mtotal / ((float)w * (float)h));
            else *normIntegral = mtotal / ((double)w * (double)h);
        }
    }
}

// Waves per row: each holds V / row_waves(V) of the scan's virtual registers.
constexpr int kRowWaves = 8;
constexpr int row_waves(int V) { return V < kRowWaves ? V : kRowWaves; }

int block_size_for(int w) {
    int numElements = (w + 3) / 4;
    return numElements <= 256 ? 256 : (numElements <= 512 ? 512 : 1024);
}

// Dynamic LDS of a workgroup: a CDF row of n floats and its marks, then the
// scan's exchange rows (V registers and WPR ghosts of 64 values of T).
template <typename T>
size_t cdf_lds_bytes(int n, int V, int WPR) {
    return wave_lds_floats(n) * sizeof(float) + (size_t)(V + WPR) * 64 * sizeof(T);
}

template <typename K>
hipError_t set_lds(K fn, size_t lds) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <int V, int WPR, typename T>
hipError_t launch_cdf_v(hipStream_t s, const float* texels, int w, int h, int ch, uint16_t* cond, uint16_t* marg,
                        float* cdfY, void* rowTotals, void* integral, double* normIntegral) {
    const size_t ldsRows = cdf_lds_bytes<T>(w, V, WPR), ldsMarg = cdf_lds_bytes<T>(h, V, WPR);
    auto rows = k_envmap_rows<V, WPR, T>;
    auto margK = k_envmap_marginal<V, WPR, T>;
    hipError_t e = set_lds(rows, ldsRows);
    if (e == hipSuccess) e = set_lds(margK, ldsMarg);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rows, dim3(h), dim3(64 * WPR), ldsRows, s, texels, w, h, ch, cond, static_cast<T*>(rowTotals));
    hipLaunchKernelGGL(margK, dim3(1), dim3(64 * WPR), ldsMarg, s, w, h, marg, cdfY, static_cast<T*>(rowTotals),
                       static_cast<T*>(integral), normIntegral);
    return hipGetLastError();
}

// The build over device buffers: texels (w*h*ch floats), outputs cond
// (w*h), marg (h), cdfY (h); scratch: rowTotals (h values of T, T = double
// in FP64 mode) and integral (one T). Two launches on `s`: the rows, then
// the marginal.
hipError_t envmap_cdf_run(hipStream_t s, bool fp64, const float* texels, int w, int h, int ch, uint16_t* cond,
                          uint16_t* marg, float* cdfY, void* rowTotals, void* integral,
                          double* normIntegral = nullptr) {
    const int B = block_size_for(w);
    if (!fp64) {
        if (B == 256)
            return launch_cdf_v<4, row_waves(4), float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
        if (B == 512)
            return launch_cdf_v<8, row_waves(8), float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
        return launch_cdf_v<16, row_waves(16), float>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
    }
    if (B == 256)
        return launch_cdf_v<4, row_waves(4), double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
    if (B == 512)
        return launch_cdf_v<8, row_waves(8), double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
    return launch_cdf_v<16, row_waves(16), double>(s, texels, w, h, ch, cond, marg, cdfY, rowTotals, integral, normIntegral);
}

// ---- the build without allocations (rbh_envmap_build_noalloc)
//
// Its scratch (the raw integral, the h weighted row totals)
// lives in this fixed device array: every GPU the library runs on holds its
// own copy until the process ends. Host code never passes the array's name
// to a kernel (in host code it is only a stand-in for the symbol): it uses
// the address hipGetSymbolAddress returned for the device, kept per device.
constexpr size_t kNoAllocIntegral = 0, kNoAllocRows = 256;
constexpr size_t kNoAllocBytes = kNoAllocRows + (size_t)RBH_ENVMAP_NOALLOC_MAX_ROWS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::once_flag g_scratchOnce[kMaxDevices];
static hipError_t g_scratchErr[kMaxDevices];

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

// The build over caller device buffers with the static scratch of `device`
// (looked up by envmap_noalloc_init): no allocation and no memset, kernel
// launches on `s` only, so the call can be captured into a graph.
hipError_t envmap_cdf_noalloc(int device, hipStream_t s, bool fp64, const float* texels, int w, int h, int ch,
                              uint16_t* cond, uint16_t* marg, float* cdfY, double* normIntegral) {
    if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return hipErrorNotInitialized;
    char* base = static_cast<char*>(g_scratch[device]);
    return envmap_cdf_run(s, fp64, texels, w, h, ch, cond, marg, cdfY, base + kNoAllocRows, base + kNoAllocIntegral,
                          normIntegral);
}

// Launches the build on the resident env map; returns a HIP error.
hipError_t envmap_cdf_launch(rbh_ctx c, bool fp64) {
    // rowTotals (h values of T) + integral (one T) live in integralDev
    char* base = c->env.integralDev.as<char>();
    return envmap_cdf_run(c->stream, fp64, c->env.texels.as<float>(), c->env.width, c->env.height, c->env.channels,
                          c->env.cond.as<uint16_t>(), c->env.marg.as<uint16_t>(), c->env.cdfY.as<float>(), base + 16,
                          base);
}

size_t envmap_scratch_bytes(int h) { return 16 + (size_t)h * sizeof(double); }

}  // namespace rbh
