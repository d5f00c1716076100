// Stable LSD radix sort and stream compaction on gfx950, for the LBVH build
// (Morton keys, node depths) and the guide's training-set compaction.
//
// One primitive does the work: ranking the 64 keys of a wave by an 8-bit
// digit with ballots. A lane's peers are the lanes whose digit agrees with
// its own on each of the eight bit ballots; its rank among them is the number
// of peers below it, and the lowest peer publishes their count. A tile of
// 4096 keys (256 threads x 16) is ranked in 16 rounds of 256 keys; per round
// the four waves' digit counts are turned into bases in wave order in LDS, so
// a key's place follows its index: the sort is stable. Across tiles, each
// tile's digit counts are written digit-major (digit d of tile t at
// d * tiles + t) and one exclusive scan turns them into every tile's first
// output slot per digit. A pass is three launches: digit counts, scan,
// ranked scatter; a sort is one pass per 8 key bits, ping-ponging between
// the output and a scratch buffer so that the last pass writes the output.
// Compaction is the same with a one-bit digit: per-tile counts of selected
// items, their scan (whose total is the number selected), ranked writes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rbh {
namespace radix {

constexpr int kBlock = 256;                 // threads per tile = digits per pass
constexpr int kWaves = kBlock / 64;
constexpr int kPerThread = 16;
constexpr int kTile = kBlock * kPerThread;  // keys per tile
constexpr int kScanBlock = 1024;

inline size_t tiles_for(size_t n) { return (n + kTile - 1) / kTile; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch of a sort of n keys (and values, valBytes 0 without): the digit
// counts, then the alternate key and value buffers.
inline size_t sort_scratch_bytes(size_t n, size_t keyBytes, size_t valBytes) {
    return align256(tiles_for(n) * kBlock * 4) + align256(n * keyBytes) + align256(n * valBytes);
}
// Scratch of a compaction of n items: the per-tile counts.
inline size_t select_scratch_bytes(size_t n) { return align256(tiles_for(n) * 4 + 4); }

__device__ __forceinline__ uint64_t lanes_below() { return (1ull << __lane_id()) - 1ull; }

template <typename K>
__global__ __launch_bounds__(kBlock) void k_digit_counts(const K* __restrict__ keys, uint32_t n, int shift,
                                                         uint32_t mask, uint32_t* __restrict__ counts, uint32_t tiles) {
    __shared__ uint32_t hist[kBlock];
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
    for (int i = 0; i < kPerThread; ++i) {
        const size_t j = base + (size_t)i * kBlock + threadIdx.x;
        if (j < n) atomicAdd(&hist[(uint32_t)(keys[j] >> shift) & mask], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive prefix sum of a[0..m) in place by one workgroup of B threads,
// four consecutive elements per thread per round; *total (when given) = the
// sum. (A template so that every translation unit may instantiate it.)
template <int B>
__global__ __launch_bounds__(B) void k_exclusive_scan(uint32_t* __restrict__ a, uint32_t m,
                                                      uint32_t* __restrict__ total) {
    constexpr int W = B / 64;
    __shared__ uint32_t wsum[W];
    const int lane = __lane_id(), w = (int)(threadIdx.x >> 6);
    uint32_t carry = 0u;
    for (uint32_t base = 0; base < m; base += B * 4) {
        const uint32_t j = base + threadIdx.x * 4;
        uint32_t v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = j + e < m ? a[j + e] : 0u;
        const uint32_t s = v[0] + v[1] + v[2] + v[3];
        uint32_t x = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(x, o);
            if (lane >= o) x += t;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t below = 0u, all = 0u;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const uint32_t ws = wsum[k];
            below += k < w ? ws : 0u;
            all += ws;
        }
        uint32_t run = carry + below + x - s;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (j + e < m) a[j + e] = run;
            run += v[e];
        }
        carry += all;
        __syncthreads();   // wsum is rewritten next round
    }
    if (total && threadIdx.x == 0) *total = carry;
}

// One pass of the sort: the keys (and values) of each tile to their slots.
template <typename K, typename VT, bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_ranked_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                           const VT* __restrict__ vin, VT* __restrict__ vout,
                                                           uint32_t n, int shift, uint32_t mask,
                                                           const uint32_t* __restrict__ offsets, uint32_t tiles) {
    __shared__ uint32_t run[kBlock];
    __shared__ uint32_t wcnt[kWaves][kBlock];
    __shared__ uint32_t wbase[kWaves][kBlock];
    const int tid = (int)threadIdx.x, w = tid >> 6;
    run[tid] = offsets[(size_t)tid * tiles + blockIdx.x];
#pragma unroll
    for (int k = 0; k < kWaves; ++k) wcnt[k][tid] = 0u;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint64_t below = lanes_below();
    for (int i = 0; i < kPerThread; ++i) {
        const size_t j = base + (size_t)i * kBlock + tid;
        const bool valid = j < n;
        const K key = valid ? kin[j] : K(0);
        VT val{};
        if (PAIRS && valid) val = vin[j];
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0u) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            // digit tid: the four waves' bases in wave order
            uint32_t r = run[tid];
#pragma unroll
            for (int k = 0; k < kWaves; ++k) {
                const uint32_t c = wcnt[k][tid];
                wbase[k][tid] = r;
                r += c;
                wcnt[k][tid] = 0u;
            }
            run[tid] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = wbase[w][d] + rank;
            kout[pos] = key;
            if (PAIRS) vout[pos] = val;
        }
    }
}

// Stable sort of n keys (with values when vin is non-null) by bits
// [beginBit, endBit): kin/vin are left unchanged, the result goes to
// kout/vout. `scratch` holds sort_scratch_bytes(n, sizeof(K), sizeof(VT))
// bytes (valBytes 0 for keys only).
template <typename K, typename VT>
hipError_t sort(void* scratch, const K* kin, K* kout, const VT* vin, VT* vout, uint32_t n, int beginBit, int endBit,
                hipStream_t s) {
    if (n == 0 || endBit <= beginBit) return hipSuccess;
    const uint32_t tiles = (uint32_t)tiles_for(n);
    char* p = static_cast<char*>(scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(p);
    K* altK = reinterpret_cast<K*>(p + align256((size_t)tiles * kBlock * 4));
    VT* altV = reinterpret_cast<VT*>(reinterpret_cast<char*>(altK) + align256((size_t)n * sizeof(K)));
    const int passes = (endBit - beginBit + 7) / 8;
    const K* src = kin;
    const VT* vsrc = vin;
    for (int pass = 0; pass < passes; ++pass) {
        const bool toOut = ((passes - 1 - pass) & 1) == 0;
        K* dst = toOut ? kout : altK;
        VT* vdst = toOut ? vout : altV;
        const int shift = beginBit + 8 * pass;
        const int bits = endBit - shift < 8 ? endBit - shift : 8;
        const uint32_t mask = (1u << bits) - 1u;
        hipLaunchKernelGGL(k_digit_counts<K>, dim3(tiles), dim3(kBlock), 0, s, src, n, shift, mask, counts, tiles);
        hipLaunchKernelGGL(k_exclusive_scan<kScanBlock>, dim3(1), dim3(kScanBlock), 0, s, counts, tiles * (uint32_t)kBlock,
                           (uint32_t*)nullptr);
        if (vin)
            hipLaunchKernelGGL((k_ranked_scatter<K, VT, true>), dim3(tiles), dim3(kBlock), 0, s, src, dst, vsrc, vdst,
                               n, shift, mask, counts, tiles);
        else
            hipLaunchKernelGGL((k_ranked_scatter<K, VT, false>), dim3(tiles), dim3(kBlock), 0, s, src, dst, vsrc,
                               vdst, n, shift, mask, counts, tiles);
        src = dst;
        vsrc = vdst;
    }
    return hipGetLastError();
}

template <typename Pred>
__global__ __launch_bounds__(kBlock) void k_select_counts(uint32_t n, Pred pred, uint32_t* __restrict__ counts) {
    __shared__ uint32_t wc[kWaves];
    const int tid = (int)threadIdx.x, w = tid >> 6;
    const size_t base = (size_t)blockIdx.x * kTile;
    uint32_t c = 0u;
    for (int i = 0; i < kPerThread; ++i) {
        const size_t j = base + (size_t)i * kBlock + tid;
        c += (uint32_t)__popcll(__ballot(j < n && pred((uint32_t)j)));
    }
    if (__lane_id() == 0) wc[w] = c;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0u;
        for (int k = 0; k < kWaves; ++k) t += wc[k];
        counts[blockIdx.x] = t;
    }
}

template <typename Pred>
__global__ __launch_bounds__(kBlock) void k_select_write(uint32_t n, Pred pred, const uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ out) {
    __shared__ uint32_t wc[kWaves];
    const int tid = (int)threadIdx.x, w = tid >> 6;
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint64_t below = lanes_below();
    uint32_t run = offsets[blockIdx.x];
    for (int i = 0; i < kPerThread; ++i) {
        const size_t j = base + (size_t)i * kBlock + tid;
        const bool flag = j < n && pred((uint32_t)j);
        const uint64_t m = __ballot(flag);
        if (__lane_id() == 0) wc[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0u, all = 0u;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            const uint32_t v = wc[k];
            before += k < w ? v : 0u;
            all += v;
        }
        if (flag) out[run + before + (uint32_t)__popcll(m & below)] = (uint32_t)j;
        run += all;
        __syncthreads();   // wc is rewritten next round
    }
}

// The indices i in [0, n) with pred(i), in increasing order, into out; their
// number into *count (device memory). `scratch` holds
// select_scratch_bytes(n) bytes.
template <typename Pred>
hipError_t select_indices(void* scratch, uint32_t n, Pred pred, uint32_t* out, uint32_t* count, hipStream_t s) {
    const uint32_t tiles = (uint32_t)tiles_for(n);
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    if (tiles > 0) hipLaunchKernelGGL(k_select_counts<Pred>, dim3(tiles), dim3(kBlock), 0, s, n, pred, counts);
    hipLaunchKernelGGL(k_exclusive_scan<kScanBlock>, dim3(1), dim3(kScanBlock), 0, s, counts, tiles, count);
    if (tiles > 0) hipLaunchKernelGGL(k_select_write<Pred>, dim3(tiles), dim3(kBlock), 0, s, n, pred, counts, out);
    return hipGetLastError();
}

}  // namespace radix
}  // namespace rbh
