// TEST INFRASTRUCTURE. Host stand-ins for the HIP device vocabulary that the
// product's device headers (ratbite_amd/csrc/device/*.hpp) use, so that the
// per-path kernel bodies compile as plain C++ and run on the CPU under
// AddressSanitizer / UndefinedBehaviorSanitizer. This header shadows
// <hip/hip_runtime.h> only in the host build of tests/hostbuild; the product
// library is always built by hipcc against the real HIP headers.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>

#define __host__
#define __device__
#define __global__
#define __constant__
#define __shared__ static thread_local
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(...)

struct float2 { float x, y; };
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(8) int2 { int x, y; };
struct alignas(16) int4 { int x, y, z, w; };
struct alignas(8) uint2 { unsigned x, y; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(4) uchar4 { unsigned char x, y, z, w; };
struct alignas(32) double4 { double x, y, z, w; };

inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline double4 make_double4(double x, double y, double z, double w) { return double4{x, y, z, w}; }

inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }
inline int __float_as_int(float f) { int u; std::memcpy(&u, &f, 4); return u; }
inline float __int_as_float(int u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }

using std::isfinite;
using std::isinf;
using std::isnan;
using std::signbit;

// HIP's integer min / max overloads
template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    using T = typename std::common_type<A, B>::type;
    return (T)a < (T)b ? (T)a : (T)b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    using T = typename std::common_type<A, B>::type;
    return (T)a > (T)b ? (T)a : (T)b;
}

template <class T>
inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
