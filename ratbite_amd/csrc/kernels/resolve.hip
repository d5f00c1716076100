// renderResolve (src/kernels/image.cu:92-121): FP64 accumulation -> mean ->
// clamp at 0 -> exposure -> XYZ-to-RGB matrix -> AgX (image.cu:27-60) ->
// RGBA8. The AgX inset/outset matrices and the contrast polynomial are the
// published AgX constants that image.cu uses.
#include "../context.hpp"

namespace rbh {

namespace {

struct M3c { float m[9]; };  // column-major

__device__ __forceinline__ float3 mul3(const float* m, float3 v) {
    return make_float3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
                       m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
__device__ __forceinline__ float agx_poly(float x) {
    const float x2 = x * x, x4 = x2 * x2;
    return -17.86f * x4 * x2 * x + 78.01f * x4 * x2 - 126.7f * x4 * x + 92.06f * x4 - 28.72f * x2 * x + 4.361f * x2 -
           0.1718f * x + 0.002857f;
}

__global__ __launch_bounds__(256) void k_resolve(uint32_t w, uint32_t h, M3c cs, float expo, const double4* __restrict__ acc,
                                                 uchar4* __restrict__ out) {
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= w || y >= h) return;
    const float agx[9] = {0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f,
                          0.0784335999999992f, 0.878468636469772f, 0.0784336f,
                          0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f};
    const float iagx[9] = {1.1968790051201738155f, -0.052896851757456180321f, -0.052971635514443794537f,
                           -0.098020881140136776078f, 1.1519031299041727435f, -0.098043450117124120312f,
                           -0.099029744079720471434f, -0.098961176844843346553f, 1.1510736726411610622f};
    const double4 d = acc[(size_t)y * w + x];
    const double nv = 1.0 / d.w;
    float3 c = make_float3((float)(d.x * nv), (float)(d.y * nv), (float)(d.z * nv));
    c = make_float3(fmaxf(c.x, 0.0f) * expo, fmaxf(c.y, 0.0f) * expo, fmaxf(c.z, 0.0f) * expo);
    c = mul3(cs.m, c);
    c = mul3(agx, c);
    const float lo = -12.47393f, hi = 4.026069f;
    c.x = (fminf(fmaxf(log2f(c.x), lo), hi) - lo) / (hi - lo);
    c.y = (fminf(fmaxf(log2f(c.y), lo), hi) - lo) / (hi - lo);
    c.z = (fminf(fmaxf(log2f(c.z), lo), hi) - lo) / (hi - lo);
    c = make_float3(agx_poly(c.x), agx_poly(c.y), agx_poly(c.z));
    c = mul3(iagx, c);
    c = make_float3(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f));
    out[(size_t)y * w + x] = make_uchar4((unsigned char)(c.x * 255.99999f), (unsigned char)(c.y * 255.99999f),
                                         (unsigned char)(c.z * 255.99999f), 255);
}

}  // namespace

hipError_t resolve_launch(rbh_ctx c, const float* colorspace, float expo, void* out) {
    M3c m;
    for (int i = 0; i < 9; ++i) m.m[i] = colorspace[i];
    const uint32_t w = c->rs.width, h = c->rs.height;
    hipLaunchKernelGGL(k_resolve, dim3((w + 15) / 16, (h + 15) / 16), dim3(16, 16), 0, c->stream, w, h, m, expo,
                       c->rs.accum.as<double4>(), static_cast<uchar4*>(out));
    return hipGetLastError();
}

}  // namespace rbh
