// post_kernels.hip — the tonemap post step on the FP32 shading output
// (hvx_post_tonemap of include/hvx.h): PipelineTonemap
// (HelloVulkan/Source/Pipelines/PipelineTonemap.cpp:34-53) running
// Shaders/Common/Tonemap.frag:17-94, one lane per pixel.
//
// main() maps each pixel with Reinhard (pure white 1.0) and then gamma 2.2;
// the shader also defines ACES, Filmic and Uncharted2 curves, selectable
// here. Every expression is evaluated in FP32 in the shader's order, with
// the shared pow of hvx_math.h, so the output equals the oracle bit for bit.
// With pure white 1.0 the Reinhard curve is the identity on luminance,
// L (1 + L) / (1 + L) = L, so main() is a gamma curve (up to rounding); a
// pixel of zero luminance gives 0/0 = NaN, as in the shader.
#include <hip/hip_runtime.h>

#include "hvx_device.h"
#include "hvx_internal.h"

namespace hvx {
namespace {

__device__ __forceinline__ float clamp01(float x) { return gmin(gmax(x, 0.0f), 1.0f); }

__device__ __forceinline__ float partial_uncharted2(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

__device__ __forceinline__ uint32_t unorm8(float x) {
    const float c = x == x ? clamp01(x) : 0.0f;  // NaN converts to 0
    return (uint32_t)__builtin_rintf(c * 255.0f);
}

__global__ __launch_bounds__(256) void tonemap_kernel(const float* __restrict__ rgb, uint32_t n, int op, int gamma,
                                                      float* __restrict__ out, uint32_t* __restrict__ bgra) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float c[3] = {rgb[i], rgb[(size_t)n + i], rgb[2 * (size_t)n + i]};
    if (op == HVX_TONEMAP_REINHARD) {  // Tonemap.frag:26-35
        const float pureWhite = 1.0f;
        const float L = (c[0] * 0.2126f + c[1] * 0.7152f) + c[2] * 0.0722f;
        const float mapped = (L * (1.0f + L / (pureWhite * pureWhite))) / (1.0f + L);
        const float k = mapped / L;
        for (float& v : c) v = k * v;
    } else if (op == HVX_TONEMAP_ACES) {  // :37-45
        const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
        for (float& x : c) x = clamp01((x * (a * x + b)) / (x * (cc * x + d) + e));
    } else if (op == HVX_TONEMAP_FILMIC) {  // :47-52
        for (float& x : c) {
            const float X = gmax(0.0f, x - 0.004f);
            x = powf_((X * (6.2f * X + 0.5f)) / (X * (6.2f * X + 1.7f) + 0.06f), 2.2f);
        }
    } else {  // HVX_TONEMAP_UNCHARTED2, :65-73
        const float exposureBias = 2.0f;
        const float whiteScale = 1.0f / partial_uncharted2(11.2f);
        for (float& x : c) x = partial_uncharted2(x * exposureBias) * whiteScale;
    }
    if (gamma) {  // GammaCorrection, :75-79
        const float inv = 1.0f / 2.2f;
        for (float& v : c) v = powf_(v, inv);
    }
    if (out) {
        out[i] = c[0];
        out[(size_t)n + i] = c[1];
        out[2 * (size_t)n + i] = c[2];
    }
    if (bgra) bgra[i] = unorm8(c[2]) | (unorm8(c[1]) << 8) | (unorm8(c[0]) << 16) | (255u << 24);
}

}  // namespace

hipError_t launch_tonemap(const float* d_rgb, uint32_t n, int op, int gamma, float* d_out, uint32_t* d_bgra,
                          hipStream_t s) {
    hipLaunchKernelGGL(tonemap_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_rgb, n, op, gamma, d_out, d_bgra);
    return hipGetLastError();
}

}  // namespace hvx
