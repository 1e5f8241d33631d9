// Test-only kernel: one wave multiplies A (16x32) by B (32x16), given as
// row-major FP32 arrays, with the 16x16x32 MFMA, distributing operands and
// collecting results by the lane layout stated in hvx_mfma.h. A wrong layout
// assumption gives a wrong product (all accesses stay inside the 16x32,
// 32x16 and 16x16 arrays).
#include <hip/hip_runtime.h>

#include "../../hellovulkan_amd/csrc/hvx_mfma.h"

using namespace hvx;

template <class V, class E>
__global__ __launch_bounds__(64) void mfma_layout_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                         float* __restrict__ C) {
    const int l = threadIdx.x, i = l & 15, g = l >> 4;
    V a, b;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        a[q] = (E)A[i * 32 + 8 * g + q];
        b[q] = (E)B[(8 * g + q) * 16 + i];
    }
    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = mfma16x16x32(a, b, c);
#pragma unroll
    for (int r = 0; r < 4; ++r) C[(4 * g + r) * 16 + i] = c[r];
}

extern "C" int hvxt_mfma_layout(int bf16, const float* d_a, const float* d_b, float* d_c) {
    if (bf16)
        hipLaunchKernelGGL((mfma_layout_kernel<bf16x8, __bf16>), dim3(1), dim3(64), 0, 0, d_a, d_b, d_c);
    else
        hipLaunchKernelGGL((mfma_layout_kernel<half8, _Float16>), dim3(1), dim3(64), 0, 0, d_a, d_b, d_c);
    return (int)hipGetLastError();
}
