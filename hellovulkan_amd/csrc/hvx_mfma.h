// hvx_mfma.h — operand types and the lane layout of the gfx950 16x16x32
// MFMA instructions used by the IBL contraction kernels.
//
// v_mfma_f32_16x16x32_{f16,bf16}: D(16x16, f32) = A(16x32) * B(32x16) + C.
// For lane l of the 64-lane wave, with i = l & 15 and g = l >> 4:
//   A operand: 8 values A[i][8g .. 8g+7]       (row i, 8 consecutive k)
//   B operand: 8 values B[8g .. 8g+7][i]       (column i, 8 consecutive k)
//   C/D:       4 floats C[4g .. 4g+3][i]       (rows 4g..4g+3 of column i)
// tests/gpu_kernels/mfma_layout.hip checks this layout on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace hvx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16x16x32(const half8& a, const half8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16x16x32(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace hvx
