// Test-only kernel: the radiance network's encoding (hvx_mlp.h, mlp_sincos)
// evaluated on the GPU for n inputs and all 16 frequencies, to be compared
// bit for bit with the same function compiled for the CPU. Thread id covers
// (input, frequency) pairs; the grid is bounded by n * 16.
#include <hip/hip_runtime.h>

#include "../../hellovulkan_amd/csrc/hvx_mlp.h"

__global__ __launch_bounds__(256) void mlp_encode_kernel(const float* __restrict__ x, uint32_t n,
                                                         float* __restrict__ sc) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n * 16) return;
    float s, c;
    hvx::mlp_sincos(x[id >> 4], (int)(id & 15), s, c);
    sc[2 * id] = s;
    sc[2 * id + 1] = c;
}

// sc: n x 16 x (sin, cos)
extern "C" int hvxt_mlp_encode(const float* d_x, uint32_t n, float* d_sc) {
    if (n == 0 || n > (1u << 24)) return 1;
    hipLaunchKernelGGL(mlp_encode_kernel, dim3((n * 16 + 255) / 256), dim3(256), 0, 0, d_x, n, d_sc);
    return (int)hipGetLastError();
}
