// Test-only kernels: exhaustive comparison of the fast correctly rounded
// square root / reciprocal of hvx_device.h against the compiler's IEEE
// sqrtf and division. Built into tests/_build/libhvx_testkernels.so.
#include <hip/hip_runtime.h>

#include "../../hellovulkan_amd/csrc/hvx_device.h"

namespace {
__global__ void check_kernel(uint32_t first_bits, uint32_t n, int which, unsigned long long* mismatches,
                             uint32_t* first_bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = hvx::u2f(first_bits + i);
    float got, ref;
    if (which == 0) {
        got = hvx::sqrt_rn_normal(x);
        ref = hvx::sqrtf_(x);
    } else {
        got = hvx::rcp_rn_normal(x);
        ref = 1.0f / x;
    }
    if (hvx::f2u(got) != hvx::f2u(ref)) {
        atomicAdd(mismatches, 1ull);
        atomicMin(first_bad, first_bits + i);
    }
}
}  // namespace

extern "C" int hvxt_fastmath_check(uint32_t first_bits, uint32_t n, int which, unsigned long long* d_mismatches,
                                   uint32_t* d_first_bad) {
    hipLaunchKernelGGL(check_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, first_bits, n, which, d_mismatches,
                       d_first_bad);
    return (int)hipGetLastError();
}
