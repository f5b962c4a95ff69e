// Checks of the correctly rounded helpers of device_math.hpp against the
// compiler's IEEE operations: div_rn against x / b (on the host and on the
// GPU) and sqrt_rn against sqrtf (on the GPU, where the hardware estimate
// it corrects lives). Built by physica_amd.build into
// tests/native/libphx_math_check.so and driven from tests/test_device_math.py.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "device_math.hpp"

namespace {

uint32_t bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}

float from_bits(uint32_t u) {
    float x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}

// mode 0: sqrt_rn(x) vs sqrtf(x); mode 1: div_rn(x, b, inv_b) vs x / b, for
// the floats with bit patterns first .. first + count - 1.
__global__ void __launch_bounds__(256)
    math_check_kernel(int mode, uint32_t first, uint32_t count, float b, float inv_b, unsigned long long* mismatches) {
    const uint32_t stride = gridDim.x * blockDim.x;
    unsigned long long wrong = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += stride) {
        const float x = __uint_as_float(first + t);
        if (mode == 0) {
            wrong += __float_as_uint(phx::sqrt_rn(x)) != __float_as_uint(sqrtf(x)) ? 1u : 0u;
        } else {
            wrong += __float_as_uint(phx::div_rn(x, b, inv_b)) != __float_as_uint(x / b) ? 1u : 0u;
        }
    }
    if (wrong != 0u) atomicAdd(mismatches, wrong);
}

}  // namespace

// Number of floats x with bit patterns in [first, first + count) for which
// div_rn(x, b, RN(1 / b)) differs from x / b, on the host.
extern "C" uint64_t phx_test_div_host(float b, uint32_t first, uint32_t count) {
    const float inv_b = 1.0F / b;
    uint64_t wrong = 0;
    for (uint32_t t = 0; t < count; ++t) {
        const float x = from_bits(first + t);
        wrong += bits(phx::div_rn(x, b, inv_b)) != bits(x / b) ? 1u : 0u;
    }
    return wrong;
}

// The same count on the GPU for mode 1, or the count of sqrt_rn(x) !=
// sqrtf(x) for mode 0. Returns a hipError_t; the count goes to *mismatches.
extern "C" int phx_test_math_gpu(int mode, uint32_t first, uint32_t count, float b, unsigned long long* mismatches) {
    if (mismatches == nullptr || (mode != 0 && mode != 1) || count == 0u || first + (count - 1u) < first) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    unsigned long long* device = nullptr;
    hipError_t e = hipMalloc(&device, sizeof(unsigned long long));
    if (e != hipSuccess) return static_cast<int>(e);
    e = hipMemset(device, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        math_check_kernel<<<2048, 256>>>(mode, first, count, b, 1.0F / b, device);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(mismatches, device, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    const hipError_t f = hipFree(device);
    return static_cast<int>(e != hipSuccess ? e : f);
}
