// Exhaustive device checks of the library's own rounding helpers
// (volumefusion_amd/csrc/vf_device.hpp), run by tests/test_gpu_numerics.py:
//
//   rcp   rcp_rn(x) against the IEEE quotient 1.0f / x, every positive float
//         from 2^-126 up to +inf (the Newton-step range and both edges of its
//         guard);
//   sqrt  sqrt_rn(s) against __builtin_sqrtf(s), every non-negative float and
//         +inf.
//
// Each kernel walks its range with a grid-stride loop and only counts
// mismatches (first few recorded); it touches no memory beyond two small
// result buffers.  Prints one line per check: "<name> checked N mismatches M".
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "vf_device.hpp"

namespace {

__global__ void k_rcp(uint32_t lo, uint32_t hi, unsigned long long* bad, uint32_t* first) {
    const uint32_t n = hi - lo;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float x = __uint_as_float(lo + i);
        if (__float_as_uint(vf::rcp_rn(x)) != __float_as_uint(1.0f / x)) {
            const unsigned long long c = atomicAdd(bad, 1ull);
            if (c < 8) first[c] = lo + i;
        }
    }
}

__global__ void k_sqrt(uint32_t lo, uint32_t hi, unsigned long long* bad, uint32_t* first) {
    const uint32_t n = hi - lo;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float s = __uint_as_float(lo + i);
        if (__float_as_uint(vf::sqrt_rn(s)) != __float_as_uint(__builtin_sqrtf(s))) {
            const unsigned long long c = atomicAdd(bad, 1ull);
            if (c < 8) first[c] = lo + i;
        }
    }
}

int run(const char* name, void (*kern)(uint32_t, uint32_t, unsigned long long*, uint32_t*), uint32_t lo,
        uint32_t hi) {
    unsigned long long* bad = nullptr;
    uint32_t* first = nullptr;
    if (hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess || hipMalloc(&first, 8 * sizeof(uint32_t)) != hipSuccess) {
        printf("%s: allocation failed\n", name);
        return 1;
    }
    if (hipMemset(bad, 0, sizeof(unsigned long long)) != hipSuccess || hipMemset(first, 0, 8 * sizeof(uint32_t)) != hipSuccess)
        return 1;
    hipLaunchKernelGGL(kern, dim3(8192), dim3(256), 0, 0, lo, hi, bad, first);
    const hipError_t e = hipDeviceSynchronize();
    unsigned long long h = 0;
    uint32_t f[8] = {0};
    if (e != hipSuccess || hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(f, first, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) {
        printf("%s: %s\n", name, hipGetErrorString(e));
        return 1;
    }
    printf("%s checked %u mismatches %llu", name, hi - lo, h);
    for (unsigned k = 0; k < 8 && k < h; ++k) printf(" 0x%08x", f[k]);
    printf("\n");
    (void)hipFree(bad);
    (void)hipFree(first);
    return h != 0;
}

}  // namespace

int main() {
    int st = 0;
    // positive floats from 2^-126 (0x00800000) through +inf (0x7f800000)
    st |= run("rcp", k_rcp, 0x00800000u, 0x7f800001u);
    // +0 through +inf
    st |= run("sqrt", k_sqrt, 0x00000000u, 0x7f800001u);
    return st;
}
