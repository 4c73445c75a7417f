// kscr_selftest.hip -- on-device check of the fast correctly rounded
// arithmetic the kernels use in place of IEEE division and square root
// (kscr_math.h), against the compiler's IEEE sequences, on the device the
// library runs on:
//   0: sqrt_rn(x) != sqrtf(x), over all 2^32 bit patterns;
//   1: rsqrt_rn(x) != 1.0f / sqrtf(x), over all 2^32 patterns (expected
//      only at +0, -0 and +inf, where a vector scaled by either is NaN);
//   2: sqrt_cr / rcp_cr (the reverse sweep's) on clamp_sq(x) against
//      sqrtf / 1.0f / sqrtf of the same clamped argument, over all 2^31
//      non-negative patterns;
//   3: div_rn(a, b, 1.0f / b) != a / b for every a (2^32 patterns) and each
//      of the given divisors b.
// NaN results on both sides count as equal.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kscr.h"
#include "kscr_math.h"

namespace kscr {
namespace {

__device__ inline bool differ(float a, float b) { return __float_as_uint(a) != __float_as_uint(b) && !(a != a && b != b); }

__global__ void rounding_kernel(uint32_t base, unsigned long long* __restrict__ counts) {
    const uint32_t u = base + blockIdx.x * blockDim.x + threadIdx.x;
    const float x = __uint_as_float(u);
    unsigned c0 = differ(sqrt_rn(x), sqrtf(x));
    unsigned c1 = differ(rsqrt_rn(x), 1.0f / sqrtf(x));
    unsigned c2 = 0;
    if (u <= 0x7FFFFFFFu) {
        const float xc = clamp_sq(x);
        const float s = sqrt_cr(xc), si = sqrtf(xc);
        c2 = differ(s, si) || differ(rcp_cr(s), 1.0f / si);
    }
    // one atomic per wave and counter
    const unsigned long long b0 = __ballot(c0), b1 = __ballot(c1), b2 = __ballot(c2);
    if ((threadIdx.x & 63) == 0) {
        if (b0) atomicAdd(&counts[0], (unsigned long long)__popcll(b0));
        if (b1) atomicAdd(&counts[1], (unsigned long long)__popcll(b1));
        if (b2) atomicAdd(&counts[2], (unsigned long long)__popcll(b2));
    }
}

__global__ void division_kernel(uint32_t base, float b, unsigned long long* __restrict__ count) {
    const uint32_t u = base + blockIdx.x * blockDim.x + threadIdx.x;
    const float a = __uint_as_float(u);
    const f3 q = div_rn(f3{a, a, a}, b, 1.0f / b);
    const unsigned long long m = __ballot(differ(q.x, a / b));
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned long long)__popcll(m));
}

}  // namespace
}  // namespace kscr

extern "C" int kscr_selftest_rounding(int device, const float* divisors, int num_divisors,
                                      unsigned long long counts[4]) {
    if (!counts || num_divisors < 0 || (num_divisors > 0 && !divisors)) return KSCR_ERR_INVALID_ARGUMENT;
    hipError_t e;
    if (device >= 0 && (e = hipSetDevice(device)) != hipSuccess) return KSCR_ERR_HIP;
    unsigned long long* d = nullptr;
    if ((e = hipMalloc((void**)&d, 4 * sizeof(unsigned long long))) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
    int rc = KSCR_OK;
    constexpr uint32_t kChunk = 1u << 26;  // patterns per launch
    constexpr int kBlock = 256;
    if ((e = hipMemset(d, 0, 4 * sizeof(unsigned long long))) != hipSuccess) rc = KSCR_ERR_HIP;
    for (uint64_t base = 0; rc == KSCR_OK && base < (1ull << 32); base += kChunk)
        hipLaunchKernelGGL(kscr::rounding_kernel, dim3(kChunk / kBlock), dim3(kBlock), 0, nullptr, (uint32_t)base, d);
    for (int k = 0; rc == KSCR_OK && k < num_divisors; k++)
        for (uint64_t base = 0; base < (1ull << 32); base += kChunk)
            hipLaunchKernelGGL(kscr::division_kernel, dim3(kChunk / kBlock), dim3(kBlock), 0, nullptr, (uint32_t)base,
                               divisors[k], d + 3);
    if (rc == KSCR_OK && ((e = hipGetLastError()) != hipSuccess ||
                          (e = hipMemcpy(counts, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost)) != hipSuccess))
        rc = KSCR_ERR_HIP;
    (void)hipFree(d);
    return rc;
}
