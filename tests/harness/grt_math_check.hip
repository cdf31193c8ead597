// grt_math_check.hip -- test infrastructure: checks on the GPU that the
// tracing kernel's sqrt_rn / rcp_rn (grt_device.hpp) return the same bits as
// the compiler's sqrt() and 1.0 / d over pseudo-random finite arguments
// spread over binary exponents [-700, 700], plus fixed edge values; and that
// the step controller's fast (1/err)^(1/4) (quarter_root_rcp) equals the
// generic quarter_root(1.0 / err), the oracle's form, over err spread across
// the fast path's range.
#include <hip/hip_runtime.h>

#include "grt_device.hpp"

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void check(uint64_t n, uint64_t seed, unsigned long long* bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = splitmix64(seed + i);
    // mantissa from the low 52 bits, exponent uniform in [-700, 700]
    const uint64_t e = (uint64_t)(1023 - 700) + (b >> 52) % 1401u;
    const double x = __longlong_as_double((long long)((e << 52) | (b & 0xFFFFFFFFFFFFFull)));
    unsigned long long s = 0, r = 0;
    if (__double_as_longlong(grt::sqrt_rn(x)) != __double_as_longlong(sqrt(x))) s = 1;
    if (__double_as_longlong(grt::rcp_rn(x)) != __double_as_longlong(1.0 / x)) r = 1;
    if (s) atomicAdd(&bad[0], 1ull);
    if (r) atomicAdd(&bad[1], 1ull);
}

__global__ void check_values(const double* v, uint64_t n, unsigned long long* bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = v[i];
    if (__double_as_longlong(grt::sqrt_rn(x)) != __double_as_longlong(sqrt(x))) atomicAdd(&bad[0], 1ull);
    if (__double_as_longlong(grt::rcp_rn(x)) != __double_as_longlong(1.0 / x)) atomicAdd(&bad[1], 1ull);
}

// err with exponent uniform in [-499, 499]: the range of quarter_root_rcp's
// short sequences (2^-500, 2^500).
__global__ void check_quarter_root(uint64_t n, uint64_t seed, unsigned long long* bad)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = splitmix64(seed ^ (i * 0xD1B54A32D192ED03ull));
    const uint64_t e = (uint64_t)(1023 - 499) + (b >> 52) % 999u;
    const double err = __longlong_as_double((long long)((e << 52) | (b & 0xFFFFFFFFFFFFFull)));
    const double fast = grt::quarter_root_rcp(err);
    const double ref = grt::quarter_root(1.0 / err);
    if (__double_as_longlong(fast) != __double_as_longlong(ref)) atomicAdd(bad, 1ull);
}

}  // namespace

extern "C" int grt_quarter_root_check(uint64_t n, uint64_t seed, unsigned long long* mismatches)
{
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemset(d, 0, sizeof(unsigned long long)) != hipSuccess) return -1;
    if (n)
        hipLaunchKernelGGL(check_quarter_root, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                           n, seed, d);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : -2;
}

extern "C" int grt_math_check(uint64_t n, uint64_t seed, const double* host_values, uint64_t nv,
                              unsigned long long* mismatches)
{
    unsigned long long* d = nullptr;
    double* dv = nullptr;
    if (hipMalloc(&d, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemset(d, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (n) hipLaunchKernelGGL(check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, seed, d);
    if (nv) {
        if (hipMalloc(&dv, nv * sizeof(double)) != hipSuccess) return -1;
        if (hipMemcpy(dv, host_values, nv * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return -1;
        hipLaunchKernelGGL(check_values, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, 0, dv,
                           nv, d + 2);
    }
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(mismatches, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (dv) (void)hipFree(dv);
    return e == hipSuccess ? 0 : -2;
}
