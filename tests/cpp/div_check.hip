// Test kernel: the correctly rounded division of the SOA sub-step
// (mam4xx_amd/csrc/scalar.hpp, num::div) against IEEE a / b, bit for bit,
// over operands spanning the ranges the sub-step feeds it (numerators +0 and
// 1e-40..1e40 of either sign, denominators 1e-25..1e25).  Test
// infrastructure only.
#include <hip/hip_runtime.h>

#include "../../mam4xx_amd/csrc/scalar.hpp"

namespace {

__device__ unsigned long long mix(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in [0, 1) from 53 random bits
__device__ double unit(unsigned long long z) { return (double)(z >> 11) * 0x1.0p-53; }

__global__ void div_check_kernel(long n, unsigned long long seed, unsigned long long *bad,
                                 double *first) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long z1 = mix(seed + 2 * (unsigned long long)i);
  const unsigned long long z2 = mix(seed + 2 * (unsigned long long)i + 1);
  // log-uniform magnitudes with a random mantissa
  double a = exp10(-40.0 + 80.0 * unit(z1)) * (1.0 + unit(z2));
  const double b = exp10(-25.0 + 50.0 * unit(z2)) * (1.0 + unit(z1));
  if (z2 & 1) a = -a;
  if ((z1 & 15) == 0) a = 0.0;  // +0: a -0 numerator gives +0 (see scalar.hpp)
  const double q0 = a / b;
  const double q1 = mam4x::num::div(a, b);
  if (__double_as_longlong(q0) != __double_as_longlong(q1)) {
    if (atomicAdd(bad, 1ull) == 0) {
      first[0] = a;
      first[1] = b;
    }
  }
}

}  // namespace

// Returns the number of operand pairs (out of n) where the two quotients
// differ in any bit; the first such pair goes to first[0..1].  -1 on a HIP
// error.
extern "C" long long div_check(long n, unsigned long long seed, double *first) {
  unsigned long long *d_bad = nullptr;
  double *d_first = nullptr;
  if (hipMalloc(&d_bad, sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMalloc(&d_first, 2 * sizeof(double)) != hipSuccess) return -1;
  long long rc = 0;
  if (hipMemset(d_bad, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(d_first, 0, 2 * sizeof(double)) != hipSuccess)
    rc = -1;
  unsigned long long bad = 0;
  if (rc == 0) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(div_check_kernel, dim3(grid), dim3(256), 0, 0, n, seed, d_bad, d_first);
    if (hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(first, d_first, 2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = -1;
  }
  if (hipFree(d_bad) != hipSuccess || hipFree(d_first) != hipSuccess) rc = -1;
  return rc ? rc : (long long)bad;
}
