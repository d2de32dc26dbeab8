// Test kernel: the SOA sub-step's fast division (mam4xx_amd/csrc/scalar.hpp,
// num::div_fast) against IEEE a / b, bit for bit.  Operands span the ranges
// the sub-step feeds it: numerators +0 and 1e-40..1e40 of either sign, or
// (tiny) +0 and 2^-1074..2^-960 of either sign (the subnormals and the small
// normals either side of the 2^-968 bound); denominators 1e-25..1e25.  Pairs
// inside the range the division is exact on (|a| = 0 or >= 2^-968, |a / b|
// >= 2^-1021) are counted apart from those outside it.  Test infrastructure
// only.
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

// bad[0]: mismatches inside the range, bad[1]: outside it, bad[2]: pairs
// outside it
__global__ void div_check_kernel(long n, unsigned long long seed, int tiny,
                                 unsigned long long *bad, double *first) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long z1 = mix(seed + 2 * (unsigned long long)i);
  const unsigned long long z2 = mix(seed + 2 * (unsigned long long)i + 1);
  // log-uniform magnitudes with a random mantissa
  double a = tiny ? exp2(-1074.0 + 114.0 * unit(z1)) * (1.0 + unit(z2))
                  : exp10(-40.0 + 80.0 * unit(z1)) * (1.0 + unit(z2));
  const double b = exp10(-25.0 + 50.0 * unit(z2)) * (1.0 + unit(z1));
  if (z2 & 1) a = -a;
  if ((z1 & 15) == 0) a = 0.0;  // +0: a -0 numerator gives +0 (see scalar.hpp)
  const double q0 = a / b;
  const double q1 = mam4x::num::div_fast(a, b);
  const bool inside = a == 0.0 || (fabs(a) >= 0x1p-968 && fabs(q0) >= 0x1p-1021);
  if (!inside) atomicAdd(&bad[2], 1ull);
  if (__double_as_longlong(q0) != __double_as_longlong(q1)) {
    if (atomicAdd(&bad[inside ? 0 : 1], 1ull) == 0 && inside) {
      first[0] = a;
      first[1] = b;
    }
  }
}

}  // namespace

// Returns the number of operand pairs (out of n) inside the exact range where
// the two quotients differ in any bit; the first such pair goes to
// first[0..1].  outside[0] receives the mismatches outside the range,
// outside[1] the number of pairs there.  -1 on a HIP error.  tiny != 0 draws
// the numerators from 2^-1074..2^-960.
extern "C" long long div_check(long n, unsigned long long seed, int tiny, double *first,
                               unsigned long long *outside) {
  unsigned long long *d_bad = nullptr;
  double *d_first = nullptr;
  if (hipMalloc(&d_bad, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMalloc(&d_first, 2 * sizeof(double)) != hipSuccess) return -1;
  long long rc = 0;
  if (hipMemset(d_bad, 0, 3 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(d_first, 0, 2 * sizeof(double)) != hipSuccess)
    rc = -1;
  unsigned long long bad[3] = {0, 0, 0};
  if (rc == 0) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(div_check_kernel, dim3(grid), dim3(256), 0, 0, n, seed, tiny, d_bad,
                       d_first);
    if (hipMemcpy(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(first, d_first, 2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = -1;
  }
  if (hipFree(d_bad) != hipSuccess || hipFree(d_first) != hipSuccess) rc = -1;
  outside[0] = bad[1];
  outside[1] = bad[2];
  return rc ? rc : (long long)bad[0];
}
