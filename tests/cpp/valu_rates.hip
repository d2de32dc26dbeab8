// Throughput probe of the scalar VALU forms the chain's precision policies
// compile to: FP64 and FP32 fused multiply-add, and the device-libm exp and
// log in both precisions.  Built with the library's flags (packed FP32 forms
// disabled), so an FP32 FMA issues as v_fma_f32, one per lane.  Each lane
// runs 8 independent dependency chains so the VALU, not latency, bounds the
// rate.  Test and measurement infrastructure only.
#include <hip/hip_runtime.h>

#include <cmath>

namespace {

constexpr int kChains = 8;

// one fused multiply-add in the operands' own precision (__builtin_fma is
// the double form: float operands would be promoted to it)
__device__ __forceinline__ double dfma(double x, double a, double b) { return __builtin_fma(x, a, b); }
__device__ __forceinline__ float dfma(float x, float a, float b) { return __builtin_fmaf(x, a, b); }

template <class T>
__global__ void fma_kernel(T *out, T a, T b, int iters) {
  T x[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) x[j] = T(threadIdx.x + j) * T(1e-3);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < kChains; ++j) x[j] = dfma(x[j], a, b);
  }
  T s = 0;
#pragma unroll
  for (int j = 0; j < kChains; ++j) s += x[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__device__ __forceinline__ double dexp(double v) { return exp(v); }
__device__ __forceinline__ float dexp(float v) { return expf(v); }
__device__ __forceinline__ double dlog(double v) { return log(v); }
__device__ __forceinline__ float dlog(float v) { return logf(v); }

// x <- exp(a x): converges to a fixed point near 1 for a = -1e-3
template <class T>
__global__ void exp_kernel(T *out, T a, int iters) {
  T x[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) x[j] = T(1) + T(threadIdx.x + j) * T(1e-6);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < kChains; ++j) x[j] = dexp(x[j] * a);
  }
  T s = 0;
#pragma unroll
  for (int j = 0; j < kChains; ++j) s += x[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// x <- log(x + b): converges to a fixed point near 1.15 for b = 2
template <class T>
__global__ void log_kernel(T *out, T b, int iters) {
  T x[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) x[j] = T(1) + T(threadIdx.x + j) * T(1e-6);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < kChains; ++j) x[j] = dlog(x[j] + b);
  }
  T s = 0;
#pragma unroll
  for (int j = 0; j < kChains; ++j) s += x[j];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

constexpr int kBlocks = 8192, kThreads = 256;

template <class K>
float time_ms(K launch) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  launch();  // warm-up
  (void)hipEventRecord(e0, 0);
  launch();
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

}  // namespace

// ops[6] <- operations per second of: FP64 FMA, FP32 FMA, FP64 exp, FP32
// exp, FP64 log, FP32 log (one operation = one FMA or one function call per
// lane).  check[6] <- a finite sum of each kernel's results.  Returns 0, or
// -1 on a HIP error.
extern "C" int valu_rates(double *ops, double *check) {
  const long n = (long)kBlocks * kThreads;
  double *d64 = nullptr;
  float *d32 = nullptr;
  if (hipMalloc(&d64, n * sizeof(double)) != hipSuccess) return -1;
  if (hipMalloc(&d32, n * sizeof(float)) != hipSuccess) return -1;
  const int fma_iters = 16384, tr_iters = 512;
  const dim3 g(kBlocks), b(kThreads);
  float ms[6];
  ms[0] = time_ms([&] { hipLaunchKernelGGL(fma_kernel<double>, g, b, 0, 0, d64, 0.999999,
                                           1e-6, fma_iters); });
  ms[1] = time_ms([&] { hipLaunchKernelGGL(fma_kernel<float>, g, b, 0, 0, d32, 0.999999f,
                                           1e-6f, fma_iters); });
  ms[2] = time_ms([&] { hipLaunchKernelGGL(exp_kernel<double>, g, b, 0, 0, d64, -1e-3,
                                           tr_iters); });
  ms[3] = time_ms([&] { hipLaunchKernelGGL(exp_kernel<float>, g, b, 0, 0, d32, -1e-3f,
                                           tr_iters); });
  ms[4] = time_ms([&] { hipLaunchKernelGGL(log_kernel<double>, g, b, 0, 0, d64, 2.0,
                                           tr_iters); });
  ms[5] = time_ms([&] { hipLaunchKernelGGL(log_kernel<float>, g, b, 0, 0, d32, 2.0f,
                                           tr_iters); });
  const int iters[6] = {fma_iters, fma_iters, tr_iters, tr_iters, tr_iters, tr_iters};
  for (int k = 0; k < 6; ++k) ops[k] = (double)n * kChains * iters[k] / (ms[k] * 1e-3);
  // a checksum of the last kernel of each precision, so no result is dead
  double *h64 = new double[n];
  float *h32 = new float[n];
  (void)hipMemcpy(h64, d64, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipMemcpy(h32, d32, n * sizeof(float), hipMemcpyDeviceToHost);
  double s64 = 0, s32 = 0;
  for (long i = 0; i < n; ++i) {
    s64 += h64[i];
    s32 += h32[i];
  }
  for (int k = 0; k < 6; ++k) check[k] = (k & 1) ? s32 : s64;
  delete[] h64;
  delete[] h32;
  const bool ok = hipGetLastError() == hipSuccess;
  (void)hipFree(d64);
  (void)hipFree(d32);
  return ok ? 0 : -1;
}
