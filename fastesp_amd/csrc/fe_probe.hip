// VALU rate calibration: the sustained FMA rate of this device in FP64 and in
// FP32 as this library is compiled (packed FP32 disabled), measured with
// independent FMA chains.  It backs the roofline peaks the benchmark divides by
// and shows what the mixed-precision contraction can gain per instruction.
#include "fe_internal.h"

template <typename T>
__global__ __launch_bounds__(256) void k_fma_chains(const int iters, const T b, const T c,
                                                    T* __restrict__ out) {
  T a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (T)(threadIdx.x + i);
  for (int k = 0; k < iters; ++k) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = fma(a[i], b, c);
  }
  T s = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += a[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <typename T>
static int run_chains(fe_ctx* c, const int iters, double* tflops) {
  const int blocks = c->n_cu * 4 * 8 / 4;  // 8 waves per SIMD
  T* out = nullptr;
  FE_HIPCHK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(T)));
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipFree(out);
    return fe_set_error(FE_ERR_HIP, "fe_valu_rate: event creation failed");
  }
  // b, c keep the chains finite: a -> a b + c converges to c / (1 - b)
  const T b = (T)0.5, cc = (T)0.25;
  hipLaunchKernelGGL(k_fma_chains<T>, dim3(blocks), dim3(256), 0, 0, iters, b, cc, out);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_fma_chains<T>, dim3(blocks), dim3(256), 0, 0, iters, b, cc, out);
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_HIP, "fe_valu_rate: %s", hipGetErrorString(e));
  }
  *tflops = (double)blocks * 256 * iters * 8 * 2 / (ms * 1e-3) / 1e12;
  return FE_OK;
}

extern "C" int fe_valu_rate(fe_ctx* c, int precision, double* tflops) {
  if (!c || !tflops) return fe_set_error(FE_ERR_ARG, "fe_valu_rate: null argument");
  FE_HIPCHK(hipSetDevice(c->device));
  const int iters = 1 << 16;
  if (precision == FE_PREC_FP64) return run_chains<double>(c, iters, tflops);
  if (precision == FE_PREC_MIXED) return run_chains<float>(c, iters, tflops);
  return fe_set_error(FE_ERR_ARG, "fe_valu_rate: precision must be FE_PREC_FP64 or FE_PREC_MIXED");
}
