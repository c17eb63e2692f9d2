// VALU rate calibration for the benchmark (libfastesp_probe.so, not part of
// the product library): the sustained FMA rate of a device in FP64 and in FP32
// as the product library is compiled (packed FP32 disabled), measured with
// independent FMA chains.  It shows what the roofline peaks mean on this
// instruction mix and what the mixed-precision contraction can gain per
// instruction.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {
thread_local char g_msg[256] = "";

int fail(const char* what, hipError_t e) {
  (void)hipGetLastError();
  snprintf(g_msg, sizeof g_msg, "%s: %s", what, hipGetErrorString(e));
  return -2;
}
}  // namespace

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
static int run_chains(const int n_cu, const int iters, double* tflops) {
  const int blocks = n_cu * 4 * 8 / 4;  // 8 waves per SIMD
  T* out = nullptr;
  hipError_t e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(T));
  if (e != hipSuccess) return fail("hipMalloc", e);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipFree(out);
    return fail("hipEventCreate", e);
  }
  // b, c keep the chains finite: a -> a b + c converges to c / (1 - b)
  const T b = (T)0.5, cc = (T)0.25;
  hipLaunchKernelGGL(k_fma_chains<T>, dim3(blocks), dim3(256), 0, 0, iters, b, cc, out);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_fma_chains<T>, dim3(blocks), dim3(256), 0, 0, iters, b, cc, out);
  (void)hipEventRecord(e1, 0);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (e != hipSuccess) return fail("k_fma_chains", e);
  *tflops = (double)blocks * 256 * iters * 8 * 2 / (ms * 1e-3) / 1e12;
  return 0;
}

// precision 0: FP64 chains, 1: FP32 chains; returns 0 or a negative code
extern "C" int fe_probe_valu_rate(int device, int precision, double* tflops) {
  if (!tflops || (precision != 0 && precision != 1)) {
    snprintf(g_msg, sizeof g_msg, "fe_probe_valu_rate: bad argument");
    return -1;
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail("hipSetDevice", e);
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess)
    return fail("hipGetDeviceProperties", e);
  const int iters = 1 << 16;
  return precision == 0 ? run_chains<double>(prop.multiProcessorCount, iters, tflops)
                        : run_chains<float>(prop.multiProcessorCount, iters, tflops);
}

extern "C" const char* fe_probe_last_error(void) { return g_msg; }
