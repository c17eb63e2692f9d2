// Mixed-precision support (FE_PREC_MIXED): the FP32 copy of the table coefficients.
//
// Per (point, pair): R = C - P, T = p R^2 and the Boys vector stay FP64; the
// monomials R^a and the per-order polynomial sums a_m = sum_a g_{a,m} R^a are
// formed in FP32 from the pair-local R (|R| is the point-pair distance, so no
// large absolute coordinates enter the FP32 arithmetic); sum_m a_m F_m and the
// per-point accumulation over pairs are FP64, the latter with Neumaier
// compensated summation.
//
// Precision routing: a pair is evaluated in FP32 only when every lane of the
// wavefront sees a small contribution, estimated by the screening bound
// K*bound / max(|R|, 0.05) < tau (default 1e-4 a.u.); otherwise the whole
// wavefront takes the FP64 contraction for that pair.  Tables uploaded from
// pairs.bin carry no bound and are therefore evaluated in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"

#include "fe_tile.h"

template <int L>
__global__ __launch_bounds__(256) void k_make_g32(const double* __restrict__ rec, const int n,
                                                  float* __restrict__ g32) {
  // the g_k and the scaled coefficients that follow them in the record
  constexpr int S = FeL<L>::STRIDE, NG = FeL<L>::NT + FeL<L>::ND, S32 = FE_S32(NG);
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * S32) return;
  const int pr = (int)(e / S32), k = (int)(e - (int64_t)pr * S32);
  g32[e] = k < NG ? (float)rec[(size_t)pr * S + 8 + k] : 0.0f;
}


// ---------------------------------------------------------------------------
// FP32 copies of every allocated record of class L (records past the kept
// count are converted too and never read)
static void launch_make_g32(fe_table* t, int L, hipStream_t st) {
  const int S32 = FE_S32(fe_host::NTERM[L] + fe_host::NDCOEF[L]);
  const int64_t ne = t->cap_L[L] * S32;
  const dim3 grid((unsigned)((ne + 255) / 256));
  switch (L) {
#define FE_G32(l)                                                                             \
  case l:                                                                                     \
    hipLaunchKernelGGL(k_make_g32<l>, grid, dim3(256), 0, st, t->d_rec[l], (int)t->cap_L[l], \
                       t->d_rec32[l]);                                                        \
    break;
    FE_G32(0) FE_G32(1) FE_G32(2) FE_G32(3) FE_G32(4) FE_G32(5) FE_G32(6)
#undef FE_G32
  }
}

int fe_ensure_g32(fe_table* t, hipStream_t st) {
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!t->cap_L[L] || t->d_rec32[L]) continue;
    const int S32 = FE_S32(fe_host::NTERM[L] + fe_host::NDCOEF[L]);
    const size_t b = (size_t)t->cap_L[L] * S32 * sizeof(float);
    if (hipMalloc(&t->d_rec32[L], b) != hipSuccess) {
      (void)hipGetLastError();
      return fe_set_error(FE_ERR_ALLOC, "mixed: FP32 table allocation failed");
    }
    t->dev_bytes += b;
    launch_make_g32(t, L, st);
  }
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

int fe_refresh_g32(fe_table* t, hipStream_t st) {
  for (int L = 0; L <= FE_LMAX; ++L)
    if (t->cap_L[L] && t->d_rec32[L]) launch_make_g32(t, L, st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
