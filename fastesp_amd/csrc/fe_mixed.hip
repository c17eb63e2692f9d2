// Mixed-precision evaluation (FE_PREC_MIXED).
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

#define FE_S32(NT) ((((NT) + 15) / 16) * 16)

__device__ __forceinline__ void fe_kahan_add(double& s, double& c, const double y) {
  const double t = s + y;
  c += (fabs(s) >= fabs(y)) ? (s - t) + y : (y - t) + s;
  s = t;
}

template <int L>
__global__ __launch_bounds__(256) void k_make_g32(const double* __restrict__ rec, const int n,
                                                  float* __restrict__ g32) {
  constexpr int S = FeL<L>::STRIDE, NT = FeL<L>::NT, S32 = FE_S32(NT);
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * S32) return;
  const int pr = (int)(e / S32), k = (int)(e - (int64_t)pr * S32);
  g32[e] = k < NT ? (float)rec[(size_t)pr * S + 8 + k] : 0.0f;
}

template <int L, int PPL>
__global__ __launch_bounds__(FE_WG) void k_esp_fwd_mixed(
    const double* __restrict__ rec, const float* __restrict__ g32, const int n_pairs, const int ppc,
    const double* __restrict__ pts, const int n_pts, const double* __restrict__ boys_tab,
    double* __restrict__ partial, const int row0, const int ld, const float tau) {
  __shared__ double s_tab[FE_BOYS_NR * FE_BOYS_COLS];
  for (int i = threadIdx.x; i < FE_BOYS_NR * FE_BOYS_COLS; i += FE_WG) s_tab[i] = boys_tab[i];
  __syncthreads();
  constexpr int S = FeL<L>::STRIDE, S32 = FE_S32(FeL<L>::NT);
  const int chunk = blockIdx.y;
  const int p0 = chunk * ppc;
  const int p1 = min(n_pairs, p0 + ppc);
  double cx[PPL], cy[PPL], cz[PPL], acc[PPL], cmp[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = (blockIdx.x * PPL + q) * FE_WG + threadIdx.x;
    const int ii = min(i, n_pts - 1);
    cx[q] = pts[3 * ii];
    cy[q] = pts[3 * ii + 1];
    cz[q] = pts[3 * ii + 2];
    acc[q] = cmp[q] = 0.0;
  }
  for (int ip = p0; ip < p1; ++ip) {
    const double* __restrict__ r = rec + (size_t)ip * S;
    const float* __restrict__ g = g32 + (size_t)ip * S32;
    const double Px = r[0], Py = r[1], Pz = r[2], p = r[3], K = r[4];
    const float kb = (float)r[5];
    double Rx[PPL], Ry[PPL], Rz[PPL], r2[PPL];
    int big = 0;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      Rx[q] = cx[q] - Px;
      Ry[q] = cy[q] - Py;
      Rz[q] = cz[q] - Pz;
      r2[q] = fma(Rx[q], Rx[q], fma(Ry[q], Ry[q], Rz[q] * Rz[q]));
      big |= kb * rsqrtf(fmaxf((float)r2[q], 0.0025f)) >= tau;
    }
    const bool fp64 = __any(big);
    int near = 0;
#pragma unroll
    for (int q = 0; q < PPL; ++q) near |= p * r2[q] < fe_tfar(L);
    const bool far = !__any(near);
    const double hp = r[6], KF = K * r[7];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      double F[L + 1];
      double w = K;
      if (far) {
        const double ri = rsqrt(r2[q]);
        fe_boys_far_scaled<L>(hp * (ri * ri), F);
        w = KF * ri;
      } else {
        fe_boys<L>(p * r2[q], F, s_tab);
      }
      double s;
      if (fp64)
        s = fe_contract<L>(Rx[q], Ry[q], Rz[q], F, r + 8);
      else
        s = fe_contract_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g);
      fe_kahan_add(acc[q], cmp[q], w * s);
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = (blockIdx.x * PPL + q) * FE_WG + threadIdx.x;
    if (i < n_pts) partial[(size_t)(row0 + chunk) * ld + i] = acc[q] + cmp[q];
  }
}

template <int L, int PPL>
__global__ __launch_bounds__(FE_WG) void k_esp_fwd_grad_mixed(
    const double* __restrict__ rec, const float* __restrict__ g32, const int n_pairs, const int ppc,
    const double* __restrict__ pts, const int n_pts, const double* __restrict__ boys_tab,
    double* __restrict__ partial, const int row0, const int ld, const size_t plane,
    const float tau) {
  __shared__ double s_tab[FE_BOYS_NR * FE_BOYS_COLS];
  for (int i = threadIdx.x; i < FE_BOYS_NR * FE_BOYS_COLS; i += FE_WG) s_tab[i] = boys_tab[i];
  __syncthreads();
  constexpr int S = FeL<L>::STRIDE, S32 = FE_S32(FeL<L>::NT);
  const int chunk = blockIdx.y;
  const int p0 = chunk * ppc;
  const int p1 = min(n_pairs, p0 + ppc);
  double cx[PPL], cy[PPL], cz[PPL], av[PPL], cv[PPL], ax[PPL], ay[PPL], az[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = (blockIdx.x * PPL + q) * FE_WG + threadIdx.x;
    const int ii = min(i, n_pts - 1);
    cx[q] = pts[3 * ii];
    cy[q] = pts[3 * ii + 1];
    cz[q] = pts[3 * ii + 2];
    av[q] = cv[q] = ax[q] = ay[q] = az[q] = 0.0;
  }
  for (int ip = p0; ip < p1; ++ip) {
    const double* __restrict__ r = rec + (size_t)ip * S;
    const float* __restrict__ g = g32 + (size_t)ip * S32;
    const double Px = r[0], Py = r[1], Pz = r[2], p = r[3], K = r[4];
    const float kb = (float)r[5];
    double Rx[PPL], Ry[PPL], Rz[PPL], r2[PPL];
    int big = 0;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      Rx[q] = cx[q] - Px;
      Ry[q] = cy[q] - Py;
      Rz[q] = cz[q] - Pz;
      r2[q] = fma(Rx[q], Rx[q], fma(Ry[q], Ry[q], Rz[q] * Rz[q]));
      // the field decays one power faster; use the same (conservative) estimate
      big |= kb * rsqrtf(fmaxf((float)r2[q], 0.0025f)) >= tau;
    }
    const bool fp64 = __any(big);
    int near = 0;
#pragma unroll
    for (int q = 0; q < PPL; ++q) near |= p * r2[q] < fe_tfar(L + 1);
    const bool far = !__any(near);
    const double hp = r[6], KF = K * r[7];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      double F[L + 2];
      double w = K;
      if (far) {
        const double ri = rsqrt(r2[q]);
        fe_boys_far_scaled<L + 1>(hp * (ri * ri), F);
        w = KF * ri;
      } else {
        fe_boys<L + 1>(p * r2[q], F, s_tab);
      }
      double I, Gx, Gy, Gz;
      if (fp64)
        fe_contract_grad<L>(Rx[q], Ry[q], Rz[q], F, r + 8, p, I, Gx, Gy, Gz);
      else
        fe_contract_grad_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g, p, Rx[q], Ry[q],
                                Rz[q], I, Gx, Gy, Gz);
      fe_kahan_add(av[q], cv[q], w * I);
      ax[q] = fma(w, Gx, ax[q]);
      ay[q] = fma(w, Gy, ay[q]);
      az[q] = fma(w, Gz, az[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = (blockIdx.x * PPL + q) * FE_WG + threadIdx.x;
    if (i < n_pts) {
      const size_t o = (size_t)(row0 + chunk) * ld + i;
      partial[o] = av[q] + cv[q];
      partial[plane + o] = ax[q];
      partial[2 * plane + o] = ay[q];
      partial[3 * plane + o] = az[q];
    }
  }
}

// ---------------------------------------------------------------------------
int fe_ensure_g32(fe_table* t, hipStream_t st) {
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!t->n_L[L] || t->d_rec32[L]) continue;
    const int S32 = FE_S32(fe_host::NTERM[L]);
    const size_t b = (size_t)t->n_L[L] * S32 * sizeof(float);
    if (hipMalloc(&t->d_rec32[L], b) != hipSuccess) {
      (void)hipGetLastError();
      return fe_set_error(FE_ERR_ALLOC, "mixed: FP32 table allocation failed");
    }
    t->dev_bytes += b;
    const int64_t ne = t->n_L[L] * S32;
    const dim3 grid((unsigned)((ne + 255) / 256));
    switch (L) {
#define FE_G32(l)                                                                             \
  case l:                                                                                     \
    hipLaunchKernelGGL(k_make_g32<l>, grid, dim3(256), 0, st, t->d_rec[l], (int)t->n_L[l],   \
                       t->d_rec32[l]);                                                        \
    break;
      FE_G32(0) FE_G32(1) FE_G32(2) FE_G32(3) FE_G32(4) FE_G32(5) FE_G32(6)
#undef FE_G32
    }
  }
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

template <int L, int PPL>
static void launch_mixed_L(const fe_table* t, const FwdPlan& pl, const double* pts, int n,
                           double* partial, float tau, hipStream_t st) {
  const double* tab = t->ctx->d_boys + (size_t)L * FE_BOYS_NR * FE_BOYS_COLS;
  hipLaunchKernelGGL((k_esp_fwd_mixed<L, PPL>), dim3(pl.n_pb[L], pl.chunks[L]), dim3(FE_WG), 0, st,
                     t->d_rec[L], t->d_rec32[L], (int)t->n_L[L], pl.ppc[L], pts, n, tab, partial,
                     pl.row0[L], pl.ld, tau);
}

template <int L, int PPL>
static void launch_mixed_grad_L(const fe_table* t, const FwdPlan& pl, const double* pts, int n,
                                double* partial, size_t plane, float tau, hipStream_t st) {
  const double* tab = t->ctx->d_boys + (size_t)(L + 1) * FE_BOYS_NR * FE_BOYS_COLS;
  hipLaunchKernelGGL((k_esp_fwd_grad_mixed<L, PPL>), dim3(pl.n_pb[L], pl.chunks[L]), dim3(FE_WG),
                     0, st, t->d_rec[L], t->d_rec32[L], (int)t->n_L[L], pl.ppc[L], pts, n, tab,
                     partial, pl.row0[L], pl.ld, plane, tau);
}

int fe_launch_fwd_mixed(fe_table* t, const FwdPlan& pl, const double* pts, int n, double* partial,
                        hipStream_t st) {
  int rc = fe_ensure_g32(t, st);
  if (rc) return rc;
  const float tau = (float)t->ctx->mixed_tau;
  FE_FOR_EACH_CLASS(pl, launch_mixed_L, t, pl, pts, n, partial, tau, st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

int fe_launch_fwd_grad_mixed(fe_table* t, const FwdPlan& pl, const double* pts, int n,
                             double* partial, size_t plane, hipStream_t st) {
  int rc = fe_ensure_g32(t, st);
  if (rc) return rc;
  const float tau = (float)t->ctx->mixed_tau;
  FE_FOR_EACH_CLASS(pl, launch_mixed_grad_L, t, pl, pts, n, partial, plane, tau, st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
