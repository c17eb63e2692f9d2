// The per-(point tile, pair range) evaluation shared by every evaluation kernel
// (single table and batched; value or value + field; FP64 or mixed).
//
// A wavefront holds PPL points per lane and walks the pairs [p0, p1) of class
// L.  The pair record is wave-uniform (scalar loads); per point the kernel
// forms R = C - P and T = p R^2, takes the far-field path when every lane of
// the wavefront has T >= T_far(M) (M = L, or L+1 for the field), else the
// tabulated Boys function, and contracts with the generated per-L code.
//
// MIXED: the contraction runs in FP32 (pair-local R) for pairs whose screening
// estimate K*bound/max(|R|, 0.05) is below tau on every lane; R, T, Boys and
// the point sums stay FP64, the value is accumulated with Neumaier compensation.
#pragma once

#include "fe_boys.h"
#include "fe_lclass.inc"

#define FE_S32(NT) ((((NT) + 15) / 16) * 16)

__device__ __forceinline__ void fe_kahan_add(double& s, double& c, const double y) {
  const double t = s + y;
  c += (fabs(s) >= fabs(y)) ? (s - t) + y : (y - t) + s;
  s = t;
}

// R = C - P and |R|^2 of the wavefront's PPL points per lane against record r;
// returns whether every lane is past the Boys switch point of the top order
// (wave-uniform, read from a scalar so branches on it are scalar branches),
// and under MIXED whether some lane needs the FP64 contraction.
// The record values the geometry needs: P and the far threshold on |R|^2
// (T = p r^2 < T_far(M)  <=>  r^2 < T_far(M) / p, stored in the record's
// far-field block).
struct FeHdr {
  double x, y, z, thr;
};
template <int L, bool GRAD>
__device__ __forceinline__ FeHdr fe_hdr(const fe_gdptr r) {
  return FeHdr{r[0], r[1], r[2], r[FeL<L>::KOFF + (GRAD ? 1 : 0)]};
}

template <int L, int PPL, bool GRAD, bool MIXED>
__device__ __forceinline__ bool fe_pair_geometry(const fe_gdptr r, const FeHdr h,
                                                 const double* cx, const double* cy,
                                                 const double* cz, const float tau, double* Rx,
                                                 double* Ry, double* Rz, double* r2, bool& fp64) {
  const double Px = h.x, Py = h.y, Pz = h.z, r2far = h.thr;
  int near = 0, big = 0;
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    Rx[q] = cx[q] - Px;
    Ry[q] = cy[q] - Py;
    Rz[q] = cz[q] - Pz;
    r2[q] = fma(Rx[q], Rx[q], fma(Ry[q], Ry[q], Rz[q] * Rz[q]));
    near |= r2[q] < r2far;
    if constexpr (MIXED) big |= (float)r[5] * rsqrtf(fmaxf((float)r2[q], 0.0025f)) >= tau;
  }
  fp64 = true;
  if constexpr (MIXED) fp64 = __any(big);
  return __builtin_amdgcn_ballot_w64(near != 0) == 0;
}

// The points of one record by its Boys vector (tabulated, or the per-lane
// asymptotic form past the switch point; with far = true every lane takes the
// latter) and the keyed contraction.
template <int L, int PPL, bool GRAD, bool MIXED>
__device__ __forceinline__ void fe_keyed_points(const fe_gdptr r, const fe_gfptr g,
                                                const bool far, const bool fp64,
                                                const double* Rx, const double* Ry,
                                                const double* Rz, const double* r2,
                                                const double* s_tab, double* av, double* cv,
                                                double* ax, double* ay, double* az) {
  constexpr int M = GRAD ? L + 1 : L;
  const double p = r[3], K = r[4], hp = r[6], KF = r[7];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    double F[M + 1];
    double w = K;
    if (far) {
      // F_m = F_0 F'_m with K F_0 = KF / |R| folded into the weight
      const double ri = fe_rsqrt_pos(r2[q]);
      fe_boys_far_scaled<M>(hp * (ri * ri), F);
      w = KF * ri;
    } else {
      fe_boys<M>(p * r2[q], F, s_tab);
    }
    if constexpr (GRAD && L <= FE_GRAD_ACC_LMAX && !MIXED) {
      // the weight folded into F before the contraction: fewer operations
      // than weighting its four results for the light classes
      fe_contract_grad_acc<L>(Rx[q], Ry[q], Rz[q], F, r + 8, p, w, av[q], ax[q], ay[q], az[q]);
    } else if constexpr (GRAD) {
      double I, Gx, Gy, Gz;
      if (fp64)
        fe_contract_grad<L>(Rx[q], Ry[q], Rz[q], F, r + 8, p, I, Gx, Gy, Gz);
      else
        fe_contract_grad_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g, p, Rx[q], Ry[q],
                                Rz[q], I, Gx, Gy, Gz);
      if constexpr (MIXED)
        fe_kahan_add(av[q], cv[q], w * I);
      else
        av[q] = fma(w, I, av[q]);
      ax[q] = fma(w, Gx, ax[q]);
      ay[q] = fma(w, Gy, ay[q]);
      az[q] = fma(w, Gz, az[q]);
    } else {
      double s;
      if (fp64)
        s = fe_contract<L>(Rx[q], Ry[q], Rz[q], F, r + 8);
      else
        s = fe_contract_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g);
      if constexpr (MIXED)
        fe_kahan_add(av[q], cv[q], w * s);
      else
        av[q] = fma(w, s, av[q]);
    }
  }
}

// The points of one record as the factors of their accumulation,
//   value += A B,   field += C (Dx, Dy, Dz),
// so that the far and the near paths below end in one shared accumulation
// (with each path adding into the running sums itself, the compiler kept the
// sums in other registers in the two paths and copied them at every join).
//
// Far (every Boys order asymptotic on every lane): the pair's integral is
// |u| H(u), u = R/|R|^2, from the record's far-field block, and its gradient
// |u|^3 (grad H - (H + 2 u.grad H) R).  FP64 in every mode.
template <int L, int PPL, bool GRAD>
__device__ __forceinline__ void fe_far_terms(const fe_gdptr r, const double* Rx,
                                             const double* Ry, const double* Rz,
                                             const double* r2, double* A, double* B, double* C,
                                             double* Dx, double* Dy, double* Dz) {
  const fe_gdptr kel = r + FeL<L>::KCOEF;
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const double ri = fe_rsqrt_pos(r2[q]);
    const double ri2 = ri * ri;
    A[q] = ri;
    if constexpr (L == 0) {
      // class 0 (see fe_eval_tile): value A = h_0/|R|, field C R with
      // C = -h_0/|R|^3
      A[q] = kel[0] * ri;
      B[q] = 1.0;
      if constexpr (GRAD) {
        C[q] = -kel[0] * (ri * ri2);
        Dx[q] = Rx[q];
        Dy[q] = Ry[q];
        Dz[q] = Rz[q];
      }
    } else if constexpr (!GRAD) {
      B[q] = fe_far_value<L>(Rx[q] * ri2, Ry[q] * ri2, Rz[q] * ri2, kel);
    } else {
      double H, Hx, Hy, Hz, D;
      fe_far_grad<L>(Rx[q] * ri2, Ry[q] * ri2, Rz[q] * ri2, kel, H, Hx, Hy, Hz, D);
      const double t = fma(2.0, D, H);
      B[q] = H;
      C[q] = ri * ri2;
      Dx[q] = fma(-t, Rx[q], Hx);
      Dy[q] = fma(-t, Ry[q], Hy);
      Dz[q] = fma(-t, Rz[q], Hz);
    }
  }
}

// Near (some lane short of the switch point): the Boys vector per lane
// (tabulated, or the asymptotic form for the lanes past it) and the keyed
// contraction, FP64 or (MIXED, fp64 false) FP32; A = C = K.
template <int L, int PPL, bool GRAD, bool MIXED>
__device__ __forceinline__ void fe_near_terms(const fe_gdptr r, const fe_gfptr g, const bool fp64,
                                              const double* Rx, const double* Ry,
                                              const double* Rz, const double* r2,
                                              const double* s_tab, double* A, double* B,
                                              double* C, double* Dx, double* Dy, double* Dz) {
  constexpr int M = GRAD ? L + 1 : L;
  const double p = r[3], K = r[4];
  if constexpr (L == 0 && !MIXED) {
    // class 0: I = g_0 F_0 and grad I = -2p g_0 F_1 R, so value A = K g_0 F_0
    // and field C R with C = -2p K g_0 F_1
    const double kg = K * r[8];
    const double kg2 = -2.0 * p * kg;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      double F[M + 1];
      fe_boys<M>(p * r2[q], F, s_tab);
      A[q] = kg * F[0];
      if constexpr (GRAD) C[q] = kg2 * F[1];
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    double F[M + 1];
    fe_boys<M>(p * r2[q], F, s_tab);
    A[q] = K;
    if constexpr (GRAD) {
      C[q] = K;
      if (!MIXED || fp64)
        fe_contract_grad<L>(Rx[q], Ry[q], Rz[q], F, r + 8, p, B[q], Dx[q], Dy[q], Dz[q]);
      else
        fe_contract_grad_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g, p, Rx[q], Ry[q],
                                Rz[q], B[q], Dx[q], Dy[q], Dz[q]);
    } else {
      if (!MIXED || fp64)
        B[q] = fe_contract<L>(Rx[q], Ry[q], Rz[q], F, r + 8);
      else
        B[q] = fe_contract_f32<L>((float)Rx[q], (float)Ry[q], (float)Rz[q], F, g);
    }
  }
}

template <int L, int PPL, bool GRAD, bool MIXED>
__device__ __forceinline__ void fe_eval_tile(fe_gdptr __restrict__ rec,
                                             fe_gfptr __restrict__ g32, const int p0,
                                             const int p1, const double* cx, const double* cy,
                                             const double* cz, const double* s_tab,
                                             const float tau, double* av, double* cv, double* ax,
                                             double* ay, double* az) {
  constexpr int S = FeL<L>::STRIDE, S32 = FE_S32(FeL<L>::NT + FeL<L>::ND);
  double Rx[PPL], Ry[PPL], Rz[PPL], r2[PPL];
  bool fp64;
  if constexpr (L >= FE_KELVIN_LMIN) {
    // the next record's header is read one trip ahead, so its scalar loads
    // are in flight while this record is evaluated (classes in
    // FE_HDR_PREFETCH; elsewhere the extra live SGPRs cost more spills than
    // the latency it hides)
    constexpr bool PF = (FE_HDR_PREFETCH >> L) & 1;
    FeHdr h = fe_hdr<L, GRAD>(rec + (size_t)min(p0, max(p1 - 1, 0)) * S);
    for (int ip = p0; ip < p1; ++ip) {
      const fe_gdptr r = rec + (size_t)ip * S;
      if constexpr (!PF) h = fe_hdr<L, GRAD>(r);
      const FeHdr hn = PF ? fe_hdr<L, GRAD>(rec + (size_t)min(ip + 1, p1 - 1) * S) : h;
      double A[PPL], B[PPL], C[PPL], Dx[PPL], Dy[PPL], Dz[PPL];
      const bool far =
          fe_pair_geometry<L, PPL, GRAD, MIXED>(r, h, cx, cy, cz, tau, Rx, Ry, Rz, r2, fp64);
      h = hn;
      if constexpr (L > FE_TAIL_LMAX) {
        // the heavy classes: each path adds its products into contributions
        // that start from zero, so only those four are live at the join
        // (fewer registers than the six factors), at one more operation per sum
        double dv[PPL], dc[PPL], dx[PPL], dy[PPL], dz[PPL];
#pragma unroll
        for (int q = 0; q < PPL; ++q) dv[q] = dc[q] = dx[q] = dy[q] = dz[q] = 0.0;
        if (far) {
          fe_far_terms<L, PPL, GRAD>(r, Rx, Ry, Rz, r2, A, B, C, Dx, Dy, Dz);
#pragma unroll
          for (int q = 0; q < PPL; ++q) {
            dv[q] = A[q] * B[q];
            if constexpr (GRAD) {
              dx[q] = C[q] * Dx[q];
              dy[q] = C[q] * Dy[q];
              dz[q] = C[q] * Dz[q];
            }
          }
        } else {
          fe_keyed_points<L, PPL, GRAD, MIXED>(r, g32 + (size_t)ip * S32, false, fp64, Rx, Ry,
                                               Rz, r2, s_tab, dv, dc, dx, dy, dz);
        }
#pragma unroll
        for (int q = 0; q < PPL; ++q) {
          if constexpr (MIXED) {
            fe_kahan_add(av[q], cv[q], dv[q]);
            cv[q] += dc[q];
          } else {
            av[q] += dv[q];
          }
          if constexpr (GRAD) {
            ax[q] += dx[q];
            ay[q] += dy[q];
            az[q] += dz[q];
          }
        }
        continue;
      }
      if (far)
        fe_far_terms<L, PPL, GRAD>(r, Rx, Ry, Rz, r2, A, B, C, Dx, Dy, Dz);
      else
        fe_near_terms<L, PPL, GRAD, MIXED>(r, g32 + (size_t)ip * S32, fp64, Rx, Ry, Rz, r2,
                                           s_tab, A, B, C, Dx, Dy, Dz);
#pragma unroll
      for (int q = 0; q < PPL; ++q) {
        if constexpr (L == 0 && !MIXED) {
          // class 0: the value is A itself and the field C R in both paths
          av[q] += A[q];
          if constexpr (GRAD) {
            ax[q] = fma(C[q], Rx[q], ax[q]);
            ay[q] = fma(C[q], Ry[q], ay[q]);
            az[q] = fma(C[q], Rz[q], az[q]);
          }
          continue;
        }
        if constexpr (MIXED)
          fe_kahan_add(av[q], cv[q], A[q] * B[q]);
        else
          av[q] = fma(A[q], B[q], av[q]);
        if constexpr (GRAD) {
          ax[q] = fma(C[q], Dx[q], ax[q]);
          ay[q] = fma(C[q], Dy[q], ay[q]);
          az[q] = fma(C[q], Dz[q], az[q]);
        }
      }
    }
  } else {
    for (int ip = p0; ip < p1; ++ip) {
      const fe_gdptr r = rec + (size_t)ip * S;
      const bool far = fe_pair_geometry<L, PPL, GRAD, MIXED>(r, fe_hdr<L, GRAD>(r), cx, cy, cz,
                                                             tau, Rx, Ry, Rz, r2, fp64);
      fe_keyed_points<L, PPL, GRAD, MIXED>(r, g32 + (size_t)ip * S32, far, fp64, Rx, Ry, Rz, r2,
                                           s_tab, av, cv, ax, ay, az);
    }
  }
}

// Load PPL points per lane of the block starting at `base` (clamped), run the
// tile, and write the partial sums (value plane; three more planes with GRAD)
// for the points below `end`, or add them to the sums already there (`add`).
template <int L, int PPL, bool GRAD, bool MIXED>
__device__ __forceinline__ void fe_eval_block(fe_gdptr __restrict__ rec,
                                              fe_gfptr __restrict__ g32, const int p0,
                                              const int p1, const double* __restrict__ pts,
                                              const int base, const int end,
                                              const double* s_tab, const float tau,
                                              double* __restrict__ out, const size_t plane,
                                              const bool add) {
  double cx[PPL], cy[PPL], cz[PPL], av[PPL], cv[PPL], ax[PPL], ay[PPL], az[PPL];
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = min(base + q * FE_WG + (int)threadIdx.x, end - 1);
    cx[q] = pts[3 * (size_t)i];
    cy[q] = pts[3 * (size_t)i + 1];
    cz[q] = pts[3 * (size_t)i + 2];
    av[q] = cv[q] = ax[q] = ay[q] = az[q] = 0.0;
  }
  fe_eval_tile<L, PPL, GRAD, MIXED>(rec, g32, p0, p1, cx, cy, cz, s_tab, tau, av, cv, ax, ay, az);
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int o = q * FE_WG + (int)threadIdx.x;
    if (base + o < end) {
      const double v = av[q] + cv[q];
      out[o] = add ? out[o] + v : v;
      if constexpr (GRAD) {
        out[plane + o] = add ? out[plane + o] + ax[q] : ax[q];
        out[2 * plane + o] = add ? out[2 * plane + o] + ay[q] : ay[q];
        out[3 * plane + o] = add ? out[3 * plane + o] + az[q] : az[q];
      }
    }
  }
}
