// Adjoint device code shared by the single-table and the batched backward.
//
// fe_dg_sum<L> / fe_dg_block<L>: one lane per pair of class L; the points
//   (x, y, z, gV) of [q0, q1) are wave-uniform scalar loads, so there is no
//   cross-lane reduction.  fe_dg_block writes
//   out[k * stride + pair] = -sum_i gV_i K R_i^a F_m(T_i).
// fe_det_from_dg<L> (generated) and fe_dp_task<L>: dL/dG -> dL/dP,
//   d etab_t = sum_k dG_k (2 pi/p)(2p)^m M0[k, t] per kept record, then per
//   entry (mu, nu) of a shell pair d blk = sum over its primitive pairs of
//   c_i c_j sum_t d etab_t Ex Ey Ez; both triangles of P receive d blk because
//   the block is folded P_ij + P_ji^T (fastesp/esp.py:185-187).
#pragma once

#include "fe_internal.h"
#include "fe_boys.h"
#include "fe_lclass.inc"

struct DPArgs {
  const int* sh_l; const int* sh_fn0; const int* sh_nfn; const int* sh_prim0; const int* sh_nprim;
  const double* sh_c; const double* pexp; const double* pcoef;
  const int* fn_lx; const int* fn_ly; const int* fn_lz;
  const int* sp_i; const int* sp_j; const int* sp_slot;
  const int* sp_by_class;  // shell pairs grouped by class: [sp_cls_off[L], sp_cls_off[L+1])
  int sp_cls_off[FE_LMAX + 2];
  const uint32_t* tasks;  // dL/dP tasks grouped by class: [task_cls_off[L], task_cls_off[L+1])
  int task_cls_off[FE_LMAX + 2];
  const double* det[FE_LMAX + 1];  // d etab per kept record (FeL<L>::NM each), fe_det_from_dg
  const int* slot2kept[FE_LMAX + 1];
  const double* dG[FE_LMAX + 1];
  int n_func;
  double* dP;
};

// acc[k] = sum_{i in [q0, q1)} gV_i K R_i^a F_m(T_i) for the lane's pair (the
// pair index is clamped to the last pair; an empty range gives zeros)
template <int L>
__device__ __forceinline__ void fe_dg_sum(const double* __restrict__ rec, const int n_pairs,
                                          const int pair, const double* __restrict__ pw,
                                          const int q0, const int q1, const double* s_tab,
                                          double* acc) {
  constexpr int S = FeL<L>::STRIDE, NT = FeL<L>::NT;
  const int ip = min(pair, n_pairs - 1);
  const double* r = rec + (size_t)ip * S;
  const double Px = r[0], Py = r[1], Pz = r[2], p = r[3], K = r[4];
  const double hp = r[6], KF = r[7];
  // Per-order weights without the pair's constants.  In the far field
  //   K F_m = C_m |R|^-(2m+1),  C_m = K sqrt(pi/p)/2 (2m-1)!! (1/(2p))^m,
  // so the far weights are w |R|^-(2m+1) (one product per order) and the
  // tabulated ones w K F_m / C_m; the sums are scaled by C_m once at the end.
  // The per-lane factors 1/C_m cost L + 1 live registers, so the class whose
  // kernel runs under a register cap (two waves per SIMD) keeps the weights
  // w K F_m with K F_0 folded in (no final scale).
  constexpr bool SPLIT = fe_dg_minw(L) == 1;
  // T = p r^2 >= T_far(L)  <=>  r^2 >= T_far(L) / p: one compare per point
  // (the product p r^2 is formed only on the tabulated path).  The class
  // under a register cap tests p r^2 itself: one live value fewer.
  const double r2far = fe_tfar(L) * (hp + hp);
  double D[L + 1];
  if constexpr (SPLIT) {
    double c = KF;
#pragma unroll
    for (int m = 0; m <= L; ++m) {
      D[m] = c != 0.0 ? K / c : 0.0;
      c *= (2 * m + 1) * hp;
    }
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) acc[k] = 0.0;
  // far-field weights of one point (lanes with T >= T_far)
  auto far_weights = [&](const double r2, const double w, double* wm) {
    const double ri = fe_rsqrt_pos(r2);
    if constexpr (SPLIT) {
      const double u = ri * ri;
      wm[0] = w * ri;
#pragma unroll
      for (int m = 1; m <= L; ++m) wm[m] = wm[m - 1] * u;
    } else {
      double F[L + 1];
      fe_boys_far_scaled<L>(hp * (ri * ri), F);
      const double wf = w * (KF * ri);
#pragma unroll
      for (int m = 0; m <= L; ++m) wm[m] = wf * F[m];
    }
  };
  auto point = [&](const double x, const double y, const double z, const double w) {
    const double Rx = x - Px, Ry = y - Py, Rz = z - Pz;
    const double r2 = fma(Rx, Rx, fma(Ry, Ry, Rz * Rz));
    double wm[L + 1];
    const bool near = SPLIT ? r2 < r2far : p * r2 < fe_tfar(L);
    if (!__any(near)) {
      far_weights(r2, w, wm);
    } else if constexpr (SPLIT) {
      // a tabulated wavefront: its near lanes take the table, the others
      // the same far-field weights as an all-far wavefront
      if (near) {
        double F[L + 1];
        fe_boys_tab<L>(p * r2, F, s_tab);
#pragma unroll
        for (int m = 0; m <= L; ++m) wm[m] = (w * F[m]) * D[m];
      } else {
        far_weights(r2, w, wm);
      }
    } else {
      // the class under a register cap: one Boys evaluation for all lanes
      // (table or asymptotic form per lane) keeps fewer values live
      double F[L + 1];
      fe_boys<L>(p * r2, F, s_tab);
#pragma unroll
      for (int m = 0; m <= L; ++m) wm[m] = (w * K) * F[m];
    }
    fe_dg_accum<L>(Rx, Ry, Rz, wm, acc);
  };
  // two wave-uniform points per iteration so their scalar loads overlap; the
  // odd tail point repeats the last point with zero weight
  for (int q = q0; q < q1; q += 2) {
    const double* __restrict__ c0 = pw + 4 * (size_t)q;
    const double* __restrict__ c1 = pw + 4 * (size_t)min(q + 1, q1 - 1);
    const double x0 = c0[0], y0 = c0[1], z0 = c0[2], w0 = c0[3];
    const double x1 = c1[0], y1 = c1[1], z1 = c1[2];
    const double w1 = (q + 1 < q1) ? c1[3] : 0.0;
    point(x0, y0, z0, w0);
    point(x1, y1, z1, w1);
  }
  if constexpr (SPLIT) {
    // C_m from the record again (not kept live through the point loop)
    double C[L + 1];
    C[0] = r[7];
#pragma unroll
    for (int m = 1; m <= L; ++m) C[m] = C[m - 1] * ((2 * m - 1) * r[6]);
    fe_dg_scale<L>(acc, C);
  }
}

// one lane per pair: writes out[k * stride + pair] = -sum over [q0, q1)
template <int L>
__device__ __forceinline__ void fe_dg_block(const double* __restrict__ rec, const int n_pairs,
                                            const int pair, const double* __restrict__ pw,
                                            const int q0, const int q1, const double* s_tab,
                                            double* __restrict__ out, const int stride) {
  constexpr int NT = FeL<L>::NT;
  double acc[NT];
  fe_dg_sum<L>(rec, n_pairs, pair, pw, q0, q1, s_tab, acc);
  if (pair < n_pairs) {
#pragma unroll
    for (int k = 0; k < NT; ++k) out[(size_t)k * stride + pair] = -acc[k];
  }
}

// 1-D Hermite coefficients E^{ab}_t, t = 0..L (zero above a + b), of
// (u+xa)^a (u+xb)^b e^{-p u^2} for a, b <= 3, a + b <= L, by the recurrence
// E^{i+1,j}_t = E^{ij}_{t-1} h + xa E^{ij}_t + (t+1) E^{ij}_{t+1} (fastesp/ints.py:62-81)
// raising a, then b; h = 1/(2p).  Compile-time bounds and predicated steps
// keep e in registers.
template <int L>
__device__ __forceinline__ void fe_e1d_ab(const int a, const int b, const double xa,
                                          const double xb, const double h, double* e) {
  constexpr int LI = L < 3 ? L : 3;
  e[0] = 1.0;
#pragma unroll
  for (int t = 1; t <= L; ++t) e[t] = 0.0;
#pragma unroll
  for (int s = 0; s < 2 * LI; ++s) {
    const bool on = s < LI ? s < a : s - LI < b;
    const double x = s < LI ? xa : xb;
    if (on) {
      double n[L + 1];
#pragma unroll
      for (int t = 0; t <= L; ++t) {
        double v = x * e[t];
        if (t >= 1) v = fma(h, e[t - 1], v);
        if (t + 1 <= L) v = fma((double)(t + 1), e[t + 1], v);
        n[t] = v;
      }
#pragma unroll
      for (int t = 0; t <= L; ++t) e[t] = n[t];
    }
  }
}

// One dL/dP task of class L: entry (mi, mj) of shell pair sp,
//   sum over the kept primitive pairs of c_i c_j sum_t d etab_t Ex Ey Ez,
// one lane each (no barriers: most shell pairs have one or two primitive
// pairs, so the work is the latency of a few gathers).  Both triangles of P
// receive it (the table folds P_ij + P_ji^T, fastesp/esp.py:185-187).
template <int L>
__device__ __forceinline__ void fe_dp_task(const DPArgs& a, const uint32_t task) {
  constexpr int NM = FeL<L>::NM;
  const int sp = (int)(task >> 8), mi = (int)(task >> 4) & 15, mj = (int)task & 15;
  const int si = a.sp_i[sp], sj = a.sp_j[sp];
  const int fi = a.sh_fn0[si] + mi, fj = a.sh_fn0[sj] + mj;
  const int npi = a.sh_nprim[si], npj = a.sh_nprim[sj];
  const int pi0 = a.sh_prim0[si], pj0 = a.sh_prim0[sj];
  const int slot0 = a.sp_slot[sp];
  const int ax = a.fn_lx[fi], ay = a.fn_ly[fi], az = a.fn_lz[fi];
  const int bx = a.fn_lx[fj], by = a.fn_ly[fj], bz = a.fn_lz[fj];
  const double Ax = a.sh_c[3 * si], Ay = a.sh_c[3 * si + 1], Az = a.sh_c[3 * si + 2];
  const double Bx = a.sh_c[3 * sj], By = a.sh_c[3 * sj + 1], Bz = a.sh_c[3 * sj + 2];
  const int* __restrict__ s2k = a.slot2kept[L];
  const double* __restrict__ det = a.det[L];
  double acc = 0.0;
  for (int ii = 0; ii < npi; ++ii)
    for (int jj = 0; jj < npj; ++jj) {
      const int kept = s2k[slot0 + ii * npj + jj];
      if (kept < 0) continue;
      const double al = a.pexp[pi0 + ii], be = a.pexp[pj0 + jj];
      const double p = al + be, ip = 1.0 / p, h = 0.5 * ip;
      const double Px = (al * Ax + be * Bx) * ip;
      const double Py = (al * Ay + be * By) * ip;
      const double Pz = (al * Az + be * Bz) * ip;
      double ex[L + 1], ey[L + 1], ez[L + 1];
      fe_e1d_ab<L>(ax, bx, Px - Ax, Px - Bx, h, ex);
      fe_e1d_ab<L>(ay, by, Py - Ay, Py - By, h, ey);
      fe_e1d_ab<L>(az, bz, Pz - Az, Pz - Bz, h, ez);
      const double* __restrict__ dt = det + (size_t)kept * NM;
      // the |t| <= L simplex in its sorted (tx, ty, tz) order
      double s = 0.0;
      int q = 0;
#pragma unroll
      for (int tx = 0; tx <= L; ++tx) {
        double sy = 0.0;
#pragma unroll
        for (int ty = 0; ty <= L - tx; ++ty) {
          double sz = 0.0;
#pragma unroll
          for (int tz = 0; tz <= L - tx - ty; ++tz) sz = fma(ez[tz], dt[q++], sz);
          sy = fma(ey[ty], sz, sy);
        }
        s = fma(ex[tx], sy, s);
      }
      acc = fma(a.pcoef[pi0 + ii] * a.pcoef[pj0 + jj], s, acc);
    }
  a.dP[(size_t)fi * a.n_func + fj] = acc;
  if (si != sj) a.dP[(size_t)fj * a.n_func + fi] = acc;
}

// DPArgs of a table whose dG of class L starts at dG + dg_off[L]
inline DPArgs fe_dp_args(const fe_table* t, double* dG, const size_t* dg_off, double* dP) {
  DPArgs pa{};
  pa.sh_l = t->b.l; pa.sh_fn0 = t->b.fn0; pa.sh_nfn = t->b.nfn; pa.sh_prim0 = t->b.prim0;
  pa.sh_nprim = t->b.nprim; pa.sh_c = t->b.c; pa.pexp = t->b.pexp; pa.pcoef = t->b.pcoef;
  pa.fn_lx = t->b.fn_lx; pa.fn_ly = t->b.fn_ly; pa.fn_lz = t->b.fn_lz;
  pa.sp_i = t->b.sp_i; pa.sp_j = t->b.sp_j; pa.sp_slot = t->b.sp_slot;
  pa.sp_by_class = t->b.sp_by_class;
  for (int L = 0; L <= FE_LMAX + 1; ++L) pa.sp_cls_off[L] = t->sp_cls_off[L];
  pa.tasks = t->b.tasks;
  for (int L = 0; L <= FE_LMAX + 1; ++L) pa.task_cls_off[L] = t->task_cls_off[L];
  for (int L = 0; L <= FE_LMAX; ++L) {
    pa.slot2kept[L] = t->d_slot2kept[L];
    pa.dG[L] = dG + dg_off[L];
  }
  pa.n_func = t->n_func;
  pa.dP = dP;
  return pa;
}
