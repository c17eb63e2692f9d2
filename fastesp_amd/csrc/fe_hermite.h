// 1-D Hermite expansion coefficients, shared by the table build and the
// density adjoint.  E[(i*(lj+1)+j)*(L+1)+t] for (u+xa)^i (u+xb)^j e^{-p u^2},
// i <= li, j <= lj, t <= L = li + lj, by
//   E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + xa E^{ij}_t + (t+1) E^{ij}_{t+1}
// (fastesp/ints.py:62-81; fastesp/native/build_pairs.cpp:250-273).
#pragma once

__device__ __forceinline__ void fe_e1d_table(double* E, const int li, const int lj,
                                             const double xa, const double xb, const double p) {
  const int L = li + lj, T = L + 1;
  const double h = 0.5 / p;
  const int nE = (li + 1) * (lj + 1) * T;
  for (int q = 0; q < nE; ++q) E[q] = 0.0;
  E[0] = 1.0;
  for (int i = 0; i <= li; ++i)
    for (int j = 0; j <= lj; ++j) {
      if (i == 0 && j == 0) continue;
      const double* src = (i > 0) ? &E[((i - 1) * (lj + 1) + j) * T] : &E[(i * (lj + 1) + j - 1) * T];
      const double x = (i > 0) ? xa : xb;
      double* dst = &E[(i * (lj + 1) + j) * T];
      for (int t = 0; t <= L; ++t) {
        double v = x * src[t];
        if (t >= 1) v += src[t - 1] * h;
        if (t + 1 <= L) v += (t + 1) * src[t + 1];
        dst[t] = v;
      }
    }
}
