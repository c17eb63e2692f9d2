// Boys functions F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt on the device.
//
// Scheme (standard; Helgaker, Jorgensen & Olsen, "Molecular Electronic-Structure
// Theory", sec. 9.8.2; Gill, Johnson & Pople, IJQC 40, 745 (1991)):
//  * T < fe_tfar(M): nearest grid point T_r = r * dT (dT = 1/8), a degree-7
//    polynomial in dt = T - T_r (|dt| <= 1/16) for the top order F_M, exp(-T)
//    = exp(-T_r) * a degree-7 polynomial of -dt, then the stable downward
//    recursion F_{m-1} = (2T F_m + exp(-T)) / (2m-1).  Both polynomials are
//    the degree-9 Taylor polynomials economized to degree 7 over |dt| <= 1/16
//    (Chebyshev economization: the dt^9 and dt^8 terms are rewritten with
//    T_9 and T_8, whose remainders h^9/256 and h^8/128 are dropped), so each
//    is within 5e-17 relative of its function on the interval.
//  * T >= fe_tfar(M) (35.5 .. 59 for top order M = 0 .. 8): asymptotic
//    F_0 = sqrt(pi/T)/2 and the upward recursion F_m = F_{m-1} (2m-1)/(2T);
//    the neglected incomplete-gamma tail is below 2^-54 relative for m <= M.
//    The grid covers T < FE_BOYS_TSW = 60 for every order.
// The reference evaluator switches at T = M - 0.5 between an 80-term Taylor
// series at T = 0 and erf + upward recursion (fastesp/native/esp_gpu.cu:36-66);
// that series cancels just below the switch (2.4e-13 relative at M = 7), which
// the tabulated expansion avoids.
//
// The grid values F_j(T_r) are computed on the host in 80-bit long double from
// the all-positive series F_j(T) = exp(-T) sum_i (2T)^i / prod_{k<=i} (2j+2k+1),
// and the economization is done there in long double.
#pragma once

#define FE_BOYS_NR 481        // grid rows, T_r = r/8, r = 0..480 (T < 60)
#define FE_BOYS_COLS 9        // 8 polynomial coefficients + exp(-T_r)
#define FE_BOYS_NK 8          // polynomial coefficients per row (degree 7)
#define FE_BOYS_NTAYLOR 10    // Taylor terms the economization starts from
#define FE_BOYS_MMAX 8        // highest top order with a table
#define FE_BOYS_TSW 60.0
#define FE_BOYS_DT 0.125
#define FE_BOYS_INVDT 8.0

// exp(x) for |x| <= 1/16: the degree-9 Taylor polynomial economized to degree
// 7 (see above); max relative error 4.8e-17 on the interval (exact rational
// coefficients rounded to double).
__device__ __forceinline__ double fe_exp_small(const double x) {
  double e = 1.9843691871279762e-04;
  e = fma(e, x, 1.3890826512896826e-03);
  e = fma(e, x, 8.3333332623754235e-03);
  e = fma(e, x, 4.1666666193613928e-02);
  e = fma(e, x, 1.6666666666674365e-01);
  e = fma(e, x, 5.0000000000036959e-01);
  e = fma(e, x, 1.0);
  return fma(e, x, 1.0);
}

// Far field.  For T >= fe_tfar(M) the incomplete-gamma tail is below 2^-54 of
// F_m for every m <= M (thresholds from mpmath at 40 digits), so
//   F_m(T) = F_0(T) prod_{j=1..m} (2j-1)/(2T),   F_0(T) = sqrt(pi/T)/2
// exactly to double precision.  A pair of class L then contributes
//   K F_0 sum_m F'_m a_m(R),  F'_m = prod_{j<=m} (2j-1) x,  x = 1/(2 p R^2),
// with K F_0 = K sqrt(pi/p)/2 * |R|^-1: no table, no exp, no division per point
// (1/(2p) and K sqrt(pi/p)/2 are stored per pair).
__host__ __device__ constexpr double fe_tfar(int M) {
  return M <= 0 ? 35.5 : M == 1 ? 39.5 : M == 2 ? 43.0 : M == 3 ? 46.0 : M == 4 ? 49.0
       : M == 5 ? 51.5 : M == 6 ? 54.0 : M == 7 ? 56.5 : 59.0;
}

// Grid rows a kernel of top order M reads: T < fe_tfar(M) takes the table
// (nearest row r = floor(8 T + 1/2) <= 8 fe_tfar(M)), larger T the asymptotic
// form, so only these rows are staged in LDS.
__host__ __device__ constexpr int fe_boys_rows(int M) {
  return (int)(fe_tfar(M) * FE_BOYS_INVDT) + 2 < FE_BOYS_NR ? (int)(fe_tfar(M) * FE_BOYS_INVDT) + 2
                                                             : FE_BOYS_NR;
}

// 1/sqrt(x) for finite x > 0 (callers' arguments are bounded away from zero,
// infinity and denormals): the hardware estimate refined by the third-order
// Newton step y + y e (1/2 + 3/8 e), e = 1 - x y^2 -- the same step as the
// math library's rsqrt, without its selects for special operands.
__device__ __forceinline__ double fe_rsqrt_pos(const double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(0.375, e, 0.5), y);
}

template <int M>
__device__ __forceinline__ void fe_boys_far_scaled(const double x, double* F) {
  F[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= M; ++m) F[m] = F[m - 1] * ((2 * m - 1) * x);
}

// F[0..M] at T; `tab` holds at least the first fe_boys_rows(M) rows of the
// table of top order M (normally staged in LDS).  The asymptotic form is used
// from fe_tfar(M) on, where it is exact to double precision.
// The tabulated form alone, for T < fe_tfar(M) (rounding may put T a few ulp
// above it: the rows staged reach one row past the threshold).
template <int M>
__device__ __forceinline__ void fe_boys_tab(const double T, double* F, const double* tab) {
  const int r = (int)fma(T, FE_BOYS_INVDT, 0.5);
  const double dt = fma(-FE_BOYS_DT, (double)r, T);
  const double* c = tab + r * FE_BOYS_COLS;
  double f = c[7];
  f = fma(f, dt, c[6]);
  f = fma(f, dt, c[5]);
  f = fma(f, dt, c[4]);
  f = fma(f, dt, c[3]);
  f = fma(f, dt, c[2]);
  f = fma(f, dt, c[1]);
  f = fma(f, dt, c[0]);
  F[M] = f;
  if constexpr (M > 0) {
    const double e = c[8] * fe_exp_small(-dt);
    const double t2 = T + T;
#pragma unroll
    for (int m = M; m > 0; --m) F[m - 1] = fma(t2, F[m], e) * (1.0 / (2 * m - 1));
  }
}

template <int M>
__device__ __forceinline__ void fe_boys(const double T, double* F, const double* tab) {
  if (T < fe_tfar(M)) {
    fe_boys_tab<M>(T, F, tab);
  } else {
    const double r = fe_rsqrt_pos(T);
    F[0] = 0.88622692545275801365 * r;  // sqrt(pi)/2
    if constexpr (M > 0) {
      // 1/(2T) from the reciprocal square root already at hand (a few ulp;
      // a division would cost a dozen instructions on every lane of the branch)
      const double x = 0.5 * (r * r);
#pragma unroll
      for (int m = 1; m <= M; ++m) F[m] = F[m - 1] * ((2 * m - 1) * x);
    }
  }
}
