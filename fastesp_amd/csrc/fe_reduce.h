// Per-point combination of the partial rows and the nuclear term, shared by
// the single-table and the batched reductions.  Rows are added in a fixed
// order with independent chains, so a result is bitwise reproducible for a
// given launch shape.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

// sum_r p[r * ld], four chains
__device__ __forceinline__ double fe_sum_rows(const double* __restrict__ p, const int n_rows,
                                              const size_t ld) {
  double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
  int r = 0;
  for (; r + 3 < n_rows; r += 4) {
    e0 += p[(size_t)r * ld];
    e1 += p[(size_t)(r + 1) * ld];
    e2 += p[(size_t)(r + 2) * ld];
    e3 += p[(size_t)(r + 3) * ld];
  }
  for (; r < n_rows; ++r) e0 += p[(size_t)r * ld];
  return (e0 + e1) + (e2 + e3);
}

// value and the three field planes (`plane` apart), two chains each
__device__ __forceinline__ void fe_sum_rows4(const double* __restrict__ p, const int n_rows,
                                             const size_t ld, const size_t plane, double& e,
                                             double& gx, double& gy, double& gz) {
  double e1 = 0.0, gx1 = 0.0, gy1 = 0.0, gz1 = 0.0;
  e = gx = gy = gz = 0.0;
  int r = 0;
  for (; r + 1 < n_rows; r += 2) {
    const size_t o = (size_t)r * ld, o1 = o + ld;
    e += p[o];
    gx += p[plane + o];
    gy += p[2 * plane + o];
    gz += p[3 * plane + o];
    e1 += p[o1];
    gx1 += p[plane + o1];
    gy1 += p[2 * plane + o1];
    gz1 += p[3 * plane + o1];
  }
  if (r < n_rows) {
    const size_t o = (size_t)r * ld;
    e += p[o];
    gx += p[plane + o];
    gy += p[2 * plane + o];
    gz += p[3 * plane + o];
  }
  e += e1;
  gx += gx1;
  gy += gy1;
  gz += gz1;
}

// Split reduction for 256-thread blocks of 64 points: wave s of the block adds
// the rows r = s (mod 4) of its 64 points (coalesced rows), the four slices
// meet in LDS and are added in the fixed order (s0 + s1) + (s2 + s3), so the
// result stays bitwise reproducible.  NP planes `plane` apart.  Every thread of
// the block must call it (it synchronises); returns the sums in the threads of
// wave 0.
template <int NP>
__device__ __forceinline__ void fe_sum_rows_split(const double* __restrict__ p, const bool valid,
                                                  const int n_rows, const size_t ld,
                                                  const size_t plane, double* out) {
  __shared__ double s_red[NP][4][64];
  const int lane = threadIdx.x & 63, s = threadIdx.x >> 6;
  double a[NP], b[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) a[q] = b[q] = 0.0;
  if (valid) {
    int r = s;
    for (; r + 4 < n_rows; r += 8) {
      const size_t o = (size_t)r * ld, o1 = o + 4 * ld;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        a[q] += p[q * plane + o];
        b[q] += p[q * plane + o1];
      }
    }
    if (r < n_rows) {
#pragma unroll
      for (int q = 0; q < NP; ++q) a[q] += p[q * plane + (size_t)r * ld];
    }
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) s_red[q][s][lane] = a[q] + b[q];
  __syncthreads();
  if (s == 0) {
#pragma unroll
    for (int q = 0; q < NP; ++q)
      out[q] = (s_red[q][0][lane] + s_red[q][1][lane]) + (s_red[q][2][lane] + s_red[q][3][lane]);
  }
}

// First stage of the reduction for few points and many partial rows (a short
// grid of point blocks would leave most of the chip idle while each block
// walks thousands of rows): block (b, g) adds rows [g rpg, (g+1) rpg) of the
// 64 points of block b, split over its waves as above, and writes row g of
// `out` (NP planes `oplane` apart).  The final stage then adds those rows in
// order, so the result stays bitwise reproducible for a given launch shape.
template <int NP>
__global__ __launch_bounds__(256) void k_rows_stage(const double* __restrict__ partial,
                                                    const int n_rows, const int ld,
                                                    const size_t plane, const int n_pts,
                                                    const int rpg, double* __restrict__ out,
                                                    const size_t oplane) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int r0 = blockIdx.y * rpg;
  const int nr = max(0, min(n_rows - r0, rpg));
  double sum[NP];
  fe_sum_rows_split<NP>(partial + (size_t)r0 * ld + i, i < n_pts, nr, ld, plane, sum);
  if (threadIdx.x < 64 && i < n_pts) {
#pragma unroll
    for (int q = 0; q < NP; ++q) out[q * oplane + (size_t)blockIdx.y * ld + i] = sum[q];
  }
}

// Row groups of the first stage for n points and n_rows partial rows (0: one
// stage suffices): enough blocks to fill the chip, at least 16 rows a group.
inline int fe_row_groups(const int n, const int n_rows) {
  const int nb = (n + 63) / 64;
  if (nb >= 128 || n_rows <= 64) return 0;
  int g = std::min((n_rows + 15) / 16, std::max(1, 2048 / nb));
  const int rpg = (n_rows + g - 1) / g;
  return (n_rows + rpg - 1) / rpg;
}

// sum_A Z_A / |C - R_A|; atoms as (x, y, z, Z)
__device__ __forceinline__ double fe_nuclear(const double x, const double y, const double z,
                                             const double* __restrict__ atoms, const int n_atoms) {
  double nuc = 0.0;
  for (int a = 0; a < n_atoms; ++a) {
    const double dx = x - atoms[4 * a], dy = y - atoms[4 * a + 1], dz = z - atoms[4 * a + 2];
    nuc += atoms[4 * a + 3] / sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
  }
  return nuc;
}

// nuclear potential and its gradient -sum_A Z_A (C - R_A) / |C - R_A|^3
__device__ __forceinline__ void fe_nuclear_grad(const double x, const double y, const double z,
                                                const double* __restrict__ atoms,
                                                const int n_atoms, double& nuc, double& nx,
                                                double& ny, double& nz) {
  nuc = nx = ny = nz = 0.0;
  for (int a = 0; a < n_atoms; ++a) {
    const double dx = x - atoms[4 * a], dy = y - atoms[4 * a + 1], dz = z - atoms[4 * a + 2];
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir = 1.0 / sqrt(r2);
    const double zr = atoms[4 * a + 3] * ir;
    nuc += zr;
    const double w = zr * ir * ir;
    nx -= w * dx;
    ny -= w * dy;
    nz -= w * dz;
  }
}
