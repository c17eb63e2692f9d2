// Spatial ordering of the ESP points (Morton / Z-order).
//
// The evaluation kernels give each wavefront 64 consecutive points.  Surface
// points arrive in arbitrary order, so a wavefront would otherwise see points
// from all over the molecule: for a given pair some lanes would sit in the
// near field (T < 60, tabulated Boys) and others in the far field (asymptotic
// Boys), and both branches would execute.  Sorting the points along a 30-bit
// Morton curve makes every wavefront a compact surface patch, so the branch is
// nearly always uniform.  Results are scattered back to the caller's order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "fe_internal.h"

__device__ __forceinline__ uint32_t fe_spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(256) void k_morton(const double* __restrict__ pts, const int n,
                                                const double ox, const double oy, const double oz,
                                                const double scale, uint32_t* __restrict__ key,
                                                int32_t* __restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t q[3];
  const double o[3] = {ox, oy, oz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double v = (pts[3 * i + a] - o[a]) * scale;
    v = fmin(fmax(v, 0.0), 1023.0);
    q[a] = (uint32_t)v;
  }
  key[i] = fe_spread10(q[0]) | (fe_spread10(q[1]) << 1) | (fe_spread10(q[2]) << 2);
  idx[i] = i;
}

__global__ __launch_bounds__(256) void k_gather_pts(const double* __restrict__ pts,
                                                    const int32_t* __restrict__ perm, const int n,
                                                    double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  out[3 * i] = pts[3 * j];
  out[3 * i + 1] = pts[3 * j + 1];
  out[3 * i + 2] = pts[3 * j + 2];
}

// Sorted copy of the points and the permutation (sorted position -> caller
// index), both in the context's FE_WS_SORT slot.  The quantisation frame is the
// nuclei's bounding box grown by 16 Bohr (points outside are clamped, which
// only costs locality, never correctness).
int fe_sort_points(fe_ctx* c, const fe_table* t, const double* d_pts, int n, hipStream_t st,
                   const double** d_sorted, const int32_t** d_perm, void* buf, size_t buf_bytes) {
  double lo[3] = {-16.0, -16.0, -16.0}, hi[3] = {16.0, 16.0, 16.0};
  if (t->n_atoms > 0) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = INFINITY;
      hi[a] = -INFINITY;
    }
    for (int i = 0; i < t->n_atoms; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], t->h_atoms[4 * i + a]);
        hi[a] = std::max(hi[a], t->h_atoms[4 * i + a]);
      }
    for (int a = 0; a < 3; ++a) {
      lo[a] -= 16.0;
      hi[a] += 16.0;
    }
  }
  const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  const double scale = 1023.0 / ext;
  size_t temp_bytes = 0;
  FE_HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (uint32_t*)nullptr,
                                               (uint32_t*)nullptr, (int32_t*)nullptr,
                                               (int32_t*)nullptr, n, 0, 30, st));
  const size_t al = 256;
  auto up = [&](size_t b) { return (b + al - 1) / al * al; };
  const size_t b_key = up((size_t)n * 4), b_pts = up((size_t)n * 3 * sizeof(double));
  const size_t need = 4 * b_key + b_pts + up(temp_bytes);
  char* base = (char*)buf;
  if (buf) {
    if (buf_bytes < need)
      return fe_set_error(FE_ERR_ARG, "point sort: %zu bytes of scratch needed, %zu given", need,
                          buf_bytes);
  } else {
    int rc = fe_ctx_reserve(c, FE_WS_SORT, need);
    if (rc) return rc;
    base = (char*)c->ws[FE_WS_SORT];
  }
  uint32_t* k_in = (uint32_t*)base;
  uint32_t* k_out = (uint32_t*)(base + b_key);
  int32_t* v_in = (int32_t*)(base + 2 * b_key);
  int32_t* v_out = (int32_t*)(base + 3 * b_key);
  double* sorted = (double*)(base + 4 * b_key);
  void* temp = base + 4 * b_key + b_pts;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_morton, grid, dim3(256), 0, st, d_pts, n, lo[0], lo[1], lo[2], scale, k_in,
                     v_in);
  FE_HIPCHK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k_in, k_out, v_in, v_out, n, 0,
                                               30, st));
  hipLaunchKernelGGL(k_gather_pts, grid, dim3(256), 0, st, d_pts, v_out, n, sorted);
  FE_HIPCHK(hipGetLastError());
  *d_sorted = sorted;
  *d_perm = v_out;
  return FE_OK;
}
