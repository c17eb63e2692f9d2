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
// index), in the context's FE_WS_SORT slot or the caller's buffer.  The
// quantisation frame is the nuclei's bounding box grown by 16 Bohr (points
// outside are clamped, which only costs locality, never correctness).  The
// sort is the library's own stable radix sort (fe_prim.hip): no memsets and no
// host synchronisation, so the sort can be recorded into a hipGraph.
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
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_key = up((size_t)n * 4), b_pts = up((size_t)n * 3 * sizeof(double));
  const size_t need = 2 * b_key + b_pts + fe_radix_sort_scratch(n);
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
  uint32_t* key = (uint32_t*)base;
  int32_t* perm = (int32_t*)(base + b_key);
  double* sorted = (double*)(base + 2 * b_key);
  void* tmp = base + 2 * b_key + b_pts;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_morton, grid, dim3(256), 0, st, d_pts, n, lo[0], lo[1], lo[2], scale, key,
                     perm);
  if (int rc = fe_radix_sort(key, perm, n, 30, tmp, st)) return rc;
  hipLaunchKernelGGL(k_gather_pts, grid, dim3(256), 0, st, d_pts, perm, n, sorted);
  FE_HIPCHK(hipGetLastError());
  *d_sorted = sorted;
  *d_perm = perm;
  return FE_OK;
}
