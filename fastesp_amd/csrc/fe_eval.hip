// Evaluation kernels: single table (grid = point blocks x pair chunks of one
// class) and batched (persistent workgroups over a work list spanning many
// tables).  Both run fe_eval_block; see fe_tile.h for the arithmetic.
#include <hip/hip_runtime.h>

#include "fe_internal.h"
#include "fe_tile.h"

template <int L, int PPL, bool GRAD, bool MIXED>
__global__ __launch_bounds__(FE_WG) void k_eval(const double* __restrict__ rec,
                                                const float* __restrict__ g32, const int n_pairs,
                                                const int ppc, const double* __restrict__ pts,
                                                const int n_pts,
                                                const double* __restrict__ boys_tab,
                                                double* __restrict__ partial, const int row0,
                                                const int ld, const size_t plane,
                                                const float tau) {
  __shared__ double s_tab[FE_BOYS_NR * FE_BOYS_COLS];
  for (int i = threadIdx.x; i < FE_BOYS_NR * FE_BOYS_COLS; i += FE_WG) s_tab[i] = boys_tab[i];
  __syncthreads();
  const int chunk = blockIdx.y;
  const int p0 = chunk * ppc;
  const int p1 = min(n_pairs, p0 + ppc);
  const int base = blockIdx.x * PPL * FE_WG;
  fe_eval_block<L, PPL, GRAD, MIXED>(rec, g32, p0, p1, pts, base, n_pts, s_tab, tau,
                                     partial + (size_t)(row0 + chunk) * ld + base, plane);
}

template <int L, int PPL, bool GRAD, bool MIXED>
__global__ __launch_bounds__(FE_WG) void k_batch_eval(const FeBatchItem* __restrict__ items,
                                                      const int n_items,
                                                      const double* __restrict__ rec,
                                                      const float* __restrict__ g32,
                                                      const double* __restrict__ pts,
                                                      const double* __restrict__ boys_tab,
                                                      double* __restrict__ partial,
                                                      const size_t plane, const float tau) {
  __shared__ double s_tab[FE_BOYS_NR * FE_BOYS_COLS];
  for (int i = threadIdx.x; i < FE_BOYS_NR * FE_BOYS_COLS; i += FE_WG) s_tab[i] = boys_tab[i];
  __syncthreads();
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const FeBatchItem b = items[it];
    fe_eval_block<L, PPL, GRAD, MIXED>(rec, g32, b.p0, b.p1, pts, b.base, b.end, s_tab, tau,
                                       partial + b.out, plane);
  }
}

// ---------------------------------------------------------------------------
template <int L, int PPL>
static void launch_eval_LP(const fe_table* t, const FwdPlan& pl, const double* pts, int n,
                           double* partial, size_t plane, bool grad, bool mixed, float tau,
                           hipStream_t st) {
  const int M = grad ? L + 1 : L;
  const double* tab = t->ctx->d_boys + (size_t)M * FE_BOYS_NR * FE_BOYS_COLS;
  const dim3 grid(pl.n_pb[L], pl.chunks[L]);
  const float* g32 = t->d_rec32[L];
#define FE_K(G, X)                                                                            \
  hipLaunchKernelGGL((k_eval<L, PPL, G, X>), grid, dim3(FE_WG), 0, st, t->d_rec[L], g32,      \
                     (int)t->n_L[L], pl.ppc[L], pts, n, tab, partial, pl.row0[L], pl.ld,      \
                     plane, tau)
  if (grad) {
    if (mixed) FE_K(true, true); else FE_K(true, false);
  } else {
    if (mixed) FE_K(false, true); else FE_K(false, false);
  }
#undef FE_K
}

int fe_launch_eval(fe_table* t, const FwdPlan& pl, const double* pts, int n, double* partial,
                   size_t plane, bool grad, bool mixed, hipStream_t st) {
  if (mixed) {
    int rc = fe_ensure_g32(t, st);
    if (rc) return rc;
  }
  const float tau = (float)t->ctx->mixed_tau;
  FE_FOR_EACH_CLASS(pl, launch_eval_LP, t, pl, pts, n, partial, plane, grad, mixed, tau, st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

template <int L, int PPL>
static void launch_batch_LP(const FeBatchItem* items, int n_items, int grid, const double* rec,
                            const float* g32, const double* pts, const double* boys,
                            double* partial, size_t plane, bool grad, bool mixed, float tau,
                            hipStream_t st) {
  const int M = grad ? L + 1 : L;
  const double* tab = boys + (size_t)M * FE_BOYS_NR * FE_BOYS_COLS;
#define FE_K(G, X)                                                                            \
  hipLaunchKernelGGL((k_batch_eval<L, PPL, G, X>), dim3(grid), dim3(FE_WG), 0, st, items,     \
                     n_items, rec, g32, pts, tab, partial, plane, tau)
  if (grad) {
    if (mixed) FE_K(true, true); else FE_K(true, false);
  } else {
    if (mixed) FE_K(false, true); else FE_K(false, false);
  }
#undef FE_K
}

template <int L>
static void launch_batch_L(int ppl, const FeBatchItem* items, int n_items, int grid,
                           const double* rec, const float* g32, const double* pts,
                           const double* boys, double* partial, size_t plane, bool grad,
                           bool mixed, float tau, hipStream_t st) {
  if (ppl == 2)
    launch_batch_LP<L, 2>(items, n_items, grid, rec, g32, pts, boys, partial, plane, grad, mixed,
                          tau, st);
  else
    launch_batch_LP<L, 1>(items, n_items, grid, rec, g32, pts, boys, partial, plane, grad, mixed,
                          tau, st);
}

void fe_launch_batch_eval(int L, int ppl, const FeBatchItem* items, int n_items, int grid,
                          const double* rec, const float* g32, const double* pts,
                          const double* boys, double* partial, size_t plane, bool grad,
                          bool mixed, float tau, hipStream_t st) {
  switch (L) {
#define FE_C(l)                                                                               \
  case l:                                                                                     \
    launch_batch_L<l>(ppl, items, n_items, grid, rec, g32, pts, boys, partial, plane, grad,   \
                      mixed, tau, st);                                                        \
    break;
    FE_C(0) FE_C(1) FE_C(2) FE_C(3) FE_C(4) FE_C(5) FE_C(6)
#undef FE_C
  }
}
