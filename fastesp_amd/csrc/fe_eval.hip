// Evaluation kernels: single table (grid = point blocks x pair chunks of one
// class) and batched (persistent workgroups over a work list spanning many
// tables).  Both run fe_eval_block; see fe_tile.h for the arithmetic.
#include <hip/hip_runtime.h>

#include "fe_internal.h"
#include "fe_tile.h"

template <int L, int PPL, bool GRAD, bool MIXED>
__global__ __launch_bounds__(FE_WG, fe_eval_minw(L, PPL)) void k_eval(const double* __restrict__ rec,
                                                const float* __restrict__ g32,
                                                const int* __restrict__ n_pairs_p,
                                                const double* __restrict__ pts,
                                                const int n_pts,
                                                const double* __restrict__ boys_tab,
                                                double* __restrict__ partial, const int nadd,
                                                const int ld, const size_t plane,
                                                const float tau) {
  constexpr int NTAB = fe_boys_rows(GRAD ? L + 1 : L) * FE_BOYS_COLS;
  __shared__ double s_tab[NTAB];
  for (int i = threadIdx.x; i < NTAB; i += FE_WG) s_tab[i] = boys_tab[i];
  __syncthreads();
  // the kept count is read on the device (a rebuild may have changed it since
  // the launch was planned); the chunks split it evenly
  const int n_pairs = *n_pairs_p;
  const int ppc = (n_pairs + (int)gridDim.y - 1) / (int)gridDim.y;
  const int chunk = blockIdx.y;
  const int p0 = chunk * ppc;
  const int p1 = min(n_pairs, p0 + ppc);
  const int base = blockIdx.x * PPL * FE_WG;
  fe_eval_block<L, PPL, GRAD, MIXED>((fe_gdptr)rec, (fe_gfptr)g32, p0, p1, pts, base, n_pts,
                                     s_tab, tau, partial + (size_t)chunk * ld + base, plane,
                                     chunk < nadd);
}

// Persistent: each workgroup takes the next item of the queue until it is
// drained (every workgroup reaches the exit test after at most n_items + 1
// increments of the counter).
template <int L, int PPL, bool GRAD, bool MIXED>
__global__ __launch_bounds__(FE_WG, fe_eval_minw(L, PPL)) void k_batch_eval(const FeBatchItem* __restrict__ items,
                                                      const int n_items, int* __restrict__ counter,
                                                      const double* const* __restrict__ recs,
                                                      const int* const* __restrict__ kept,
                                                      const float* const* __restrict__ g32s,
                                                      const double* __restrict__ pts,
                                                      const double* __restrict__ boys_tab,
                                                      double* __restrict__ partial,
                                                      const size_t plane, const float tau) {
  constexpr int NTAB = fe_boys_rows(GRAD ? L + 1 : L) * FE_BOYS_COLS;
  __shared__ double s_tab[NTAB];
  __shared__ int s_it;
  for (int i = threadIdx.x; i < NTAB; i += FE_WG) s_tab[i] = boys_tab[i];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_it = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_it);
    if (it >= n_items) break;
    const FeBatchItem b = items[it];
    // items are cut from the tables' capacities when the batch rebuilds its
    // tables; the pairs past a table's kept count are not evaluated (an item
    // with none left still writes its zero row slot)
    const int p1 = min(b.p1, *kept[b.conf]);
    const int p0 = min(b.p0, p1);
    fe_eval_block<L, PPL, GRAD, MIXED>((fe_gdptr)recs[b.conf],
                                       (fe_gfptr)(MIXED ? g32s[b.conf] : nullptr), p0, p1,
                                       pts, b.base, b.end, s_tab, tau, partial + b.out, plane,
                                       false);
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
  mixed = mixed && t->ctx->mixed_L[L];
  const float* g32 = t->d_rec32[L];
  const int ph = fe_prof_begin(t->ctx, FE_PH_EVAL0 + L, st);
#define FE_K(G, X)                                                                            \
  hipLaunchKernelGGL((k_eval<L, PPL, G, X>), grid, dim3(FE_WG), 0, st, t->d_rec[L], g32,      \
                     t->d_stats + FE_ST_KEPT + L, pts, n, tab, partial, pl.nadd[L], pl.ld,    \
                     plane, tau)
  if (grad) {
    if (mixed) FE_K(true, true); else FE_K(true, false);
  } else {
    if (mixed) FE_K(false, true); else FE_K(false, false);
  }
#undef FE_K
  fe_prof_end(t->ctx, ph, st);
}

static void launch_eval_class(int L, const fe_table* t, const FwdPlan& pl, const double* pts,
                              int n, double* partial, size_t plane, bool grad, bool mixed,
                              float tau, hipStream_t st) {
  if (pl.chunks[L] == 0 || pl.n_pb[L] == 0) return;  // no pairs (or no points)
  double* rows = partial + (size_t)pl.row0[L] * pl.ld;
  switch (L) {
    case 0: FE_DISPATCH_PPL(pl, 0, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    case 1: FE_DISPATCH_PPL(pl, 1, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    case 2: FE_DISPATCH_PPL(pl, 2, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    case 3: FE_DISPATCH_PPL(pl, 3, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    case 4: FE_DISPATCH_PPL(pl, 4, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    case 5: FE_DISPATCH_PPL(pl, 5, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
    default: FE_DISPATCH_PPL(pl, 6, launch_eval_LP, t, pl, pts, n, rows, plane, grad, mixed, tau, st); break;
  }
}

int fe_launch_eval(fe_table* t, const FwdPlan& pl, const double* pts, int n, double* partial,
                   size_t plane, bool grad, bool mixed, hipStream_t st) {
  if (mixed) {
    int rc = fe_ensure_g32(t, st);
    if (rc) return rc;
  }
  const float tau = (float)t->ctx->mixed_tau;
  if (!pl.split) {
    for (int k = 0; FE_ORDER_ONE[k] >= 0; ++k)
      if (FE_ORDER_ONE[k] <= FE_LMAX)
        launch_eval_class(FE_ORDER_ONE[k], t, pl, pts, n, partial, plane, grad, mixed, tau, st);
  } else if (pl.split == 2) {
    // classes 6 and 5 on the second stream beside classes 0 and 1; joined
    // before class 2, so the large classes run alone
    hipStream_t s2 = st;
    int rc = fe_fork(t->ctx, st, &s2);
    if (rc) return rc;
    launch_eval_class(6, t, pl, pts, n, partial, plane, grad, mixed, tau, s2);
    launch_eval_class(0, t, pl, pts, n, partial, plane, grad, mixed, tau, st);
    launch_eval_class(5, t, pl, pts, n, partial, plane, grad, mixed, tau, s2);
    launch_eval_class(1, t, pl, pts, n, partial, plane, grad, mixed, tau, st);
    rc = fe_join(t->ctx, st);
    if (rc) return rc;
    for (int L = 2; L <= 4; ++L)
      launch_eval_class(L, t, pl, pts, n, partial, plane, grad, mixed, tau, st);
  } else {
    // the two row sets are independent: the first on the caller's stream,
    // the second on the context's second stream
    const auto& order = pl.split == 2 ? FE_ORDER_SMALL : FE_ORDER_SPLIT;
    hipStream_t s2 = st;
    int rc = fe_fork(t->ctx, st, &s2);
    if (rc) return rc;
    const hipStream_t sg[2] = {st, s2};
    // interleaved, so both streams have work queued from the start
    for (int k = 0; k <= FE_LMAX; ++k)
      for (int g = 0; g < 2; ++g) {
        int len = 0;
        while (order[g][len] >= 0) ++len;
        if (k < len && order[g][k] <= FE_LMAX)
          launch_eval_class(order[g][k], t, pl, pts, n, partial, plane, grad, mixed, tau, sg[g]);
      }
    rc = fe_join(t->ctx, st);
    if (rc) return rc;
  }
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
template <int L, int PPL, bool G, bool X>
static int batch_LPGX(const FeBatchEvalArgs* a, hipStream_t st) {
  if (!a) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_eval<L, PPL, G, X>, FE_WG, 0) !=
        hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    return nb;
  }
  const int M = G ? L + 1 : L;
  const double* tab = a->boys + (size_t)M * FE_BOYS_NR * FE_BOYS_COLS;
  hipLaunchKernelGGL((k_batch_eval<L, PPL, G, X>), dim3(a->grid), dim3(FE_WG), 0, st, a->items,
                     a->n_items, a->counter, a->recs, a->kept, a->g32s, a->pts, tab, a->partial,
                     a->plane, a->tau);
  return 0;
}

template <int L, int PPL>
static int batch_LP(bool grad, bool mixed, const FeBatchEvalArgs* a, hipStream_t st) {
  if (grad) return mixed ? batch_LPGX<L, PPL, true, true>(a, st) : batch_LPGX<L, PPL, true, false>(a, st);
  return mixed ? batch_LPGX<L, PPL, false, true>(a, st) : batch_LPGX<L, PPL, false, false>(a, st);
}

template <int L>
static int batch_L(int ppl, bool grad, bool mixed, const FeBatchEvalArgs* a, hipStream_t st) {
  switch (ppl) {
    case 4: return batch_LP<L, 4>(grad, mixed, a, st);
    case 2: return batch_LP<L, 2>(grad, mixed, a, st);
    default: return batch_LP<L, 1>(grad, mixed, a, st);
  }
}

int fe_batch_eval(int L, int ppl, bool grad, bool mixed, const FeBatchEvalArgs* a,
                  hipStream_t st) {
  switch (L) {
    case 0: return batch_L<0>(ppl, grad, mixed, a, st);
    case 1: return batch_L<1>(ppl, grad, mixed, a, st);
    case 2: return batch_L<2>(ppl, grad, mixed, a, st);
    case 3: return batch_L<3>(ppl, grad, mixed, a, st);
    case 4: return batch_L<4>(ppl, grad, mixed, a, st);
    case 5: return batch_L<5>(ppl, grad, mixed, a, st);
    default: return batch_L<6>(ppl, grad, mixed, a, st);
  }
}
