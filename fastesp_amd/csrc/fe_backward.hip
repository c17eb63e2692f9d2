// Adjoint of the ESP evaluation: the field dV/dC and dL/dP for a loss L(V).
// The reference has no gradients; this is the hand-written backward behind
// the torch.autograd.Function of fastesp_amd.autograd.
//
//  fe_esp_forward_grad : V and dV/dC in one pass (k_eval<L, PPL, GRAD=true>): per
//      (point, pair) the Boys vector up to L+1 and
//      dI/dC = K sum_k g_k [ d(R^a)/dR F_m - 2p R R^a F_{m+1} ],
//      nuclear part dV/dC = -sum_A Z_A (C-R_A)/|C-R_A|^3.
//  fe_esp_backward     : dL/dG_k = -sum_i gV_i K R_i^a F_m(T_i) with one lane per
//      pair and the points (x, y, z, gV) streamed as wave-uniform scalar
//      loads (k_dG<L>; no cross-lane reduction), chunked over points with a
//      fixed-order reduction (bitwise reproducible); then dL/dP by the adjoint
//      of the table contraction: d etab_t = sum_k dG_k (2 pi/p)(2p)^m M0[k,t]
//      per kept record (k_det<L>), then one lane per (shell pair, function
//      pair) task (k_dp_tasks<L>): d blk_{mu nu} = c_i c_j sum over the kept
//      primitive pairs and t of d etab_t Ex Ey Ez, and both triangles of P
//      receive d blk because the block is folded P_ij + P_ji^T
//      (fastesp/esp.py:185-187).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"
#include "fe_adjoint.h"
#include "fe_reduce.h"

// 64 points per 256-thread block, each point's rows split over the 4 waves
__global__ __launch_bounds__(256) void k_esp_reduce_grad(const double* __restrict__ partial,
                                                          const int n_rows, const int ld,
                                                          const size_t plane,
                                                          const double* __restrict__ pts,
                                                          const int n_pts,
                                                          const double* __restrict__ atoms,
                                                          const int n_atoms,
                                                          const int* __restrict__ perm,
                                                          double* __restrict__ V,
                                                          double* __restrict__ dVdC) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  double sum[4];
  fe_sum_rows_split<4>(partial + i, i < n_pts, n_rows, ld, plane, sum);
  if (threadIdx.x >= 64 || i >= n_pts) return;
  double nuc, nx, ny, nz;
  fe_nuclear_grad(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], atoms, n_atoms, nuc, nx, ny, nz);
  const int o = perm ? perm[i] : i;
  V[o] = nuc - sum[0];
  dVdC[3 * o] = nx - sum[1];
  dVdC[3 * o + 1] = ny - sum[2];
  dVdC[3 * o + 2] = nz - sum[3];
}

extern "C" int fe_esp_forward_grad(fe_table* t, const double* pts, int64_t n, int precision,
                                   int ptr_kind, void* stream, double* V, double* dVdC) {
  if (!t || (n > 0 && (!pts || !V || !dVdC)))
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_grad: null argument");
  if (n < 0 || n > (int64_t)INT32_MAX / 4)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_grad: bad point count");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_grad: bad ptr_kind");
  if (precision != FE_PREC_FP64 && precision != FE_PREC_MIXED)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_grad: bad precision");
  if (n == 0) return FE_OK;
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const double* d_pts = nullptr;
  int rc = fe_stage_points(c, pts, n, ptr_kind, st, &d_pts);
  if (rc) return rc;
  double *d_V = V, *d_G = dVdC;
  if (ptr_kind == FE_PTR_HOST) {
    rc = fe_ctx_reserve(c, FE_WS_OUT, (size_t)n * 4 * sizeof(double));
    if (rc) return rc;
    d_V = (double*)c->ws[FE_WS_OUT];
    d_G = d_V + n;
  }
  const int32_t* perm = nullptr;
  if (c->sort_points && n >= 2 * FE_WG) {
    const int ps = fe_prof_begin(c, FE_PH_SORT, st);
    rc = fe_sort_points(c, t, d_pts, (int)n, st, &d_pts, &perm);
    if (rc) return rc;
    fe_prof_end(c, ps, st);
  }
  FwdPlan pl = fe_plan(t, n, c->ppl_grad_L, c->target_wg, c->split_eval_grad);
  const size_t plane = (size_t)std::max(pl.rows, 1) * pl.ld;
  const int groups = fe_row_groups((int)n, pl.rows);
  const size_t gplane = (size_t)groups * pl.ld;
  rc = fe_ctx_reserve(c, FE_WS_PARTIAL, 4 * (plane + gplane) * sizeof(double));
  if (rc) return rc;
  double* partial = (double*)c->ws[FE_WS_PARTIAL];
  rc = fe_launch_eval(t, pl, d_pts, (int)n, partial, plane, true, precision == FE_PREC_MIXED, st);
  if (rc) return rc;
  const int pr = fe_prof_begin(c, FE_PH_REDUCE, st);
  const double* red = partial;
  int nrows = pl.rows;
  size_t rplane = plane;
  if (groups) {
    // few points: the rows are first added in groups over a 2-D grid
    double* staged = partial + 4 * plane;
    const int rpg = (pl.rows + groups - 1) / groups;
    hipLaunchKernelGGL(k_rows_stage<4>, dim3((unsigned)((n + 63) / 64), (unsigned)groups),
                       dim3(256), 0, st, partial, pl.rows, pl.ld, plane, (int)n, rpg, staged,
                       gplane);
    red = staged;
    nrows = groups;
    rplane = gplane;
  }
  hipLaunchKernelGGL(k_esp_reduce_grad, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st,
                     red, nrows, pl.ld, rplane, d_pts, (int)n, t->d_atoms, t->n_atoms, perm,
                     d_V, d_G);
  fe_prof_end(c, pr, st);
  FE_HIPCHK(hipGetLastError());
  if (ptr_kind == FE_PTR_HOST) {
    FE_HIPCHK(hipMemcpyAsync(V, d_V, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipMemcpyAsync(dVdC, d_G, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
  }
  return FE_OK;
}

// ---------------------------------------------------------------------------
// dL/dG
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack_pw(const double* __restrict__ pts,
                                                 const double* __restrict__ gV, const int n,
                                                 double* __restrict__ pw) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  pw[4 * i] = pts[3 * i];
  pw[4 * i + 1] = pts[3 * i + 1];
  pw[4 * i + 2] = pts[3 * i + 2];
  pw[4 * i + 3] = gV[i];
}

// one lane per pair of class L; points (x, y, z, gV) of chunk blockIdx.y are
// wave-uniform; partial[(chunk * NT + k) * stride + pair].  The kept count is
// read on the device; `stride` (>= it) is the count the launch was planned for.
template <int L>
__global__ __launch_bounds__(256, fe_dg_minw(L)) void k_dG(const double* __restrict__ rec,
                                            const int* __restrict__ n_pairs_p, const int stride,
                                            const double* __restrict__ pw, const int n_pts,
                                            const int ppc, const double* __restrict__ boys_tab,
                                            double* __restrict__ partial) {
  const int n_pairs = min(*n_pairs_p, stride);
  if ((int)(blockIdx.x * 256) >= n_pairs) return;  // block-uniform
  constexpr int NTAB = fe_boys_rows(L) * FE_BOYS_COLS;
  __shared__ double s_tab[NTAB];
  for (int i = threadIdx.x; i < NTAB; i += 256) s_tab[i] = boys_tab[i];
  __syncthreads();
  const int q0 = blockIdx.y * ppc;
  fe_dg_block<L>(rec, n_pairs, blockIdx.x * 256 + threadIdx.x, pw, q0, min(n_pts, q0 + ppc),
                 s_tab, partial + (size_t)blockIdx.y * FeL<L>::NT * stride, stride);
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ x, const int64_t n,
                                              const double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}

// dG[pair * NT + k] = sum over chunks; optional flat scatter.  64 (k, pair)
// elements per 256-thread block: wave s adds the chunks c = s (mod 4) with four
// chains (four loads in flight per lane: the high classes have few elements
// and hundreds of chunks, so the sum is latency-bound), and the four slices are
// combined in a fixed order in LDS (bitwise reproducible).
template <int L>
__global__ __launch_bounds__(256) void k_dG_reduce(const double* __restrict__ partial,
                                                   const int n_pairs, const int* __restrict__ kept,
                                                   const int n_chunks, double* __restrict__ dG,
                                                   const int* __restrict__ flatoff,
                                                   double* __restrict__ flat) {
  // n_pairs: the planned count (layout stride); pairs from the device kept
  // count on are not part of the table
  constexpr int NT = FeL<L>::NT;
  __shared__ double s_red[4][64];
  const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  const int nk = min(*kept, n_pairs);
  const int k0 = e < (int64_t)n_pairs * NT ? (int)(e / n_pairs) : 0;
  const bool valid = e < (int64_t)n_pairs * NT && (int)(e - (int64_t)k0 * n_pairs) < nk;
  const int k = valid ? k0 : 0;
  const int pair = valid ? (int)(e - (int64_t)k * n_pairs) : 0;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (valid) {
    const size_t stride = (size_t)NT * n_pairs;
    const double* p = partial + (size_t)k * n_pairs + pair;
    int c = sl;
    for (; c + 12 < n_chunks; c += 16) {
      const double x0 = p[(size_t)c * stride], x1 = p[(size_t)(c + 4) * stride];
      const double x2 = p[(size_t)(c + 8) * stride], x3 = p[(size_t)(c + 12) * stride];
      a0 += x0;
      a1 += x1;
      a2 += x2;
      a3 += x3;
    }
    for (; c < n_chunks; c += 4) a0 += p[(size_t)c * stride];
  }
  s_red[sl][lane] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (sl != 0 || !valid) return;
  const double s = (s_red[0][lane] + s_red[1][lane]) + (s_red[2][lane] + s_red[3][lane]);
  dG[(size_t)pair * NT + k] = s;
  if (flat) flat[flatoff[pair] + k] = s;
}

struct DGPlan {
  int chunks[FE_LMAX + 1];
  int ppc[FE_LMAX + 1];
  size_t part_off[FE_LMAX + 1];  // doubles
  size_t dg_off[FE_LMAX + 1];    // doubles
  size_t part_total, dg_total;
};

static DGPlan plan_dG(const fe_table* t, int64_t n, int target_wg) {
  DGPlan d{};
  for (int L = 0; L <= FE_LMAX; ++L) {
    const int64_t nL = fe_plan_n(t, L);
    d.part_off[L] = d.part_total;
    d.dg_off[L] = d.dg_total;
    if (!nL) continue;
    const int NT = fe_host::NTERM[L];
    // the point chunks from the pairs expected to be kept (blocks past the
    // kept count exit at once), the buffers for all nL
    const int64_t nE = fe_plan_est(t, L);
    const int64_t nblk = (nE + 255) / 256;
    int64_t ch = std::max<int64_t>(1, target_wg / nblk);
    // bound the partial buffer (written and re-read once) and keep >= 64
    // points per chunk so the reduction stays small next to the kernel
    ch = std::min<int64_t>(ch, std::max<int64_t>(1, (16ll << 20) / (nE * NT)));
    ch = std::min<int64_t>(ch, std::max<int64_t>(1, n / 64));
    ch = std::min<int64_t>(ch, 65535);
    const int64_t ppc = (n + ch - 1) / ch;
    ch = (n + ppc - 1) / ppc;
    d.chunks[L] = (int)ch;
    d.ppc[L] = (int)ppc;
    d.part_total += (size_t)ch * nL * NT;
    d.dg_total += (size_t)nL * NT;
  }
  return d;
}

template <int L>
static void launch_dG_L(const fe_table* t, const DGPlan& d, const double* pw, int n,
                        double* part, double* dG, double* flat, hipStream_t st) {
  const int64_t nL = fe_plan_n(t, L);
  if (!nL) return;
  const double* tab = t->ctx->d_boys + (size_t)L * FE_BOYS_NR * FE_BOYS_COLS;
  const int* kept = t->d_stats + FE_ST_KEPT + L;
  const int pg = fe_prof_begin(t->ctx, FE_PH_DG0 + L, st);
  hipLaunchKernelGGL(k_dG<L>, dim3((unsigned)((nL + 255) / 256), d.chunks[L]), dim3(256), 0, st,
                     t->d_rec[L], kept, (int)nL, pw, n, d.ppc[L], tab, part + d.part_off[L]);
  fe_prof_end(t->ctx, pg, st);
  const int64_t ne = nL * FeL<L>::NT;
  const int pr = fe_prof_begin(t->ctx, FE_PH_DGRED0 + L, st);
  hipLaunchKernelGGL(k_dG_reduce<L>, dim3((unsigned)((ne + 63) / 64)), dim3(256), 0, st,
                     part + d.part_off[L], (int)nL, kept, d.chunks[L], dG + d.dg_off[L],
                     t->d_flatoff[L], flat);
  fe_prof_end(t->ctx, pr, st);
}

// ---------------------------------------------------------------------------
// dL/dG -> dL/dP, one workgroup (64 lanes) per shell pair
// ---------------------------------------------------------------------------
// d etab of every kept record of class L (one lane each) from its dL/dG
template <int L>
__global__ __launch_bounds__(256) void k_det(const double* __restrict__ rec,
                                             const int* __restrict__ kept, const int cap,
                                             const double* __restrict__ dG,
                                             double* __restrict__ det) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= min(*kept, cap)) return;
  fe_det_from_dg<L>(dG + (size_t)i * FeL<L>::NT, rec[(size_t)i * FeL<L>::STRIDE + 3],
                    det + (size_t)i * FeL<L>::NM);
}

// the dL/dP tasks of class L (one lane each; their d etab is ready)
template <int L>
__global__ __launch_bounds__(256) void k_dp_tasks(const DPArgs a, const int n_tasks) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_tasks) return;
  fe_dp_task<L>(a, a.tasks[a.task_cls_off[L] + i]);
}

template <int L>
static void launch_dp_L(const fe_table* t, const DPArgs& pa, hipStream_t s) {
  const int64_t nL = fe_plan_n(t, L);
  const int nt = t->task_cls_off[L + 1] - t->task_cls_off[L];
  if (nL > 0)
    hipLaunchKernelGGL(k_det<L>, dim3((unsigned)((nL + 255) / 256)), dim3(256), 0, s,
                       t->d_rec[L], t->d_stats + FE_ST_KEPT + L, (int)nL, pa.dG[L],
                       (double*)pa.det[L]);
  if (nt > 0)
    hipLaunchKernelGGL(k_dp_tasks<L>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, pa,
                       nt);
}

// ---------------------------------------------------------------------------
extern "C" int fe_esp_backward(fe_table* t, const double* pts, int64_t n, const double* gV,
                               int precision, int ptr_kind, void* stream, double* dLdG,
                               double* dLdP) {
  if (!t || (n > 0 && (!pts || !gV)))
    return fe_set_error(FE_ERR_ARG, "fe_esp_backward: null argument");
  if (n < 0 || n > (int64_t)INT32_MAX / 8)
    return fe_set_error(FE_ERR_ARG, "fe_esp_backward: bad point count");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_esp_backward: bad ptr_kind");
  if (precision != FE_PREC_FP64 && precision != FE_PREC_MIXED)
    return fe_set_error(FE_ERR_ARG, "fe_esp_backward: bad precision");
  if (dLdP && !t->has_basis)
    return fe_set_error(FE_ERR_STATE,
                        "fe_esp_backward: dL/dP needs a table built from a basis (fe_table_build)");
  // the flat dL/dG layout (n_terms values) needs the host-side counts
  if (dLdG)
    if (int rc = fe_table_sync_counts(t)) return rc;
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // staging: points + weights packed as (x, y, z, gV); with no points nothing
  // is read from pts or gV (either may be NULL) and dL/dG is zero
  const double* d_pts = nullptr;
  int rc = FE_OK;
  if (n > 0) {
    rc = fe_stage_points(c, pts, n, ptr_kind, st, &d_pts);
    if (rc) return rc;
  }
  const DGPlan d = plan_dG(t, std::max<int64_t>(n, 1), c->target_wg);
  const size_t pw_bytes = (size_t)std::max<int64_t>(n, 1) * 4 * sizeof(double);
  rc = fe_ctx_reserve(c, FE_WS_AUX0, pw_bytes + (size_t)std::max<int64_t>(n, 1) * sizeof(double));
  if (rc) return rc;
  double* pw = (double*)c->ws[FE_WS_AUX0];
  const double* d_gV = gV;
  if (ptr_kind == FE_PTR_HOST && n > 0) {
    double* tmp = pw + 4 * n;
    FE_HIPCHK(hipMemcpyAsync(tmp, gV, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    d_gV = tmp;
  }
  if (n > 0) {
    const int pp = fe_prof_begin(c, FE_PH_PACK, st);
    hipLaunchKernelGGL(k_pack_pw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_pts,
                       d_gV, (int)n, pw);
    fe_prof_end(c, pp, st);
  }
  rc = fe_ctx_reserve(c, FE_WS_AUX1, std::max<size_t>(d.part_total, 1) * sizeof(double));
  if (rc) return rc;
  rc = fe_ctx_reserve(c, FE_WS_AUX2, std::max<size_t>(d.dg_total, 1) * sizeof(double));
  if (rc) return rc;
  double* part = (double*)c->ws[FE_WS_AUX1];
  double* dG = (double*)c->ws[FE_WS_AUX2];
  double* flat = nullptr;
  if (dLdG) {
    if (ptr_kind == FE_PTR_DEVICE) {
      flat = dLdG;
    } else {
      rc = fe_ctx_reserve(c, FE_WS_AUX3, std::max<size_t>(t->n_terms, 1) * sizeof(double));
      if (rc) return rc;
      flat = (double*)c->ws[FE_WS_AUX3];
    }
  }
  // dL/dP: every shell pair writes its own block of dL/dP (and its mirror), so
  // the shell pairs of class L run as soon as class L's dL/dG is reduced, on
  // the same stream
  double* d_dP = dLdP;
  if (dLdP && ptr_kind == FE_PTR_HOST) {
    rc = fe_ctx_reserve(c, FE_WS_OUT, (size_t)t->n_func * t->n_func * sizeof(double));
    if (rc) return rc;
    d_dP = (double*)c->ws[FE_WS_OUT];
  }
  DPArgs pa = dLdP ? fe_dp_args(t, dG, d.dg_off, d_dP) : DPArgs{};
  // the d etab of class L go to its partial rows, which its dL/dG reduction
  // has consumed by then (same stream); they hold at least NT >= NM values
  // per record
  for (int L = 0; L <= FE_LMAX; ++L) pa.det[L] = part + d.part_off[L];
  auto launch_dp = [&](int L, hipStream_t s) {
    // classes without shell pairs have no entries of dL/dP
    if (!dLdP || t->task_cls_off[L + 1] == t->task_cls_off[L]) return;
    const int pd = fe_prof_begin(c, FE_PH_DP, s);
    switch (L) {
      case 0: launch_dp_L<0>(t, pa, s); break;
      case 1: launch_dp_L<1>(t, pa, s); break;
      case 2: launch_dp_L<2>(t, pa, s); break;
      case 3: launch_dp_L<3>(t, pa, s); break;
      case 4: launch_dp_L<4>(t, pa, s); break;
      case 5: launch_dp_L<5>(t, pa, s); break;
      case 6: launch_dp_L<6>(t, pa, s); break;
      default: break;
    }
    fe_prof_end(c, pd, s);
  };
  if (n == 0) {
    if (d.dg_total)
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((d.dg_total + 255) / 256)), dim3(256), 0, st, dG,
                         (int64_t)d.dg_total, 0.0);
    if (flat && t->n_terms)
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((t->n_terms + 255) / 256)), dim3(256), 0, st,
                         flat, (int64_t)t->n_terms, 0.0);
    for (int L = 0; L <= FE_LMAX; ++L) launch_dp(L, st);
  } else {
    // The classes are independent (own partial rows, own reduction, own
    // shell pairs), so they run on the context's dL/dG streams (tuning key
    // 6) with about the same work each: the tail of one class's launch fills
    // with another stream's workgroups, and the latency-bound dL/dP of a
    // class runs beside another stream's dL/dG.
    hipStream_t ss[3];
    const int ns = c->dg_streams;
    rc = fe_fork_n(c, st, ns, ss);
    if (rc) return rc;
    // Largest class first, each to the stream with the least estimated work
    // so far; a class's dL/dG per (pair, point) costs about 20 + 2.2 NT
    // instructions (NT its term count), measured on config 1.
    double work[FE_LMAX + 1];
    for (int L = 0; L <= FE_LMAX; ++L)
      work[L] = (double)fe_plan_est(t, L) * (20.0 + 2.2 * fe_host::NTERM[L]);
    int order[FE_LMAX + 1], q[FE_LMAX + 1];
    fe_lpt(work, ns, order, q);
    for (int k = 0; k <= FE_LMAX; ++k) {
      const int L = order[k];
      hipStream_t s = ss[q[k]];
      switch (L) {
        case 0: launch_dG_L<0>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 1: launch_dG_L<1>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 2: launch_dG_L<2>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 3: launch_dG_L<3>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 4: launch_dG_L<4>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 5: launch_dG_L<5>(t, d, pw, (int)n, part, dG, flat, s); break;
        case 6: launch_dG_L<6>(t, d, pw, (int)n, part, dG, flat, s); break;
        default: break;
      }
      launch_dp(L, s);
    }
    rc = fe_join_n(c, st, ns);
    if (rc) return rc;
  }
  FE_HIPCHK(hipGetLastError());
  if (ptr_kind == FE_PTR_HOST) {
    if (dLdG && t->n_terms)
      FE_HIPCHK(hipMemcpyAsync(dLdG, flat, (size_t)t->n_terms * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    if (dLdP)
      FE_HIPCHK(hipMemcpyAsync(dLdP, d_dP, (size_t)t->n_func * t->n_func * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
  }
  return FE_OK;
}

// ---------------------------------------------------------------------------
// Work counts (fe_eval_work): how many (point, pair) evaluations of each
// class take the far-field path and how many the tabulated one, in the
// evaluation launches (value + field, or value only) and in the dL/dG
// launches.  The far test of both is per wavefront, so the counting kernels
// repeat the launches' decompositions exactly: the same sorted points, points
// per lane, pair chunks and pair blocks; only the arithmetic is left out.
// Each wavefront-level decision is weighted by the wavefront's valid points
// (evaluation) or valid pairs (dL/dG).
// ---------------------------------------------------------------------------
template <int M, int PPL>
__global__ __launch_bounds__(FE_WG) void k_count_eval(const double* __restrict__ rec, const int S,
                                                      const int* __restrict__ n_pairs_p,
                                                      const double* __restrict__ pts,
                                                      const int n_pts,
                                                      unsigned long long* __restrict__ out) {
  const int n_pairs = *n_pairs_p;
  const int ppc = (n_pairs + (int)gridDim.y - 1) / (int)gridDim.y;
  const int p0 = blockIdx.y * ppc, p1 = min(n_pairs, p0 + ppc);
  const int base = blockIdx.x * PPL * FE_WG;
  double cx[PPL], cy[PPL], cz[PPL];
  int nv = 0;
#pragma unroll
  for (int q = 0; q < PPL; ++q) {
    const int i = base + q * FE_WG + (int)threadIdx.x;
    nv += i < n_pts;
    const int ic = min(i, n_pts - 1);
    cx[q] = pts[3 * (size_t)ic];
    cy[q] = pts[3 * (size_t)ic + 1];
    cz[q] = pts[3 * (size_t)ic + 2];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nv += __shfl_xor(nv, o, 64);
  unsigned long long far = 0, near = 0;
  for (int ip = p0; ip < p1; ++ip) {
    const double* r = rec + (size_t)ip * S;
    const double Px = r[0], Py = r[1], Pz = r[2];
    const double r2far = fe_tfar(M) * (r[6] + r[6]);  // the far test of fe_eval_tile
    int nr = 0;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
      const double x = cx[q] - Px, y = cy[q] - Py, z = cz[q] - Pz;
      nr |= fma(x, x, fma(y, y, z * z)) < r2far;
    }
    if (__any(nr)) ++near;
    else ++far;
  }
  if ((threadIdx.x & 63) == 0 && nv > 0 && p1 > p0) {
    atomicAdd(out, far * (unsigned long long)nv);
    atomicAdd(out + 1, near * (unsigned long long)nv);
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_count_dg(const double* __restrict__ rec, const int S,
                                                  const int* __restrict__ n_pairs_p,
                                                  const int stride, const double* __restrict__ pts,
                                                  const int n_pts, const int ppc,
                                                  unsigned long long* __restrict__ out) {
  const int n_pairs = min(*n_pairs_p, stride);
  if ((int)(blockIdx.x * 256) >= n_pairs) return;  // block-uniform, as in k_dG
  const int pair = blockIdx.x * 256 + threadIdx.x;
  const int ip = min(pair, n_pairs - 1);
  const double* r = rec + (size_t)ip * S;
  const double Px = r[0], Py = r[1], Pz = r[2], p = r[3], hp = r[6];
  const double r2far = fe_tfar(L) * (hp + hp);  // the far test of fe_dg_sum
  constexpr bool SPLIT = fe_dg_minw(L) == 1;
  const int nvp = __popcll(__ballot(pair < n_pairs));
  const int q0 = blockIdx.y * ppc, q1 = min(n_pts, q0 + ppc);
  unsigned long long far = 0, near = 0;
  for (int q = q0; q < q1; ++q) {
    const double x = pts[3 * (size_t)q] - Px, y = pts[3 * (size_t)q + 1] - Py,
                 z = pts[3 * (size_t)q + 2] - Pz;
    const double r2 = fma(x, x, fma(y, y, z * z));
    if (__any(SPLIT ? r2 < r2far : p * r2 < fe_tfar(L))) ++near;
    else ++far;
  }
  if ((threadIdx.x & 63) == 0 && nvp > 0) {
    atomicAdd(out, far * (unsigned long long)nvp);
    atomicAdd(out + 1, near * (unsigned long long)nvp);
  }
}

template <int L, int PPL>
static void count_eval_LP(const fe_table* t, const FwdPlan& pl, const double* pts, int n, bool grad,
                          unsigned long long* out, hipStream_t st) {
  const dim3 grid(pl.n_pb[L], pl.chunks[L]);
  const int S = fe_host::STRIDE[L];
  const int* np = t->d_stats + FE_ST_KEPT + L;
  if (grad)
    hipLaunchKernelGGL((k_count_eval<L + 1, PPL>), grid, dim3(FE_WG), 0, st, t->d_rec[L], S, np,
                       pts, n, out + 2 * L);
  else
    hipLaunchKernelGGL((k_count_eval<L, PPL>), grid, dim3(FE_WG), 0, st, t->d_rec[L], S, np, pts,
                       n, out + 2 * L);
}

template <int L>
static void count_dg_L(const fe_table* t, const DGPlan& d, const double* pts, int n,
                       unsigned long long* out, hipStream_t st) {
  const int64_t nL = fe_plan_n(t, L);
  if (!nL) return;
  hipLaunchKernelGGL(k_count_dg<L>, dim3((unsigned)((nL + 255) / 256), d.chunks[L]), dim3(256), 0,
                     st, t->d_rec[L], fe_host::STRIDE[L], t->d_stats + FE_ST_KEPT + L, (int)nL,
                     pts, n, d.ppc[L], out + 2 * (FE_LMAX + 1) + 2 * L);
}

extern "C" int fe_eval_work(fe_table* t, const double* pts, int64_t n, int grad, int ptr_kind,
                            void* stream, int64_t* counts) {
  if (!t || !counts || (n > 0 && !pts))
    return fe_set_error(FE_ERR_ARG, "fe_eval_work: null argument");
  if (n < 0 || n > (int64_t)INT32_MAX / 8)
    return fe_set_error(FE_ERR_ARG, "fe_eval_work: bad point count");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_eval_work: bad ptr_kind");
  for (int i = 0; i < 4 * (FE_LMAX + 1); ++i) counts[i] = 0;
  if (n == 0) return FE_OK;
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const double* d_pts = nullptr;
  int rc = fe_stage_points(c, pts, n, ptr_kind, st, &d_pts);
  if (rc) return rc;
  // the evaluation sees the Morton-sorted points, dL/dG the caller's order
  const double* s_pts = d_pts;
  const int32_t* perm = nullptr;
  if (c->sort_points && n >= 2 * FE_WG) {
    rc = fe_sort_points(c, t, d_pts, (int)n, st, &s_pts, &perm);
    if (rc) return rc;
  }
  rc = fe_ctx_reserve(c, FE_WS_AUX3, 4 * (FE_LMAX + 1) * sizeof(unsigned long long));
  if (rc) return rc;
  unsigned long long* d_out = (unsigned long long*)c->ws[FE_WS_AUX3];
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, (double*)d_out,
                     (int64_t)(4 * (FE_LMAX + 1)), 0.0);
  const FwdPlan pl = fe_plan(t, n, grad ? c->ppl_grad_L : c->ppl_L, c->target_wg);
  FE_FOR_EACH_CLASS(pl, count_eval_LP, t, pl, s_pts, (int)n, grad != 0, d_out, st);
  const DGPlan d = plan_dG(t, n, c->target_wg);
  count_dg_L<0>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<1>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<2>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<3>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<4>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<5>(t, d, d_pts, (int)n, d_out, st);
  count_dg_L<6>(t, d, d_pts, (int)n, d_out, st);
  FE_HIPCHK(hipGetLastError());
  unsigned long long h[4 * (FE_LMAX + 1)];
  FE_HIPCHK(hipMemcpyAsync(h, d_out, sizeof h, hipMemcpyDeviceToHost, st));
  FE_HIPCHK(hipStreamSynchronize(st));
  // counts: [far eval L][near eval L] x 7, then [far dG L][near dG L] x 7 ->
  // out[L], out[7 + L], out[14 + L], out[21 + L]
  for (int L = 0; L <= FE_LMAX; ++L) {
    counts[L] = (int64_t)h[2 * L];
    counts[FE_LMAX + 1 + L] = (int64_t)h[2 * L + 1];
    counts[2 * (FE_LMAX + 1) + L] = (int64_t)h[2 * (FE_LMAX + 1) + 2 * L];
    counts[3 * (FE_LMAX + 1) + L] = (int64_t)h[2 * (FE_LMAX + 1) + 2 * L + 1];
  }
  return FE_OK;
}
