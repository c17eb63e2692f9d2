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
//      of the table contraction, one workgroup per shell pair (k_dg_to_dp):
//      d etab_t = sum_k dG_k (2 pi/p)(2p)^m M0[k,t], d blk_{mu nu} = c_i c_j
//      sum_t d etab_t Ex Ey Ez, and both triangles of P receive d blk because
//      the block is folded P_ij + P_ji^T (fastesp/esp.py:185-187).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"
#include "fe_hermite.h"

__global__ __launch_bounds__(64) void k_esp_reduce_grad(const double* __restrict__ partial,
                                                         const int n_rows, const int ld,
                                                         const size_t plane,
                                                         const double* __restrict__ pts,
                                                         const int n_pts,
                                                         const double* __restrict__ atoms,
                                                         const int n_atoms,
                                                         const int* __restrict__ perm,
                                                         double* __restrict__ V,
                                                         double* __restrict__ dVdC) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n_pts) return;
  // two independent chains per quantity over the rows (fixed order)
  double e = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, e1 = 0.0, gx1 = 0.0, gy1 = 0.0, gz1 = 0.0;
  int r = 0;
  for (; r + 1 < n_rows; r += 2) {
    const size_t o = (size_t)r * ld + i, o1 = o + ld;
    e += partial[o];
    gx += partial[plane + o];
    gy += partial[2 * plane + o];
    gz += partial[3 * plane + o];
    e1 += partial[o1];
    gx1 += partial[plane + o1];
    gy1 += partial[2 * plane + o1];
    gz1 += partial[3 * plane + o1];
  }
  if (r < n_rows) {
    const size_t o = (size_t)r * ld + i;
    e += partial[o];
    gx += partial[plane + o];
    gy += partial[2 * plane + o];
    gz += partial[3 * plane + o];
  }
  e += e1;
  gx += gx1;
  gy += gy1;
  gz += gz1;
  const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
  double nuc = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
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
  const int o = perm ? perm[i] : i;
  V[o] = nuc - e;
  dVdC[3 * o] = nx - gx;
  dVdC[3 * o + 1] = ny - gy;
  dVdC[3 * o + 2] = nz - gz;
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
    rc = fe_sort_points(c, t, d_pts, (int)n, st, &d_pts, &perm);
    if (rc) return rc;
  }
  FwdPlan pl = fe_plan(t, n, c->ppl_L, c->target_wg);
  const size_t plane = (size_t)std::max(pl.rows, 1) * pl.ld;
  rc = fe_ctx_reserve(c, FE_WS_PARTIAL, 4 * plane * sizeof(double));
  if (rc) return rc;
  double* partial = (double*)c->ws[FE_WS_PARTIAL];
  rc = fe_launch_eval(t, pl, d_pts, (int)n, partial, plane, true, precision == FE_PREC_MIXED, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_esp_reduce_grad, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st,
                     partial, pl.rows, pl.ld, plane, d_pts, (int)n, t->d_atoms, t->n_atoms, perm,
                     d_V, d_G);
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
// wave-uniform; partial[(chunk * NT + k) * n_pairs + pair]
template <int L>
__global__ __launch_bounds__(256) void k_dG(const double* __restrict__ rec, const int n_pairs,
                                            const double* __restrict__ pw, const int n_pts,
                                            const int ppc, const double* __restrict__ boys_tab,
                                            double* __restrict__ partial) {
  __shared__ double s_tab[FE_BOYS_NR * FE_BOYS_COLS];
  for (int i = threadIdx.x; i < FE_BOYS_NR * FE_BOYS_COLS; i += 256) s_tab[i] = boys_tab[i];
  __syncthreads();
  constexpr int S = FeL<L>::STRIDE, NT = FeL<L>::NT;
  const int pair = blockIdx.x * 256 + threadIdx.x;
  const int ip = min(pair, n_pairs - 1);
  const double* r = rec + (size_t)ip * S;
  const double Px = r[0], Py = r[1], Pz = r[2], p = r[3], K = r[4];
  const double hp = r[6], cF = r[7];
  double acc[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) acc[k] = 0.0;
  const int q0 = blockIdx.y * ppc;
  const int q1 = min(n_pts, q0 + ppc);
  auto point = [&](const double x, const double y, const double z, double w) {
    const double Rx = x - Px, Ry = y - Py, Rz = z - Pz;
    const double r2 = fma(Rx, Rx, fma(Ry, Ry, Rz * Rz));
    double F[L + 1];
    if (!__any(p * r2 < fe_tfar(L))) {
      const double ri = rsqrt(r2);
      fe_boys_far_scaled<L>(hp * (ri * ri), F);
      w *= cF * ri;
    } else {
      fe_boys<L>(p * r2, F, s_tab);
    }
    fe_dg_accum<L>(Rx, Ry, Rz, F, w, acc);
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
  if (pair < n_pairs) {
    double* o = partial + (size_t)blockIdx.y * NT * n_pairs + pair;
#pragma unroll
    for (int k = 0; k < NT; ++k) o[(size_t)k * n_pairs] = -K * acc[k];
  }
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ x, const int64_t n,
                                              const double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}

// dG[pair * NT + k] = sum over chunks (fixed order); optional flat scatter
template <int L>
__global__ __launch_bounds__(256) void k_dG_reduce(const double* __restrict__ partial,
                                                   const int n_pairs, const int n_chunks,
                                                   double* __restrict__ dG,
                                                   const int* __restrict__ flatoff,
                                                   double* __restrict__ flat) {
  constexpr int NT = FeL<L>::NT;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n_pairs * NT) return;
  const int k = (int)(e / n_pairs), pair = (int)(e - (int64_t)k * n_pairs);
  double s = 0.0;
  for (int c = 0; c < n_chunks; ++c) s += partial[((size_t)c * NT + k) * n_pairs + pair];
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
    const int64_t nL = t->n_L[L];
    d.part_off[L] = d.part_total;
    d.dg_off[L] = d.dg_total;
    if (!nL) continue;
    const int NT = fe_host::NTERM[L];
    const int64_t nblk = (nL + 255) / 256;
    int64_t ch = std::max<int64_t>(1, target_wg / nblk);
    // bound the partial buffer (written and re-read once) and keep >= 128
    // points per chunk so the reduction stays small next to the kernel
    ch = std::min<int64_t>(ch, std::max<int64_t>(1, (16ll << 20) / (nL * NT)));
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
  const int64_t nL = t->n_L[L];
  if (!nL) return;
  const double* tab = t->ctx->d_boys + (size_t)L * FE_BOYS_NR * FE_BOYS_COLS;
  hipLaunchKernelGGL(k_dG<L>, dim3((unsigned)((nL + 255) / 256), d.chunks[L]), dim3(256), 0, st,
                     t->d_rec[L], (int)nL, pw, n, d.ppc[L], tab, part + d.part_off[L]);
  const int64_t ne = nL * FeL<L>::NT;
  hipLaunchKernelGGL(k_dG_reduce<L>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st,
                     part + d.part_off[L], (int)nL, d.chunks[L], dG + d.dg_off[L],
                     t->d_flatoff[L], flat);
}

// ---------------------------------------------------------------------------
// dL/dG -> dL/dP, one workgroup (64 lanes) per shell pair
// ---------------------------------------------------------------------------
struct DPArgs {
  const int* sh_l; const int* sh_fn0; const int* sh_nfn; const int* sh_prim0; const int* sh_nprim;
  const double* sh_c; const double* pexp; const double* pcoef;
  const int* fn_lx; const int* fn_ly; const int* fn_lz;
  const int* sp_i; const int* sp_j; const int* sp_slot;
  const int* slot2kept[FE_LMAX + 1];
  const double* dG[FE_LMAX + 1];
  int n_func;
  double* dP;
  LTabDev lt;
};

__global__ __launch_bounds__(64) void k_dg_to_dp(const DPArgs a) {
  __shared__ int s_ci[10][3], s_cj[10][3];
  __shared__ double s_E[3][4 * 4 * 7];
  __shared__ double s_det[84];
  const int sp = blockIdx.x;
  const int tid = threadIdx.x;
  const int si = a.sp_i[sp], sj = a.sp_j[sp];
  const int li = a.sh_l[si], lj = a.sh_l[sj];
  const int L = li + lj, T = L + 1;
  const int ni = a.sh_nfn[si], nj = a.sh_nfn[sj];
  const int fi = a.sh_fn0[si], fj = a.sh_fn0[sj];
  const int npi = a.sh_nprim[si], npj = a.sh_nprim[sj];
  const int pi0 = a.sh_prim0[si], pj0 = a.sh_prim0[sj];
  const int slot0 = a.sp_slot[sp];
  const int NT = a.lt.key_off[L + 1] - a.lt.key_off[L];
  const int nT = a.lt.mono_off[L + 1] - a.lt.mono_off[L];
  const int* tl = a.lt.tlist + 3 * a.lt.mono_off[L];
  const int* cp = a.lt.m0t_colptr + a.lt.mono_off[L];
  const int k0 = a.lt.key_off[L];
  const int* s2k = a.slot2kept[L];
  const double* dGL = a.dG[L];
  if (tid < ni) {
    s_ci[tid][0] = a.fn_lx[fi + tid];
    s_ci[tid][1] = a.fn_ly[fi + tid];
    s_ci[tid][2] = a.fn_lz[fi + tid];
  }
  if (tid < nj) {
    s_cj[tid][0] = a.fn_lx[fj + tid];
    s_cj[tid][1] = a.fn_ly[fj + tid];
    s_cj[tid][2] = a.fn_lz[fj + tid];
  }
  __syncthreads();
  const double Ax = a.sh_c[3 * si], Ay = a.sh_c[3 * si + 1], Az = a.sh_c[3 * si + 2];
  const double Bx = a.sh_c[3 * sj], By = a.sh_c[3 * sj + 1], Bz = a.sh_c[3 * sj + 2];
  double acc0 = 0.0, acc1 = 0.0;  // entries e = tid and tid + 64 of the ni x nj block
  for (int ii = 0; ii < npi; ++ii)
    for (int jj = 0; jj < npj; ++jj) {
      const int kept = s2k[slot0 + ii * npj + jj];
      if (kept < 0) continue;
      const double al = a.pexp[pi0 + ii], be = a.pexp[pj0 + jj];
      const double p = al + be;
      const double cc = a.pcoef[pi0 + ii] * a.pcoef[pj0 + jj];
      const double Px = (al * Ax + be * Bx) / p;
      const double Py = (al * Ay + be * By) / p;
      const double Pz = (al * Az + be * Bz) / p;
      if (tid < 3) {
        const double xa = (tid == 0 ? Px - Ax : tid == 1 ? Py - Ay : Pz - Az);
        const double xb = (tid == 0 ? Px - Bx : tid == 1 ? Py - By : Pz - Bz);
        fe_e1d_table(s_E[tid], li, lj, xa, xb, p);
      }
      // d etab_t = sum_k dG_k (2 pi/p)(2p)^m M0[k, t]
      const double tp = 2.0 * p, cf = 2.0 * M_PI / p;
      const double* g = dGL + (size_t)kept * NT;
      for (int t = tid; t < nT; t += 64) {
        double s = 0.0;
        for (int q = cp[t]; q < cp[t + 1]; ++q) {
          const int k = a.lt.m0t_k[q];
          double pw = 1.0;
          for (int m = 0; m < a.lt.key_m[k0 + k]; ++m) pw *= tp;
          s += a.lt.m0t_val[q] * pw * g[k];
        }
        s_det[t] = cf * s;
      }
      __syncthreads();
      for (int e = tid, slot = 0; e < ni * nj; e += 64, ++slot) {
        const int mi = e / nj, mj = e - mi * nj;
        const int ex0 = (s_ci[mi][0] * (lj + 1) + s_cj[mj][0]) * T;
        const int ey0 = (s_ci[mi][1] * (lj + 1) + s_cj[mj][1]) * T;
        const int ez0 = (s_ci[mi][2] * (lj + 1) + s_cj[mj][2]) * T;
        double s = 0.0;
        for (int t = 0; t < nT; ++t) {
          const int tx = tl[3 * t], ty = tl[3 * t + 1], tz = tl[3 * t + 2];
          s += s_det[t] * s_E[0][ex0 + tx] * s_E[1][ey0 + ty] * s_E[2][ez0 + tz];
        }
        if (slot == 0) acc0 = fma(cc, s, acc0);
        else acc1 = fma(cc, s, acc1);
      }
      __syncthreads();
    }
  for (int e = tid, slot = 0; e < ni * nj; e += 64, ++slot) {
    const int mi = e / nj, mj = e - mi * nj;
    const double v = slot == 0 ? acc0 : acc1;
    a.dP[(size_t)(fi + mi) * a.n_func + (fj + mj)] = v;
    if (si != sj) a.dP[(size_t)(fj + mj) * a.n_func + (fi + mi)] = v;
  }
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
  fe_ctx* c = t->ctx;
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // staging: points + weights packed as (x, y, z, gV)
  const double* d_pts = nullptr;
  int rc = fe_stage_points(c, pts, std::max<int64_t>(n, 1), ptr_kind, st, &d_pts);
  if (n == 0) d_pts = nullptr;
  if (rc && n > 0) return rc;
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
  if (n > 0)
    hipLaunchKernelGGL(k_pack_pw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_pts,
                       d_gV, (int)n, pw);
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
  if (n == 0) {
    if (d.dg_total)
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((d.dg_total + 255) / 256)), dim3(256), 0, st, dG,
                         (int64_t)d.dg_total, 0.0);
    if (flat && t->n_terms)
      hipLaunchKernelGGL(k_fill, dim3((unsigned)((t->n_terms + 255) / 256)), dim3(256), 0, st,
                         flat, (int64_t)t->n_terms, 0.0);
  } else {
    launch_dG_L<0>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<1>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<2>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<3>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<4>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<5>(t, d, pw, (int)n, part, dG, flat, st);
    launch_dG_L<6>(t, d, pw, (int)n, part, dG, flat, st);
  }
  FE_HIPCHK(hipGetLastError());
  double* d_dP = dLdP;
  if (dLdP) {
    const size_t pb = (size_t)t->n_func * t->n_func * sizeof(double);
    if (ptr_kind == FE_PTR_HOST) {
      rc = fe_ctx_reserve(c, FE_WS_OUT, pb);
      if (rc) return rc;
      d_dP = (double*)c->ws[FE_WS_OUT];
    }
    DPArgs pa{};
    pa.sh_l = t->b.l; pa.sh_fn0 = t->b.fn0; pa.sh_nfn = t->b.nfn; pa.sh_prim0 = t->b.prim0;
    pa.sh_nprim = t->b.nprim; pa.sh_c = t->b.c; pa.pexp = t->b.pexp; pa.pcoef = t->b.pcoef;
    pa.fn_lx = t->b.fn_lx; pa.fn_ly = t->b.fn_ly; pa.fn_lz = t->b.fn_lz;
    pa.sp_i = t->b.sp_i; pa.sp_j = t->b.sp_j; pa.sp_slot = t->b.sp_slot;
    for (int L = 0; L <= FE_LMAX; ++L) {
      pa.slot2kept[L] = t->d_slot2kept[L];
      pa.dG[L] = dG + d.dg_off[L];
    }
    pa.n_func = t->n_func;
    pa.dP = d_dP;
    pa.lt = c->ltab;
    if (t->n_sp > 0) hipLaunchKernelGGL(k_dg_to_dp, dim3(t->n_sp), dim3(64), 0, st, pa);
    FE_HIPCHK(hipGetLastError());
  }
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
