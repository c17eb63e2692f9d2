// Point-independent pair-table build on the device (fe_table_build).
//
// Restates ESPEngine._build/_make_pair (fastesp/esp.py:168-257) and
// build_table (fastesp/table.py:176-259) / build_pairs.cpp:162-425:
//
//  host   : basis functions -> shells (contiguous runs of fn_shell,
//           build_pairs.cpp:170-182); shell pairs si <= sj in reference order;
//           every primitive pair gets a candidate slot in its class-L array.
//  kernel : k_build_pairs, one 64-lane workgroup per shell pair.  Folded
//           density block P_ij (+ P_ji^T), skip if max|blk| < 1e-15; per
//           primitive pair p, P, K (skip K < 1e-30), 1-D Hermite E tables
//           (three lanes, one axis each), etab over the |t| <= L simplex (lanes
//           over t), G = (2 pi/p)(2p)^m M0 etab (lanes over terms), the
//           screening bound (wave reduction), reff and the sum-rule partial.
//  kernel : k_screen, one lane per candidate, points streamed through LDS
//           (k_screen_part + k_screen_decide split the points over blocks
//           when there are few candidates per point):
//           rule 0 (NumPy, table.py:190-194) keep = !(K bound/max(dmin,0.05) < tol)
//           rule 1 (C++, build_pairs.cpp:318-328) keep = !(K bound/max(dmin-reff,0.05) < tol)
//           rule 2 (ESPEngine, esp.py:265-270)    keep = !(K bound/max(dmax-reff,0.05) < tol)
//                  and G not all zero and c_i c_j != 0.
//  compact: kept candidates are packed in candidate (= reference) order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"
#include "fe_hermite.h"

#define FE_BUILD_MAXPRIM 32
#define FE_BUILD_MAXFN 10  // Cartesian components of an f shell

__constant__ double c_mabs[12] = {2.5066282746310002, 2.0, 1.5957691216057308, 4.0,
                                  10.212920779565312, 33.63284766793385, 136.5963625280984,
                                  665.1416330443616, 3768.863619586618, 25506.330595911374,
                                  202331.81842358326, 1872154.2572862155};
__constant__ int c_stride[FE_LMAX + 1];
__constant__ int c_nterm[FE_LMAX + 1];
__constant__ int c_nmono[FE_LMAX + 1];

struct BuildArgs {
  const int* sh_atom; const int* sh_l; const int* sh_fn0; const int* sh_nfn; const int* sh_prim0;
  const int* sh_nprim; const double* sh_c;
  const double* pexp; const double* pcoef;
  const int* fn_lx; const int* fn_ly; const int* fn_lz;
  const int* sp_i; const int* sp_j; const int* sp_slot;
  const double* dens; int n_func;
  double* cand[FE_LMAX + 1];
  double* sp_srule;
  LTabDev lt;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(64) void k_build_pairs(const BuildArgs a) {
  __shared__ double s_blk[FE_BUILD_MAXFN * FE_BUILD_MAXFN];
  __shared__ int s_ci[FE_BUILD_MAXFN][3], s_cj[FE_BUILD_MAXFN][3];
  __shared__ double s_ei[FE_BUILD_MAXPRIM], s_ej[FE_BUILD_MAXPRIM];
  __shared__ double s_wi[FE_BUILD_MAXPRIM], s_wj[FE_BUILD_MAXPRIM];
  __shared__ double s_E[3][4 * 4 * 7];
  __shared__ double s_et[84];
  __shared__ int s_nonzero;

  const int sp = blockIdx.x;
  const int tid = threadIdx.x;
  const int si = a.sp_i[sp], sj = a.sp_j[sp];
  const int li = a.sh_l[si], lj = a.sh_l[sj];
  const int L = li + lj;
  const int ni = a.sh_nfn[si], nj = a.sh_nfn[sj];
  const int fi = a.sh_fn0[si], fj = a.sh_fn0[sj];
  const int npi = a.sh_nprim[si], npj = a.sh_nprim[sj];
  const int S = c_stride[L];
  const int nterm = c_nterm[L];
  const int nT = c_nmono[L];
  double* cand = a.cand[L];
  const int slot0 = a.sp_slot[sp];

  if (tid == 0) s_nonzero = 0;
  __syncthreads();
  for (int idx = tid; idx < ni * nj; idx += 64) {
    const int mi = idx / nj, mj = idx - mi * nj;
    double d = a.dens[(size_t)(fi + mi) * a.n_func + (fj + mj)];
    if (si != sj) d += a.dens[(size_t)(fj + mj) * a.n_func + (fi + mi)];
    s_blk[idx] = d;
    if (!(fabs(d) < 1e-15)) s_nonzero = 1;  // np.abs(blk).max() < 1e-15 -> skip
  }
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
  if (tid < npi) {
    s_ei[tid] = a.pexp[a.sh_prim0[si] + tid];
    s_wi[tid] = a.pcoef[a.sh_prim0[si] + tid];
  }
  if (tid < npj) {
    s_ej[tid] = a.pexp[a.sh_prim0[sj] + tid];
    s_wj[tid] = a.pcoef[a.sh_prim0[sj] + tid];
  }
  __syncthreads();
  if (!s_nonzero) {
    for (int s = tid; s < npi * npj; s += 64) cand[(size_t)(slot0 + s) * S + 7] = 0.0;
    if (tid == 0) a.sp_srule[sp] = 0.0;
    return;
  }
  const double Ax = a.sh_c[3 * si], Ay = a.sh_c[3 * si + 1], Az = a.sh_c[3 * si + 2];
  const double Bx = a.sh_c[3 * sj], By = a.sh_c[3 * sj + 1], Bz = a.sh_c[3 * sj + 2];
  const double dx = Ax - Bx, dy = Ay - By, dz = Az - Bz;
  const double ab2 = dx * dx + dy * dy + dz * dz;
  const int T = L + 1;
  const int* tl = a.lt.tlist + 3 * a.lt.mono_off[L];
  const int k0 = a.lt.key_off[L];
  double srule = 0.0;

  for (int ii = 0; ii < npi; ++ii) {
    for (int jj = 0; jj < npj; ++jj) {
      double* rec = cand + (size_t)(slot0 + ii * npj + jj) * S;
      const double al = s_ei[ii], be = s_ej[jj];
      const double p = al + be;
      const double K = exp(-al * be / p * ab2);
      if (K < 1e-30) {
        if (tid == 0) rec[7] = 0.0;
        continue;
      }
      const double cc = s_wi[ii] * s_wj[jj];
      const double Px = (al * Ax + be * Bx) / p;
      const double Py = (al * Ay + be * By) / p;
      const double Pz = (al * Az + be * Bz) / p;
      if (tid < 3) {
        // 1-D Hermite coefficients of axis `tid`
        const double xa = (tid == 0 ? Px - Ax : tid == 1 ? Py - Ay : Pz - Az);
        const double xb = (tid == 0 ? Px - Bx : tid == 1 ? Py - By : Pz - Bz);
        fe_e1d_table(s_E[tid], li, lj, xa, xb, p);
      }
      __syncthreads();
      // etab over the |t| <= L simplex
      for (int t = tid; t < nT; t += 64) {
        const int tx = tl[3 * t], ty = tl[3 * t + 1], tz = tl[3 * t + 2];
        double s = 0.0;
        for (int mi = 0; mi < ni; ++mi) {
          const int ax = s_ci[mi][0], ay = s_ci[mi][1], az = s_ci[mi][2];
          for (int mj = 0; mj < nj; ++mj) {
            const double w = s_blk[mi * nj + mj];
            if (w == 0.0) continue;
            const double ex = s_E[0][(ax * (lj + 1) + s_cj[mj][0]) * T + tx];
            const double ey = s_E[1][(ay * (lj + 1) + s_cj[mj][1]) * T + ty];
            const double ez = s_E[2][(az * (lj + 1) + s_cj[mj][2]) * T + tz];
            s += w * ex * ey * ez;
          }
        }
        s_et[t] = s * cc;
      }
      __syncthreads();
      const double tp = 2.0 * p;
      const double cf = 2.0 * M_PI / p;
      int anyg = 0;
      for (int k = tid; k < nterm; k += 64) {
        const int r = k0 + k;
        double s = 0.0;
        for (int q = a.lt.m0_rowptr[r]; q < a.lt.m0_rowptr[r + 1]; ++q)
          s += a.lt.m0_val[q] * s_et[a.lt.m0_tidx[q]];
        double pw = 1.0;
        for (int m = 0; m < a.lt.key_m[r]; ++m) pw *= tp;
        const double g = cf * pw * s;
        rec[8 + k] = g;
        anyg |= (g != 0.0);
      }
      const double sq = sqrt(tp);
      double b = 0.0;
      for (int t = tid; t < nT; t += 64) {
        const int tx = tl[3 * t], ty = tl[3 * t + 1], tz = tl[3 * t + 2];
        double pw = 1.0;
        for (int q = 0; q < tx + ty + tz; ++q) pw *= sq;
        b += fabs(s_et[t]) * (c_mabs[tx] * c_mabs[ty] * c_mabs[tz]) * pw;
      }
      b = wave_sum(b);
      anyg = __any(anyg);
      if (tid == 0) {
        const double bound = pow(tp, -1.5) * b;
        rec[0] = Px;
        rec[1] = Py;
        rec[2] = Pz;
        rec[3] = p;
        rec[4] = K;
        rec[5] = K * bound;
        rec[6] = 1.6 * sqrt((L + 2.0) / p);
        rec[7] = 1.0 + (anyg ? 2.0 : 0.0) + (cc != 0.0 ? 4.0 : 0.0);
        srule += K * pow(M_PI / p, 1.5) * s_et[0];
      }
      for (int k = nterm + tid; k < S - 8; k += 64) rec[8 + k] = 0.0;
      __syncthreads();
    }
  }
  if (tid == 0) a.sp_srule[sp] = srule;
}

// Smallest (rule 0, 1) or largest (rule 2) |C - P|^2 over the points
// [q0, q1), staged through `s_p` 1024 at a time by the whole block.
__device__ __forceinline__ double fe_screen_d2(const double Px, const double Py, const double Pz,
                                               const double* __restrict__ pts, const int q0,
                                               const int q1, const bool use_max, double* s_p) {
  double d2sel = use_max ? -INFINITY : INFINITY;
  for (int base = q0; base < q1; base += 1024) {
    const int m = min(1024, q1 - base);
    __syncthreads();
    for (int q = threadIdx.x; q < 3 * m; q += 256) s_p[q] = pts[3 * (size_t)base + q];
    __syncthreads();
    if (use_max) {
      for (int q = 0; q < m; ++q) {
        const double x = s_p[3 * q] - Px, y = s_p[3 * q + 1] - Py, z = s_p[3 * q + 2] - Pz;
        d2sel = fmax(d2sel, x * x + y * y + z * z);
      }
    } else {
      for (int q = 0; q < m; ++q) {
        const double x = s_p[3 * q] - Px, y = s_p[3 * q + 1] - Py, z = s_p[3 * q + 2] - Pz;
        d2sel = fmin(d2sel, x * x + y * y + z * z);
      }
    }
  }
  return d2sel;
}

// The screen's decision for candidate record r given the selected squared
// distance: 1 kept, 0 a candidate the screen dropped, -1 no candidate
// (K < 1e-30 or an empty density block: not in the reference's primitive-pair
// list).
__device__ __forceinline__ int fe_screen_keep(const double* __restrict__ r, const double d2sel,
                                              const double tol, const int rule) {
  const int flags = (int)r[7];
  int k = flags & 1;
  if (k) {
    const double kb = r[5];
    double dd = sqrt(fmax(d2sel, 0.0));
    if (rule == 1) {
      dd -= r[6];
      k = (flags & 2) && (flags & 4);
    } else if (rule == 2) {
      dd -= r[6];
    }
    if (dd < 0.05) dd = 0.05;
    if (kb / dd < tol) k = 0;
  }
  return (flags & 1) ? k : -1;
}

// One lane per candidate, every point (small point sets).
__global__ __launch_bounds__(256) void k_screen(const double* __restrict__ rec, const int n_cand,
                                                const int S, const double* __restrict__ pts,
                                                const int n_pts, const double tol, const int rule,
                                                int* __restrict__ keep) {
  __shared__ double s_p[3 * 1024];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double* r = rec + (size_t)min(i, max(n_cand - 1, 0)) * S;
  const double d2sel = fe_screen_d2(r[0], r[1], r[2], pts, 0, n_pts, rule == 2, s_p);
  if (i < n_cand) keep[i] = fe_screen_keep(r, d2sel, tol, rule);
}

// One lane per candidate over the point slice blockIdx.y: the slice's
// selected squared distance is combined into d2[i] by an integer atomic min
// (max for rule 2) on its bit pattern, which orders like the value for
// non-negative doubles.  d2 starts at +inf (0 for rule 2).  Splitting the
// points over blocks fills the device for the classes with few candidates.
__global__ __launch_bounds__(256) void k_screen_part(const double* __restrict__ rec,
                                                     const int n_cand, const int S,
                                                     const double* __restrict__ pts,
                                                     const int n_pts, const int slice,
                                                     const int rule,
                                                     unsigned long long* __restrict__ d2) {
  __shared__ double s_p[3 * 1024];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double* r = rec + (size_t)min(i, max(n_cand - 1, 0)) * S;
  const int q0 = blockIdx.y * slice, q1 = min(n_pts, q0 + slice);
  const double v = fe_screen_d2(r[0], r[1], r[2], pts, q0, q1, rule == 2, s_p);
  if (i >= n_cand || q0 >= q1) return;
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  if (rule == 2) atomicMax(d2 + i, b);
  else atomicMin(d2 + i, b);
}

__global__ __launch_bounds__(256) void k_screen_decide(const double* __restrict__ rec,
                                                       const int n_cand, const int S,
                                                       const unsigned long long* __restrict__ d2,
                                                       const double tol, const int rule,
                                                       int* __restrict__ keep) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_cand) return;
  keep[i] = fe_screen_keep(rec + (size_t)i * S, __longlong_as_double((long long)d2[i]), tol, rule);
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ a, const int n,
                                                  const unsigned long long v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ src, const int S,
                                                const int* __restrict__ idx, const int n,
                                                double* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * S) return;
  const int r = (int)(e / S), c = (int)(e - (int64_t)r * S);
  const double* s = src + (size_t)idx[r] * S;
  // slots 6/7 of a candidate hold reff and the screen flags; the compacted
  // record keeps the per-pair far-field constants there instead: 1/(2p) and
  // K sqrt(pi/p)/2
  dst[e] = c == 6 ? 0.5 / s[3] : c == 7 ? s[4] * (0.88622692545275801365 / sqrt(s[3])) : s[c];
}

// the scaled coefficients c g_k of each compacted record of class L (one lane
// per record)
__global__ __launch_bounds__(256) void k_fill_dcoef(double* __restrict__ rec, const int S,
                                                    const int n, const int L) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) fe_fill_dcoef_any(L, rec + (size_t)i * S + 8);
}

// ---------------------------------------------------------------------------
namespace {
struct HostShell {
  int atom, l, fn0, nfn, prim0, nprim;
  double c[3];
};

template <class T>
size_t push_bytes(std::vector<char>& buf, const T* p, size_t n) {
  size_t off = (buf.size() + 255) & ~(size_t)255;
  buf.resize(off + n * sizeof(T));
  if (n) memcpy(buf.data() + off, p, n * sizeof(T));
  return off;
}
}  // namespace

extern "C" int fe_table_build(fe_ctx* c, int n_func, int n_prim, const int32_t* fn_atom,
                              const int32_t* fn_nprim, const int32_t* fn_off,
                              const int32_t* fn_lx, const int32_t* fn_ly, const int32_t* fn_lz,
                              const int32_t* fn_shell, const double* fc, const double* pexp,
                              const double* pcoef, const double* dens, int n_points,
                              const double* pts, int n_atoms, const double* Z, const double* xyz,
                              double tol, int rule, int ptr_kind, void* stream, fe_table** out) {
  if (!c || !out) return fe_set_error(FE_ERR_ARG, "fe_table_build: null argument");
  *out = nullptr;
  if (n_func <= 0 || n_prim <= 0 || n_points < 0 || n_atoms < 0)
    return fe_set_error(FE_ERR_ARG, "fe_table_build: bad sizes");
  if (!fn_atom || !fn_nprim || !fn_off || !fn_lx || !fn_ly || !fn_lz || !fn_shell || !fc ||
      !pexp || !pcoef || !dens || (n_points && !pts) || (n_atoms && (!Z || !xyz)))
    return fe_set_error(FE_ERR_ARG, "fe_table_build: null array");
  if (rule < 0 || rule > 2) return fe_set_error(FE_ERR_ARG, "fe_table_build: rule must be 0, 1 or 2");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_table_build: bad ptr_kind");
  if (!(tol >= 0.0)) return fe_set_error(FE_ERR_ARG, "fe_table_build: tol must be >= 0");

  // ---- shells (contiguous runs of equal fn_shell, build_pairs.cpp:170-182)
  std::vector<HostShell> sh;
  for (int i = 0; i < n_func;) {
    int j = i;
    while (j < n_func && fn_shell[j] == fn_shell[i]) ++j;
    HostShell s;
    s.atom = fn_atom[i];
    s.l = fn_lx[i] + fn_ly[i] + fn_lz[i];
    s.fn0 = i;
    s.nfn = j - i;
    s.prim0 = fn_off[i];
    s.nprim = fn_nprim[i];
    for (int k = 0; k < 3; ++k) s.c[k] = fc[3 * i + k];
    if (s.l > 3 || 2 * s.l > FE_LMAX)
      return fe_set_error(FE_ERR_UNSUPPORTED, "shell at function %d has l=%d (max %d)", i, s.l,
                          std::min(3, FE_LMAX / 2));
    if (s.nfn != (s.l + 1) * (s.l + 2) / 2)
      return fe_set_error(FE_ERR_FORMAT, "shell at function %d: %d components for l=%d", i,
                          s.nfn, s.l);
    if (s.nprim < 1 || s.nprim > FE_BUILD_MAXPRIM || s.prim0 < 0 || s.prim0 + s.nprim > n_prim)
      return fe_set_error(FE_ERR_FORMAT, "shell at function %d: bad primitive range", i);
    for (int q = i; q < j; ++q) {
      if (fn_lx[q] + fn_ly[q] + fn_lz[q] != s.l || fn_nprim[q] != s.nprim)
        return fe_set_error(FE_ERR_FORMAT, "function %d is inconsistent with its shell", q);
      for (int k = 0; k < s.nprim; ++k)
        if (pexp[fn_off[q] + k] != pexp[s.prim0 + k] || pcoef[fn_off[q] + k] != pcoef[s.prim0 + k])
          return fe_set_error(FE_ERR_FORMAT, "function %d: primitives differ within its shell", q);
    }
    sh.push_back(s);
    i = j;
  }
  const int nsh = (int)sh.size();
  // ---- shell pairs and candidate slots per class
  std::vector<int> sp_i, sp_j, sp_slot;
  int64_t ncand[FE_LMAX + 1] = {};
  for (int i = 0; i < nsh; ++i)
    for (int j = i; j < nsh; ++j) {
      const int L = sh[i].l + sh[j].l;
      sp_i.push_back(i);
      sp_j.push_back(j);
      sp_slot.push_back((int)ncand[L]);
      ncand[L] += (int64_t)sh[i].nprim * sh[j].nprim;
      if (ncand[L] > INT32_MAX / 2)
        return fe_set_error(FE_ERR_UNSUPPORTED, "too many primitive pairs in class %d", L);
    }
  const int nsp = (int)sp_i.size();

  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  fe_table* t = new fe_table();
  t->ctx = c;
  t->has_basis = true;
  t->n_func = n_func;
  t->n_shells = nsh;
  t->n_sp = nsp;
  t->h_sp_i = sp_i;
  t->h_sp_j = sp_j;
  t->h_sp_slot.assign(sp_slot.begin(), sp_slot.end());
  auto fail = [&](int rc) {
    fe_table_free(t);
    return rc;
  };

  // ---- basis + shell pairs to the device (one buffer)
  std::vector<int> v_atom(nsh), v_l(nsh), v_fn0(nsh), v_nfn(nsh), v_p0(nsh), v_np(nsh);
  std::vector<double> v_c(3 * nsh);
  for (int i = 0; i < nsh; ++i) {
    v_atom[i] = sh[i].atom; v_l[i] = sh[i].l; v_fn0[i] = sh[i].fn0; v_nfn[i] = sh[i].nfn;
    v_p0[i] = sh[i].prim0; v_np[i] = sh[i].nprim;
    for (int k = 0; k < 3; ++k) v_c[3 * i + k] = sh[i].c[k];
  }
  std::vector<char> hb;
  const size_t o_atom = push_bytes(hb, v_atom.data(), nsh), o_l = push_bytes(hb, v_l.data(), nsh),
               o_fn0 = push_bytes(hb, v_fn0.data(), nsh), o_nfn = push_bytes(hb, v_nfn.data(), nsh),
               o_p0 = push_bytes(hb, v_p0.data(), nsh), o_np = push_bytes(hb, v_np.data(), nsh),
               o_c = push_bytes(hb, v_c.data(), 3 * (size_t)nsh),
               o_pexp = push_bytes(hb, pexp, n_prim), o_pcoef = push_bytes(hb, pcoef, n_prim),
               o_lx = push_bytes(hb, fn_lx, n_func), o_ly = push_bytes(hb, fn_ly, n_func),
               o_lz = push_bytes(hb, fn_lz, n_func), o_spi = push_bytes(hb, sp_i.data(), nsp),
               o_spj = push_bytes(hb, sp_j.data(), nsp),
               o_sps = push_bytes(hb, sp_slot.data(), nsp);
  if (hipMalloc(&t->d_basis_mem, hb.size()) != hipSuccess) {
    (void)hipGetLastError();
    return fail(fe_set_error(FE_ERR_ALLOC, "fe_table_build: basis allocation failed"));
  }
  if (hipMemcpy(t->d_basis_mem, hb.data(), hb.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: basis upload failed"));
  t->dev_bytes += hb.size();
  const char* bm = (const char*)t->d_basis_mem;
  t->b.atom = (const int*)(bm + o_atom); t->b.l = (const int*)(bm + o_l);
  t->b.fn0 = (const int*)(bm + o_fn0); t->b.nfn = (const int*)(bm + o_nfn);
  t->b.prim0 = (const int*)(bm + o_p0); t->b.nprim = (const int*)(bm + o_np);
  t->b.c = (const double*)(bm + o_c); t->b.pexp = (const double*)(bm + o_pexp);
  t->b.pcoef = (const double*)(bm + o_pcoef); t->b.fn_lx = (const int*)(bm + o_lx);
  t->b.fn_ly = (const int*)(bm + o_ly); t->b.fn_lz = (const int*)(bm + o_lz);
  t->b.sp_i = (const int*)(bm + o_spi); t->b.sp_j = (const int*)(bm + o_spj);
  t->b.sp_slot = (const int*)(bm + o_sps);

  // ---- scratch in the context's build workspace (grown when needed and kept
  // for the next build, so a build allocates only its table): density and
  // points of host inputs, candidates, keep flags, screen distances, sum-rule
  // partials and gather indices; the keep flags come back through pinned
  // host memory, so their copies do not hold up the launches
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  int64_t cand_off[FE_LMAX + 2] = {};
  for (int L = 0; L <= FE_LMAX; ++L) cand_off[L + 1] = cand_off[L] + ncand[L];
  const int64_t ncand_all = cand_off[FE_LMAX + 1];
  const bool host_in = ptr_kind == FE_PTR_HOST;
  size_t wo = 0;
  const size_t o_dens = wo;
  if (host_in) wo += al((size_t)n_func * n_func * sizeof(double));
  const size_t o_pts = wo;
  if (host_in) wo += al((size_t)n_points * 3 * sizeof(double));
  size_t o_cand[FE_LMAX + 1], o_keep[FE_LMAX + 1], o_idx[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_cand[L] = wo;
    wo += al((size_t)ncand[L] * fe_host::STRIDE[L] * sizeof(double));
    o_keep[L] = wo;
    wo += al((size_t)ncand[L] * sizeof(int));
    o_idx[L] = wo;
    wo += al((size_t)ncand[L] * sizeof(int));
  }
  const size_t o_d2 = wo;
  wo += al((size_t)std::max<int64_t>(ncand_all, 1) * sizeof(unsigned long long));
  const size_t o_srule = wo;
  wo += al((size_t)std::max(nsp, 1) * sizeof(double));
  if (fe_ctx_reserve(c, FE_WS_BUILD, wo) != FE_OK ||
      fe_ctx_reserve_pinned(c, (size_t)std::max<int64_t>(ncand_all, 1) * sizeof(int)) != FE_OK)
    return fail(FE_ERR_ALLOC);
  char* wb = (char*)c->ws[FE_WS_BUILD];
  int* h_keep = (int*)c->h_pin;

  // ---- density and points
  const double* d_dens = dens;
  const double* d_pts = pts;
  if (host_in) {
    (void)hipMemcpyAsync(wb + o_dens, dens, (size_t)n_func * n_func * sizeof(double),
                         hipMemcpyHostToDevice, st);
    d_dens = (const double*)(wb + o_dens);
    if (n_points) {
      (void)hipMemcpyAsync(wb + o_pts, pts, (size_t)n_points * 3 * sizeof(double),
                           hipMemcpyHostToDevice, st);
      d_pts = (const double*)(wb + o_pts);
    }
  }

  // ---- candidates
  static bool consts_done[64] = {};
  if (!consts_done[c->device & 63]) {
    FE_HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_stride), fe_host::STRIDE, sizeof(int) * (FE_LMAX + 1)));
    FE_HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_nterm), fe_host::NTERM, sizeof(int) * (FE_LMAX + 1)));
    FE_HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_nmono), fe_host::NMONO, sizeof(int) * (FE_LMAX + 1)));
    consts_done[c->device & 63] = true;
  }
  double* cand[FE_LMAX + 1] = {};
  int* keep[FE_LMAX + 1] = {};
  double* d_srule = (double*)(wb + o_srule);
  unsigned long long* d_d2 = (unsigned long long*)(wb + o_d2);
  hipEvent_t ev[FE_LMAX + 1] = {};  // class L's keep flags are on the host
  // every exit after the first launch has synchronised the stream first
  auto free_cand = [&]() {
    for (int L = 0; L <= FE_LMAX; ++L)
      if (ev[L]) (void)hipEventDestroy(ev[L]);
  };
  for (int L = 0; L <= FE_LMAX; ++L) {
    t->n_cand_L[L] = ncand[L];
    cand[L] = (double*)(wb + o_cand[L]);
    keep[L] = (int*)(wb + o_keep[L]);
    if (!ncand[L]) continue;
    if (hipEventCreateWithFlags(&ev[L], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(st);
      free_cand();
      return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: event creation failed"));
    }
  }
  BuildArgs ba{};
  ba.sh_atom = t->b.atom; ba.sh_l = t->b.l; ba.sh_fn0 = t->b.fn0; ba.sh_nfn = t->b.nfn;
  ba.sh_prim0 = t->b.prim0; ba.sh_nprim = t->b.nprim; ba.sh_c = t->b.c;
  ba.pexp = t->b.pexp; ba.pcoef = t->b.pcoef;
  ba.fn_lx = t->b.fn_lx; ba.fn_ly = t->b.fn_ly; ba.fn_lz = t->b.fn_lz;
  ba.sp_i = t->b.sp_i; ba.sp_j = t->b.sp_j; ba.sp_slot = t->b.sp_slot;
  ba.dens = d_dens; ba.n_func = n_func;
  for (int L = 0; L <= FE_LMAX; ++L) ba.cand[L] = cand[L];
  ba.sp_srule = d_srule;
  ba.lt = c->ltab;
  if (nsp > 0) hipLaunchKernelGGL(k_build_pairs, dim3(nsp), dim3(64), 0, st, ba);
  // each class's keep flags go to the host as soon as its screen ends
  int* hk[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) hk[L] = h_keep + cand_off[L];
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!ncand[L]) continue;
    const unsigned nb = (unsigned)((ncand[L] + 255) / 256);
    // point slices of at least 1024 points, enough for about 2048 blocks
    const int max_split = (n_points + 1023) / 1024;
    const int ns = std::min<int>(max_split, (int)std::max<unsigned>(1, (2048 + nb - 1) / nb));
    if (ns < 2) {
      hipLaunchKernelGGL(k_screen, dim3(nb), dim3(256), 0, st, cand[L], (int)ncand[L],
                         fe_host::STRIDE[L], d_pts, n_points, tol, rule, keep[L]);
    } else {
      const int slice = (n_points + ns - 1) / ns;
      unsigned long long* d2 = d_d2 + cand_off[L];
      hipLaunchKernelGGL(k_fill_u64, dim3(nb), dim3(256), 0, st, d2, (int)ncand[L],
                         rule == 2 ? 0ull : 0x7FF0000000000000ull);
      hipLaunchKernelGGL(k_screen_part, dim3(nb, (unsigned)ns), dim3(256), 0, st, cand[L],
                         (int)ncand[L], fe_host::STRIDE[L], d_pts, n_points, slice, rule, d2);
      hipLaunchKernelGGL(k_screen_decide, dim3(nb), dim3(256), 0, st, cand[L], (int)ncand[L],
                         fe_host::STRIDE[L], d2, tol, rule, keep[L]);
    }
    (void)hipMemcpyAsync(hk[L], keep[L], sizeof(int) * ncand[L], hipMemcpyDeviceToHost, st);
    (void)hipEventRecord(ev[L], st);
  }
  if (hipGetLastError() != hipSuccess) {
    (void)hipStreamSynchronize(st);
    free_cand();
    return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: kernel launch failed"));
  }
  std::vector<double> srule(std::max(nsp, 1));
  (void)hipMemcpyAsync(srule.data(), d_srule, sizeof(double) * std::max(nsp, 1),
                       hipMemcpyDeviceToHost, st);
  // ---- compaction in candidate order, one host thread per class, each
  // started when its keep flags have arrived (the later classes are still
  // being screened); the device copies and gathers follow in one pass
  // global candidate order: shell pairs in order, primitives ii-major
  std::vector<int64_t> sp_global(nsp + 1, 0);
  for (int s = 0; s < nsp; ++s)
    sp_global[s + 1] = sp_global[s] + (int64_t)sh[sp_i[s]].nprim * sh[sp_j[s]].nprim;
  struct ClassOut {
    std::vector<int> idx, s2k;
    std::vector<int32_t> orig;
    int64_t n_cand = 0;
  };
  std::vector<ClassOut> co(FE_LMAX + 1);
  auto compact = [&](const int L) {
    std::vector<int>& idx = co[L].idx;
    std::vector<int>& s2k = co[L].s2k;
    s2k.assign(ncand[L], -1);
    for (int64_t q = 0; q < ncand[L]; ++q) {
      if (hk[L][q] > 0) idx.push_back((int)q);
      if (hk[L][q] >= 0) ++co[L].n_cand;
    }
    // original (reference-order) position of every kept record, for export:
    // candidate slot -> global candidate index via its shell pair (slots of
    // one class are contiguous per shell pair in shell-pair order, primitives
    // ii-major); with pair ordering on, also each record's (Px, Py, Pz, p)
    // from the basis on the host
    const bool order = c->pair_order && idx.size() > 1;
    std::vector<int32_t> orig(idx.size());
    std::vector<double> head(order ? idx.size() * 4 : 0);
    {
      std::vector<int> sps;
      for (int s = 0; s < nsp; ++s)
        if (sh[sp_i[s]].l + sh[sp_j[s]].l == L) sps.push_back(s);
      size_t cur = 0;
      for (size_t k = 0; k < idx.size(); ++k) {
        while (cur + 1 < sps.size() && sp_slot[sps[cur + 1]] <= idx[k]) ++cur;
        const int s = sps[cur];
        const int64_t local = idx[k] - sp_slot[s];
        orig[k] = (int32_t)(sp_global[s] + local);
        if (order) {
          const HostShell &A = sh[sp_i[s]], &B = sh[sp_j[s]];
          const int ii = (int)(local / B.nprim), jj = (int)(local - (int64_t)ii * B.nprim);
          const double al = pexp[A.prim0 + ii], be = pexp[B.prim0 + jj], pp = al + be;
          for (int a = 0; a < 3; ++a) head[4 * k + a] = (al * A.c[a] + be * B.c[a]) / pp;
          head[4 * k + 3] = pp;
        }
      }
    }
    if (order) {
      const std::vector<int> ord = fe_pair_order(head.data(), 4, (int64_t)idx.size());
      std::vector<int> idx2(idx.size());
      std::vector<int32_t> orig2(idx.size());
      for (size_t k = 0; k < idx.size(); ++k) {
        idx2[k] = idx[ord[k]];
        orig2[k] = orig[ord[k]];
      }
      idx.swap(idx2);
      orig.swap(orig2);
    }
    for (size_t k = 0; k < idx.size(); ++k) s2k[idx[k]] = (int)k;
    co[L].orig = std::move(orig);
  };
  std::vector<std::thread> workers;
  bool screened = true;
  for (int L = 0; L <= FE_LMAX && screened; ++L) {
    if (!ncand[L]) continue;
    if (hipEventSynchronize(ev[L]) != hipSuccess) {
      screened = false;
      break;
    }
    try {
      workers.emplace_back(compact, L);
    } catch (...) {
      compact(L);  // no thread available: compact on this one
    }
  }
  for (auto& w : workers) w.join();
  if (!screened || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipStreamSynchronize(st);
    free_cand();
    return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: build kernels failed: %s",
                             hipGetErrorString(hipGetLastError())));
  }
  double sr = 0.0;
  for (int s = 0; s < nsp; ++s) sr += srule[s];
  t->sum_rule = sr;
  int64_t total = 0;
  // the copies below read the host vectors of co: synchronise before leaving
  auto free_idx = [&]() { (void)hipStreamSynchronize(st); };
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!ncand[L]) continue;
    const std::vector<int>& idx = co[L].idx;
    t->n_candidates += co[L].n_cand;
    t->n_L[L] = (int64_t)idx.size();
    total += (int64_t)idx.size();
    if (hipMalloc(&t->d_slot2kept[L], sizeof(int) * ncand[L]) != hipSuccess) {
      (void)hipGetLastError();
      free_idx();
      free_cand();
      return fail(fe_set_error(FE_ERR_ALLOC, "fe_table_build: allocation failed"));
    }
    (void)hipMemcpyAsync(t->d_slot2kept[L], co[L].s2k.data(), sizeof(int) * ncand[L],
                         hipMemcpyHostToDevice, st);
    t->dev_bytes += sizeof(int) * ncand[L];
    if (idx.empty()) continue;
    const int S = fe_host::STRIDE[L];
    int* d_idx = (int*)(wb + o_idx[L]);
    if (hipMalloc(&t->d_rec[L], (size_t)idx.size() * S * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      free_idx();
      free_cand();
      return fail(fe_set_error(FE_ERR_ALLOC, "fe_table_build: table allocation failed"));
    }
    t->dev_bytes += (size_t)idx.size() * S * sizeof(double);
    (void)hipMemcpyAsync(d_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, st);
    const int64_t ne = (int64_t)idx.size() * S;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, cand[L], S,
                       d_idx, (int)idx.size(), t->d_rec[L]);
    if (fe_host::NDCOEF[L])
      hipLaunchKernelGGL(k_fill_dcoef, dim3((unsigned)((idx.size() + 255) / 256)), dim3(256), 0,
                         st, t->d_rec[L], S, (int)idx.size(), L);
  }
  free_idx();
  for (int L = 0; L <= FE_LMAX; ++L) t->orig[L] = std::move(co[L].orig);
  t->n_pairs = (int)total;
  int64_t nterms = 0, nvalid = 0;
  for (int L = 0; L <= FE_LMAX; ++L) nterms += t->n_L[L] * fe_host::NTERM[L];
  t->n_terms = (int)nterms;
  (void)nvalid;
  free_cand();
  if (hipGetLastError() != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: compaction failed"));
  int rc = fe_table_set_atoms(t, n_atoms, Z, xyz);
  if (rc) return fail(rc);
  rc = fe_table_index(t);
  if (rc) return fail(rc);
  *out = t;
  return FE_OK;
}

extern "C" int fe_table_stats(const fe_table* t, double* out, int n) {
  if (!t || !out) return fe_set_error(FE_ERR_ARG, "fe_table_stats: null argument");
  double v[4] = {t->sum_rule, (double)t->n_pairs, (double)t->n_func, (double)t->n_sp};
  for (int i = 0; i < n && i < 4; ++i) out[i] = v[i];
  return FE_OK;
}
