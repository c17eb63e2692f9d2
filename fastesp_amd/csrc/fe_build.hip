// Point-independent pair-table build on the device (fe_table_build) and its
// in-place, allocation-free repeat for new densities or points
// (fe_table_rebuild).
//
// Restates ESPEngine._build/_make_pair (fastesp/esp.py:168-257) and
// build_table (fastesp/table.py:176-259) / build_pairs.cpp:162-425:
//
//  host   : basis functions -> shells (contiguous runs of fn_shell,
//           build_pairs.cpp:170-182); shell pairs si <= sj in reference order;
//           every primitive pair gets a candidate slot in its class-L array
//           and its reference-order index (shell pairs in order, primitives
//           i-major).
//  kernel : k_build_pairs, one 64-lane workgroup per shell pair.  Folded
//           density block P_ij (+ P_ji^T), skip if max|blk| < 1e-15; per
//           primitive pair p, P, K (skip K < 1e-30), 1-D Hermite E tables
//           (three lanes, one axis each), etab over the |t| <= L simplex (lanes
//           over t), G = (2 pi/p)(2p)^m M0 etab (lanes over terms), the
//           screening bound (wave reduction), reff and the sum-rule partial.
//  kernel : the screen (k_screen_boxes), one lane per candidate.  The points
//           are Morton-sorted and cut into boxes of 64; the rules
//           rule 0 (NumPy, table.py:190-194) keep = !(K bound/max(dmin,0.05) < tol)
//           rule 1 (C++, build_pairs.cpp:318-328) keep = !(K bound/max(dmin-reff,0.05) < tol)
//                  and G not all zero and c_i c_j != 0
//           rule 2 (ESPEngine, esp.py:265-270)    keep = !(K bound/max(dmax-reff,0.05) < tol)
//           are decided exactly without visiting every point: the test is
//           monotone in the distance, so rules 0/1 keep a candidate iff some
//           point passes it and rule 2 iff every point does, and a box whose
//           distance bounds settle the test for all its points is decided
//           (or skipped) without reading them.
//  device : compaction and ordering.  Each candidate gets a 32-bit key
//           (class, dropped flag, exponent bucket, Morton code of P), the keys
//           are radix-sorted (stable), and the kept records of class L are the
//           first ones of its range; a prefix sum over the candidates in
//           reference order gives each kept record its flat term offset and
//           its reference-order index.  Nothing goes back to the host but the
//           counts (fe_table_build, one synchronisation to size the table);
//           fe_table_rebuild does not synchronise at all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"
#include "fe_hermite.h"

#define FE_BUILD_MAXPRIM 32
#define FE_BUILD_MAXFN 10  // Cartesian components of an f shell
#define FE_BOX 64          // sorted points per screening box
#define FE_BOX_CHUNK 256   // boxes staged in LDS at a time

// \int |He_n| e^{-x^2/2} dx as tabulated by the reference (fastesp/esp.py:25-28)
__constant__ double c_mabs[12] = {2.5066282746310002, 2.0, 1.5957691216057308, 4.0,
                                  10.212920779565312, 33.63284766793385, 136.5963625280984,
                                  665.1416330443616, 3768.863619586618, 25506.330595911374,
                                  202331.81842358326, 1872154.2572862155};

struct ClassInfo {
  int stride[FE_LMAX + 1], nterm[FE_LMAX + 1], nmono[FE_LMAX + 1];
  int64_t cand_off[FE_LMAX + 2];  // candidate slots of class L: [cand_off[L], cand_off[L+1])
};

struct BuildArgs {
  const int* sh_atom; const int* sh_l; const int* sh_fn0; const int* sh_nfn; const int* sh_prim0;
  const int* sh_nprim; const double* sh_c;
  const double* pexp; const double* pcoef;
  const int* fn_lx; const int* fn_ly; const int* fn_lz;
  const int* sp_i; const int* sp_j; const int* sp_slot; const int* sp_gidx;
  const double* dens; int n_func;
  double* cand[FE_LMAX + 1];
  int* gidx;  // per candidate (class-concatenated slot): reference-order index
  double* sp_srule;
  LTabDev lt;
  ClassInfo ci;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Candidate record slots 0-7: Px Py Pz p K K*bound reff flags, then the G
// terms.  flags: 1 candidate of the reference's pair list (K >= 1e-30 and a
// non-empty density block), 2 some G non-zero, 4 c_i c_j != 0, 8 K >= 1e-30
// (independent of the density: the capacity of a rebuildable table).
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
  const int S = a.ci.stride[L];
  const int nterm = a.ci.nterm[L];
  const int nT = a.ci.nmono[L];
  double* cand = a.cand[L];
  const int slot0 = a.sp_slot[sp];

  // reference-order index of every slot of this shell pair
  for (int s = tid; s < npi * npj; s += 64)
    a.gidx[a.ci.cand_off[L] + slot0 + s] = a.sp_gidx[sp] + s;
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
  const double Ax = a.sh_c[3 * si], Ay = a.sh_c[3 * si + 1], Az = a.sh_c[3 * si + 2];
  const double Bx = a.sh_c[3 * sj], By = a.sh_c[3 * sj + 1], Bz = a.sh_c[3 * sj + 2];
  const double dx = Ax - Bx, dy = Ay - By, dz = Az - Bz;
  const double ab2 = dx * dx + dy * dy + dz * dz;
  if (!s_nonzero) {
    // no candidates; only the K >= 1e-30 flag (the capacity) is recorded
    for (int s = tid; s < npi * npj; s += 64) {
      const double al = s_ei[s / npj], be = s_ej[s % npj];
      const double K = exp(-al * be / (al + be) * ab2);
      cand[(size_t)(slot0 + s) * S + 7] = K < 1e-30 ? 0.0 : 8.0;
    }
    if (tid == 0) a.sp_srule[sp] = 0.0;
    return;
  }
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
        rec[7] = 1.0 + (anyg ? 2.0 : 0.0) + (cc != 0.0 ? 4.0 : 0.0) + 8.0;
        srule += K * pow(M_PI / p, 1.5) * s_et[0];
      }
      for (int k = nterm + tid; k < S - 8; k += 64) rec[8 + k] = 0.0;
      __syncthreads();
    }
  }
  if (tid == 0) a.sp_srule[sp] = srule;
}

// Bounding box of each run of FE_BOX sorted points: box[6 b ..] = min xyz, max xyz
__global__ __launch_bounds__(64) void k_boxes(const double* __restrict__ pts, const int n,
                                              double* __restrict__ box) {
  const int b = blockIdx.x;
  const int i = min(b * FE_BOX + (int)threadIdx.x, n - 1);
  double lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pts[3 * (size_t)i + a];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, 64));
      hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, 64));
    }
  if (threadIdx.x < 3) {
    box[6 * (size_t)b + threadIdx.x] = lo[threadIdx.x];
    box[6 * (size_t)b + 3 + threadIdx.x] = hi[threadIdx.x];
  }
}

// The screen's test at squared distance d2 (a point's, or a bound): the
// reference's keep condition with that distance as dmin (rules 0, 1) or dmax
// (rule 2).  Non-increasing in d2.
__device__ __forceinline__ bool fe_pass(const double d2, const double kb, const double reff,
                                        const double tol, const int rule) {
  double dd = sqrt(fmax(d2, 0.0));
  if (rule != 0) dd -= reff;
  if (dd < 0.05) dd = 0.05;
  return !(kb / dd < tol);
}

struct ScreenArgs {
  const double* cand[FE_LMAX + 1];
  ClassInfo ci;
  const int* gidx;
  unsigned long long* gflag;  // per reference-order index: (nterm << 32 | 1) if kept, else 0
  uint32_t* key;
  int32_t* val;
  const double* pts;  // Morton-sorted
  int n_pts;
  const double* box;
  int n_box;
  double tol;
  int rule;
  double lo[3], kscale;  // Morton frame of P (8 bits per axis)
  double lpmin[FE_LMAX + 1], lpinv[FE_LMAX + 1];  // exponent buckets: 16 over the log range
  int* stats;  // FE_ST_* counters, zero before the launch
  int64_t n_all;
};

__global__ __launch_bounds__(256) void k_screen_boxes(const ScreenArgs a) {
  __shared__ double s_box[FE_BOX_CHUNK * 6];
  __shared__ int s_cnt[FE_ST_N];
  if (threadIdx.x < FE_ST_N) s_cnt[threadIdx.x] = 0;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool in = e < a.n_all;
  int L = 0;
  while (L < FE_LMAX && e >= a.ci.cand_off[L + 1]) ++L;
  const double* r = in ? a.cand[L] + (size_t)(e - a.ci.cand_off[L]) * a.ci.stride[L] : nullptr;
  const int flags = in ? (int)r[7] : 0;
  const bool valid = flags & 1;
  double Px = 0.0, Py = 0.0, Pz = 0.0, kb = 0.0, reff = 0.0;
  if (valid) {
    Px = r[0];
    Py = r[1];
    Pz = r[2];
    kb = r[5];
    reff = r[6];
  }
  // rules 0, 1: kept iff some point passes (the event is a pass); rule 2: kept
  // iff every point passes (the event is a fail).  Without an event the
  // decision is the opposite of the event's.
  const bool any_mode = a.rule != 2;
  const double tol = a.tol;
  const int rule = a.rule;
  bool done = !valid, keep = false;
  if (valid) {
    if (rule == 1 && !((flags & 2) && (flags & 4))) {
      done = true;
    } else if (a.n_pts == 0) {
      done = true;  // as if dmin = +inf (rules 0, 1) or dmax = 0 (rule 2)
      keep = fe_pass(any_mode ? INFINITY : 0.0, kb, reff, tol, rule);
    } else {
      keep = !any_mode;
    }
  }
  for (int b0 = 0; b0 < a.n_box; b0 += FE_BOX_CHUNK) {
    if (__syncthreads_and(done)) break;
    const int nb = min(FE_BOX_CHUNK, a.n_box - b0);
    for (int q = threadIdx.x; q < 6 * nb; q += 256) s_box[q] = a.box[6 * (size_t)b0 + q];
    __syncthreads();
    for (int j = 0; j < nb && !done; ++j) {
      const double* bx = s_box + 6 * j;
      // bounds of every point's squared distance, widened by 1e-12 relative
      // so that rounding cannot move a point across them
      const double gx = fmax(fmax(bx[0] - Px, Px - bx[3]), 0.0);
      const double gy = fmax(fmax(bx[1] - Py, Py - bx[4]), 0.0);
      const double gz = fmax(fmax(bx[2] - Pz, Pz - bx[5]), 0.0);
      const double hx = fmax(Px - bx[0], bx[3] - Px);
      const double hy = fmax(Py - bx[1], bx[4] - Py);
      const double hz = fmax(Pz - bx[2], bx[5] - Pz);
      const double lo2 = fma(gx, gx, fma(gy, gy, gz * gz)) * (1.0 - 1e-12);
      const double hi2 = fma(hx, hx, fma(hy, hy, hz * hz)) * (1.0 + 1e-12) + 1e-300;
      const bool pass_far = fe_pass(hi2, kb, reff, tol, rule);  // every point passes
      const bool pass_near = fe_pass(lo2, kb, reff, tol, rule); // some point may pass
      if (any_mode ? pass_far : !pass_near) {
        keep = any_mode;
        done = true;
        break;
      }
      if (any_mode ? !pass_near : pass_far) continue;
      const int q0 = (b0 + j) * FE_BOX, q1 = min(a.n_pts, q0 + FE_BOX);
      for (int q = q0; q < q1; ++q) {
        const double x = a.pts[3 * (size_t)q] - Px;
        const double y = a.pts[3 * (size_t)q + 1] - Py;
        const double z = a.pts[3 * (size_t)q + 2] - Pz;
        if (fe_pass(fma(x, x, fma(y, y, z * z)), kb, reff, tol, rule) == any_mode) {
          keep = any_mode;
          done = true;
          break;
        }
      }
    }
  }
  if (in) {
    const bool kept = valid && keep;
    uint32_t k = (uint32_t)L << 29;
    if (!kept) {
      k |= 1u << 28;
    } else {
      const double p = r[3];
      const int bucket = (int)fmin(fmax((log(p) - a.lpmin[L]) * a.lpinv[L], 0.0), 15.0);
      uint32_t code = 0;
      const double P[3] = {Px, Py, Pz};
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const uint32_t q = (uint32_t)fmin(fmax((P[ax] - a.lo[ax]) * a.kscale, 0.0), 255.0);
        for (int bit = 0; bit < 8; ++bit) code |= ((q >> bit) & 1u) << (3 * bit + ax);
      }
      k |= ((uint32_t)bucket << 24) | code;
    }
    a.key[e] = k;
    a.val[e] = (int32_t)e;
    a.gflag[a.gidx[e]] = kept ? (((unsigned long long)a.ci.nterm[L] << 32) | 1ull) : 0ull;
  }
  __syncthreads();
  if (in) {
    if (valid && keep) atomicAdd(&s_cnt[FE_ST_KEPT + L], 1);
    if (valid) atomicAdd(&s_cnt[FE_ST_VALID + L], 1);
    if (flags & 8) atomicAdd(&s_cnt[FE_ST_KVALID + L], 1);
  }
  __syncthreads();
  if (threadIdx.x < FE_ST_N && s_cnt[threadIdx.x])
    atomicAdd(&a.stats[threadIdx.x], s_cnt[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ a, const int64_t n, const int v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// sum rule: the shell-pair partials added in a fixed order (one block), and
// the counters copied into the table's statistics
__global__ __launch_bounds__(256) void k_build_stats(const double* __restrict__ srule, const int nsp,
                                                     const int* __restrict__ ws_stats,
                                                     int* __restrict__ stats,
                                                     double* __restrict__ sum_rule) {
  __shared__ double s_w[4];
  const int per = (nsp + 255) / 256;
  const int i0 = threadIdx.x * per, i1 = min(nsp, i0 + per);
  double s = 0.0;
  for (int i = i0; i < i1; ++i) s += srule[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *sum_rule = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  if (threadIdx.x < FE_ST_N && stats) stats[threadIdx.x] = ws_stats[threadIdx.x];
}

// Kept record k of class L: candidate sorted[k] (k < the kept count), with
// slots 6/7 holding the far-field constants 1/(2p) and K sqrt(pi/p)/2 instead
// of reff and the flags; column 0's thread also writes the flat term offset,
// the reference-order index and the slot -> record map.
__global__ __launch_bounds__(256) void k_compact(const double* __restrict__ cand, const int S,
                                                 const int64_t slot_base,
                                                 const int32_t* __restrict__ sorted,
                                                 const int* __restrict__ kept_p, const int cap,
                                                 const int* __restrict__ gidx,
                                                 const unsigned long long* __restrict__ gscan,
                                                 double* __restrict__ rec, int* __restrict__ flatoff,
                                                 int32_t* __restrict__ orig,
                                                 int* __restrict__ slot2kept) {
  const int n = min(*kept_p, cap);
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * S) return;
  const int r = (int)(e / S), c = (int)(e - (int64_t)r * S);
  const int ge = sorted[r];
  const int slot = (int)(ge - slot_base);
  const double* s = cand + (size_t)slot * S;
  rec[e] = c == 6 ? 0.5 / s[3] : c == 7 ? s[4] * (0.88622692545275801365 / sqrt(s[3])) : s[c];
  if (c == 0) {
    const unsigned long long g = gscan[gidx[ge]];
    flatoff[r] = (int)(g >> 32);
    orig[r] = (int32_t)(g & 0xffffffffull);
    slot2kept[slot] = r;
  }
}

// the scaled coefficients c g_k of each kept record of class L (one lane per record)
__global__ __launch_bounds__(256) void k_fill_dcoef(double* __restrict__ rec, const int S,
                                                    const int* __restrict__ kept_p, const int cap,
                                                    const int L) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < min(*kept_p, cap)) fe_fill_dcoef_any(L, rec + (size_t)i * S + 8);
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

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

int check_build_args(int n_points, const double* pts, const double* dens, double tol, int rule,
                     int ptr_kind, const char* fn) {
  if (n_points < 0) return fe_set_error(FE_ERR_ARG, "%s: negative point count", fn);
  if (!dens || (n_points && !pts)) return fe_set_error(FE_ERR_ARG, "%s: null array", fn);
  if (rule < 0 || rule > 2) return fe_set_error(FE_ERR_ARG, "%s: rule must be 0, 1 or 2", fn);
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "%s: bad ptr_kind %d", fn, ptr_kind);
  if (!(tol >= 0.0)) return fe_set_error(FE_ERR_ARG, "%s: tol must be >= 0", fn);
  return FE_OK;
}

// The device part of a build, from the density and the points to the table's
// statistics and records.  fe_table_build (`sized` false): the counts are
// read back once to allocate the table, then the records are compacted into
// it.  fe_table_rebuild (`sized` true): the table already holds a record for
// every K-valid candidate, and nothing is synchronised or allocated beyond the
// context's build workspace (which only grows for a larger point count).
int build_on_device(fe_table* t, const double* dens, int n_points, const double* pts, double tol,
                    int rule, int ptr_kind, hipStream_t st, bool sized) {
  fe_ctx* c = t->ctx;
  const int nsp = t->n_sp;
  ClassInfo ci{};
  for (int L = 0; L <= FE_LMAX; ++L) {
    ci.stride[L] = fe_host::STRIDE[L];
    ci.nterm[L] = fe_host::NTERM[L];
    ci.nmono[L] = fe_host::NMONO[L];
    ci.cand_off[L + 1] = ci.cand_off[L] + t->n_cand_L[L];
  }
  const int64_t n_all = ci.cand_off[FE_LMAX + 1];
  const bool host_in = ptr_kind == FE_PTR_HOST;
  const int n_box = (n_points + FE_BOX - 1) / FE_BOX;
  // ---- workspace layout
  size_t wo = 0;
  const size_t o_dens = wo;
  if (host_in) wo += al256((size_t)t->n_func * t->n_func * sizeof(double));
  const size_t o_pts = wo;
  if (host_in) wo += al256((size_t)n_points * 3 * sizeof(double));
  size_t o_cand[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_cand[L] = wo;
    wo += al256((size_t)t->n_cand_L[L] * fe_host::STRIDE[L] * sizeof(double));
  }
  const size_t o_gidx = wo;
  wo += al256((size_t)std::max<int64_t>(n_all, 1) * sizeof(int));
  const size_t o_key = wo;
  wo += al256((size_t)std::max<int64_t>(n_all, 1) * sizeof(uint32_t));
  const size_t o_val = wo;
  wo += al256((size_t)std::max<int64_t>(n_all, 1) * sizeof(int32_t));
  const size_t o_gflag = wo;
  wo += al256((size_t)std::max<int64_t>(n_all, 1) * sizeof(unsigned long long));
  const size_t o_srule = wo;
  wo += al256((size_t)std::max(nsp, 1) * sizeof(double));
  const size_t o_stats = wo;
  wo += al256(FE_ST_SUMRULE_BYTE + sizeof(double));
  const size_t o_box = wo;
  wo += al256((size_t)std::max(n_box, 1) * 6 * sizeof(double));
  const size_t o_sort = wo;  // point sort, then the candidate sort and scan
  const size_t b_key = al256((size_t)n_points * 4), b_pts = al256((size_t)n_points * 24);
  const size_t sort_pts = 2 * b_key + b_pts + fe_radix_sort_scratch(n_points);
  const size_t sort_cand = std::max(fe_radix_sort_scratch(n_all), fe_scan_scratch(n_all));
  wo += al256(std::max(sort_pts, sort_cand));
  if (int rc = fe_ctx_reserve(c, FE_WS_BUILD, wo)) return rc;
  char* wb = (char*)c->ws[FE_WS_BUILD];
  int* ws_stats = (int*)(wb + o_stats);

  // ---- density and points
  const double* d_dens = dens;
  const double* d_pts = pts;
  if (host_in) {
    FE_HIPCHK(hipMemcpyAsync(wb + o_dens, dens, (size_t)t->n_func * t->n_func * sizeof(double),
                             hipMemcpyHostToDevice, st));
    d_dens = (const double*)(wb + o_dens);
    if (n_points) {
      FE_HIPCHK(hipMemcpyAsync(wb + o_pts, pts, (size_t)n_points * 3 * sizeof(double),
                               hipMemcpyHostToDevice, st));
      d_pts = (const double*)(wb + o_pts);
    }
  }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, st, ws_stats, (int64_t)FE_ST_N, 0);

  // ---- candidates
  BuildArgs ba{};
  ba.sh_atom = t->b.atom; ba.sh_l = t->b.l; ba.sh_fn0 = t->b.fn0; ba.sh_nfn = t->b.nfn;
  ba.sh_prim0 = t->b.prim0; ba.sh_nprim = t->b.nprim; ba.sh_c = t->b.c;
  ba.pexp = t->b.pexp; ba.pcoef = t->b.pcoef;
  ba.fn_lx = t->b.fn_lx; ba.fn_ly = t->b.fn_ly; ba.fn_lz = t->b.fn_lz;
  ba.sp_i = t->b.sp_i; ba.sp_j = t->b.sp_j; ba.sp_slot = t->b.sp_slot; ba.sp_gidx = t->b.sp_gidx;
  ba.dens = d_dens;
  ba.n_func = t->n_func;
  for (int L = 0; L <= FE_LMAX; ++L) ba.cand[L] = (double*)(wb + o_cand[L]);
  ba.gidx = (int*)(wb + o_gidx);
  ba.sp_srule = (double*)(wb + o_srule);
  ba.lt = c->ltab;
  ba.ci = ci;
  if (nsp > 0) hipLaunchKernelGGL(k_build_pairs, dim3(nsp), dim3(64), 0, st, ba);

  // ---- points: Morton order, boxes of FE_BOX
  const double* sorted = nullptr;
  if (n_points > 0) {
    const int32_t* perm = nullptr;
    if (int rc = fe_sort_points(c, t, d_pts, n_points, st, &sorted, &perm, wb + o_sort, sort_pts))
      return rc;
    hipLaunchKernelGGL(k_boxes, dim3(n_box), dim3(64), 0, st, sorted, n_points,
                       (double*)(wb + o_box));
  }

  // ---- screen, sort keys, reference-order flags
  uint32_t* key = (uint32_t*)(wb + o_key);
  int32_t* val = (int32_t*)(wb + o_val);
  unsigned long long* gflag = (unsigned long long*)(wb + o_gflag);
  ScreenArgs sa{};
  for (int L = 0; L <= FE_LMAX; ++L) {
    sa.cand[L] = ba.cand[L];
    sa.lpmin[L] = t->key_lpmin[L];
    sa.lpinv[L] = t->key_lpinv[L];
  }
  sa.ci = ci;
  sa.gidx = ba.gidx;
  sa.gflag = gflag;
  sa.key = key;
  sa.val = val;
  sa.pts = sorted;
  sa.n_pts = n_points;
  sa.box = (const double*)(wb + o_box);
  sa.n_box = n_box;
  sa.tol = tol;
  sa.rule = rule;
  for (int a = 0; a < 3; ++a) sa.lo[a] = t->key_lo[a];
  sa.kscale = t->key_scale;
  sa.stats = ws_stats;
  sa.n_all = n_all;
  if (n_all > 0) {
    hipLaunchKernelGGL(k_screen_boxes, dim3((unsigned)((n_all + 255) / 256)), dim3(256), 0, st,
                       sa);
    if (int rc = fe_radix_sort(key, val, (int)n_all, 32, wb + o_sort, st)) return rc;
    if (int rc = fe_scan_excl_u64(gflag, gflag, (int)n_all, wb + o_sort, st)) return rc;
  }
  FE_HIPCHK(hipGetLastError());

  if (!sized) {
    // the one synchronisation of a build: the counts size the table
    hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(256), 0, st, ba.sp_srule, nsp, ws_stats,
                       (int*)nullptr, (double*)(wb + o_stats + FE_ST_SUMRULE_BYTE));
    int hs[FE_ST_N];
    double sr = 0.0;
    FE_HIPCHK(hipMemcpyAsync(hs, ws_stats, sizeof hs, hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipMemcpyAsync(&sr, wb + o_stats + FE_ST_SUMRULE_BYTE, sizeof sr,
                             hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
    int64_t cap[FE_LMAX + 1], np = 0, nt = 0, nc = 0;
    for (int L = 0; L <= FE_LMAX; ++L) {
      cap[L] = t->n_L[L] = hs[FE_ST_KEPT + L];
      t->kvalid_L[L] = hs[FE_ST_KVALID + L];
      np += hs[FE_ST_KEPT + L];
      nt += (int64_t)hs[FE_ST_KEPT + L] * fe_host::NTERM[L];
      nc += hs[FE_ST_VALID + L];
    }
    if (nt > INT32_MAX) return fe_set_error(FE_ERR_UNSUPPORTED, "table: %lld terms", (long long)nt);
    t->n_pairs = (int)np;
    t->n_terms = (int)nt;
    t->n_candidates = nc;
    t->sum_rule = sr;
    if (int rc = fe_table_layout(t, cap, n_all)) return rc;
  }
  hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(256), 0, st, ba.sp_srule, nsp, ws_stats,
                     t->d_stats, t->d_sum_rule);

  // ---- compaction into the table, in key order
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!t->n_cand_L[L]) continue;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((t->n_cand_L[L] + 255) / 256)), dim3(256), 0,
                       st, t->d_slot2kept[L], t->n_cand_L[L], -1);
    if (!t->cap_L[L]) continue;
    const int S = fe_host::STRIDE[L];
    const int* kept = t->d_stats + FE_ST_KEPT + L;
    const int64_t ne = t->cap_L[L] * S;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st,
                       ba.cand[L], S, ci.cand_off[L], val + ci.cand_off[L], kept,
                       (int)t->cap_L[L], ba.gidx, gflag, t->d_rec[L], t->d_flatoff[L],
                       t->d_orig[L], t->d_slot2kept[L]);
    if (fe_host::NDCOEF[L])
      hipLaunchKernelGGL(k_fill_dcoef, dim3((unsigned)((t->cap_L[L] + 255) / 256)), dim3(256), 0,
                         st, t->d_rec[L], S, kept, (int)t->cap_L[L], L);
  }
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
}  // namespace

extern "C" int fe_table_build(fe_ctx* c, int n_func, int n_prim, const int32_t* fn_atom,
                              const int32_t* fn_nprim, const int32_t* fn_off,
                              const int32_t* fn_lx, const int32_t* fn_ly, const int32_t* fn_lz,
                              const int32_t* fn_shell, const double* fc, const double* pexp,
                              const double* pcoef, const double* dens, int n_points,
                              const double* pts, int n_atoms, const double* Z, const double* xyz,
                              double tol, int rule, int ptr_kind, void* stream, fe_table** out) {
  // arguments first (no device needed to reject them)
  if (out) *out = nullptr;
  if (n_func <= 0 || n_prim <= 0 || n_atoms < 0)
    return fe_set_error(FE_ERR_ARG, "fe_table_build: bad sizes");
  if (!fn_atom || !fn_nprim || !fn_off || !fn_lx || !fn_ly || !fn_lz || !fn_shell || !fc ||
      !pexp || !pcoef || (n_atoms && (!Z || !xyz)))
    return fe_set_error(FE_ERR_ARG, "fe_table_build: null array");
  if (int rc = check_build_args(n_points, pts, dens, tol, rule, ptr_kind, "fe_table_build"))
    return rc;
  if (!c || !out) return fe_set_error(FE_ERR_ARG, "fe_table_build: null argument");

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
      if (!(pexp[fn_off[q]] > 0.0))
        return fe_set_error(FE_ERR_FORMAT, "function %d: exponents must be > 0", q);
    }
    sh.push_back(s);
    i = j;
  }
  const int nsh = (int)sh.size();
  // ---- shell pairs, candidate slots per class, reference-order indices, and
  // the log range of the product exponents of each class (pair-order keys)
  std::vector<double> emin(nsh), emax(nsh);
  for (int i = 0; i < nsh; ++i) {
    emin[i] = *std::min_element(pexp + sh[i].prim0, pexp + sh[i].prim0 + sh[i].nprim);
    emax[i] = *std::max_element(pexp + sh[i].prim0, pexp + sh[i].prim0 + sh[i].nprim);
  }
  std::vector<int> sp_i, sp_j, sp_slot, sp_gidx;
  int64_t ncand[FE_LMAX + 1] = {}, gtot = 0;
  double pmin[FE_LMAX + 1], pmax[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    pmin[L] = INFINITY;
    pmax[L] = 0.0;
  }
  for (int i = 0; i < nsh; ++i)
    for (int j = i; j < nsh; ++j) {
      const int L = sh[i].l + sh[j].l;
      const int64_t np = (int64_t)sh[i].nprim * sh[j].nprim;
      sp_i.push_back(i);
      sp_j.push_back(j);
      sp_slot.push_back((int)ncand[L]);
      sp_gidx.push_back((int)gtot);
      ncand[L] += np;
      gtot += np;
      pmin[L] = std::min(pmin[L], emin[i] + emin[j]);
      pmax[L] = std::max(pmax[L], emax[i] + emax[j]);
      if (gtot > INT32_MAX / 2)
        return fe_set_error(FE_ERR_UNSUPPORTED, "too many primitive pairs (%lld)", (long long)gtot);
    }
  const int nsp = (int)sp_i.size();

  FE_CTX_LOCK(c);
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
  for (int L = 0; L <= FE_LMAX; ++L) {
    t->n_cand_L[L] = ncand[L];
    const double lo = std::log(std::isfinite(pmin[L]) ? pmin[L] : 1.0);
    const double w = std::log(std::max(pmax[L], 1e-300)) - lo;
    t->key_lpmin[L] = lo;
    t->key_lpinv[L] = 16.0 / std::max(w, 1e-12);
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const HostShell& s : sh)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], s.c[a]);
      hi[a] = std::max(hi[a], s.c[a]);
    }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6});
  for (int a = 0; a < 3; ++a) t->key_lo[a] = lo[a];
  t->key_scale = 255.0 / ext;
  auto fail = [&](int rc) {
    (void)hipStreamSynchronize(st);
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
               o_sps = push_bytes(hb, sp_slot.data(), nsp),
               o_spg = push_bytes(hb, sp_gidx.data(), nsp);
  if (hipMalloc(&t->d_basis_mem, hb.size()) != hipSuccess) {
    (void)hipGetLastError();
    return fail(fe_set_error(FE_ERR_ALLOC, "fe_table_build: basis allocation failed"));
  }
  // synchronous: hb is a local buffer
  if (hipMemcpy(t->d_basis_mem, hb.data(), hb.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return fail(fe_set_error(FE_ERR_HIP, "fe_table_build: basis upload (hipMemcpy) failed"));
  }
  t->dev_bytes += hb.size();
  const char* bm = (const char*)t->d_basis_mem;
  t->b.atom = (const int*)(bm + o_atom); t->b.l = (const int*)(bm + o_l);
  t->b.fn0 = (const int*)(bm + o_fn0); t->b.nfn = (const int*)(bm + o_nfn);
  t->b.prim0 = (const int*)(bm + o_p0); t->b.nprim = (const int*)(bm + o_np);
  t->b.c = (const double*)(bm + o_c); t->b.pexp = (const double*)(bm + o_pexp);
  t->b.pcoef = (const double*)(bm + o_pcoef); t->b.fn_lx = (const int*)(bm + o_lx);
  t->b.fn_ly = (const int*)(bm + o_ly); t->b.fn_lz = (const int*)(bm + o_lz);
  t->b.sp_i = (const int*)(bm + o_spi); t->b.sp_j = (const int*)(bm + o_spj);
  t->b.sp_slot = (const int*)(bm + o_sps); t->b.sp_gidx = (const int*)(bm + o_spg);
  // nuclei (the Morton frame of the point sort)
  if (int rc = fe_table_set_atoms(t, n_atoms, Z, xyz)) return fail(rc);
  if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, ptr_kind, st, false))
    return fail(rc);
  *out = t;
  return FE_OK;
}

extern "C" int fe_table_rebuild(fe_table* t, const double* dens, int n_points, const double* pts,
                                double tol, int rule, int ptr_kind, void* stream) {
  if (!t) return fe_set_error(FE_ERR_ARG, "fe_table_rebuild: null table");
  if (int rc = check_build_args(n_points, pts, dens, tol, rule, ptr_kind, "fe_table_rebuild"))
    return rc;
  if (!t->has_basis)
    return fe_set_error(FE_ERR_STATE,
                        "fe_table_rebuild: the table was not built from a basis (fe_table_build)");
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  if (!t->capacity_layout) {
    // first rebuild: room for every K-valid candidate (a bound that holds for
    // any density and any points, so later rebuilds never allocate)
    if (int rc = fe_table_sync_counts(t)) return rc;
    FE_HIPCHK(hipDeviceSynchronize());
    if (int rc = fe_table_layout(t, t->kvalid_L, t->n_cand_L[0] + t->n_cand_L[1] +
                                 t->n_cand_L[2] + t->n_cand_L[3] + t->n_cand_L[4] +
                                 t->n_cand_L[5] + t->n_cand_L[6]))
      return rc;
    bool had32 = false;
    for (int L = 0; L <= FE_LMAX; ++L) {
      if (t->d_rec32[L]) {
        had32 = true;
        (void)hipFree(t->d_rec32[L]);
        t->d_rec32[L] = nullptr;
      }
    }
    t->capacity_layout = true;
    t->counts_stale = true;  // plans use the capacity from now on
    t->count_stream = st;
    if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, ptr_kind, st, true)) return rc;
    if (had32)
      if (int rc = fe_ensure_g32(t, st)) return rc;
  } else {
    if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, ptr_kind, st, true)) return rc;
    if (int rc = fe_refresh_g32(t, st)) return rc;
  }
  t->counts_stale = true;
  t->count_stream = st;
  if (ptr_kind == FE_PTR_HOST) FE_HIPCHK(hipStreamSynchronize(st));
  return FE_OK;
}

extern "C" int fe_table_stats(const fe_table* t, double* out, int n) {
  if (!t || !out) return fe_set_error(FE_ERR_ARG, "fe_table_stats: null argument");
  if (int rc = fe_table_sync_counts(const_cast<fe_table*>(t))) return rc;
  double v[4] = {t->sum_rule, (double)t->n_pairs, (double)t->n_func, (double)t->n_sp};
  for (int i = 0; i < n && i < 4; ++i) out[i] = v[i];
  return FE_OK;
}
