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

// The largest non-negative double T with fe_pass(T) (-1 if none passes), so
// that fe_pass(d2) == (d2 <= T) for every d2 >= 0: fe_pass is monotone, and
// the bit patterns of non-negative doubles are ordered like their values, so
// a bisection over the patterns finds the switch exactly (63 steps).
__device__ double fe_pass_threshold(const double kb, const double reff, const double tol,
                                    const int rule) {
  if (!fe_pass(0.0, kb, reff, tol, rule)) return -1.0;
  unsigned long long lo = 0ull, hi = 0x7FF0000000000000ull;  // pass(lo), +inf
  if (fe_pass(INFINITY, kb, reff, tol, rule)) return INFINITY;
  while (hi - lo > 1ull) {
    const unsigned long long mid = lo + (hi - lo) / 2;
    if (fe_pass(__longlong_as_double((long long)mid), kb, reff, tol, rule)) lo = mid;
    else hi = mid;
  }
  return __longlong_as_double((long long)lo);
}

// lane l's value of a double (l wave-uniform)
__device__ __forceinline__ double fe_readlane(const double v, const int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

struct ScreenArgs {
  const double* cand[FE_LMAX + 1];
  ClassInfo ci;
  const int* gidx;
  unsigned long long* gflag;  // per reference-order index: (nterm << 32 | 1) if kept, else 0
  uint32_t* key;              // final pair-order keys and candidate indices
  int32_t* val;
  uint32_t* skey;             // screening order: keys and candidate indices
  int32_t* sval;
  double* thr;                // per candidate: the squared-distance threshold
  const double* pts;  // Morton-sorted
  int n_pts;
  const double* box;
  int n_box;
  double tol;
  int rule;
  double lo[3], kscale;  // Morton frame of P (8 bits per axis)
  double lpmin[FE_LMAX + 1], lpinv[FE_LMAX + 1];  // exponent buckets: 16 over the log range
  int* stats;  // FE_ST_* counters, zero before the launch
  int* n_open;  // candidates left to the box walk (counted by k_screen_prep)
  int64_t n_all;
};

__device__ __forceinline__ int fe_class_of(const ClassInfo& ci, const int64_t e) {
  int L = 0;
  while (L < FE_LMAX && e >= ci.cand_off[L + 1]) ++L;
  return L;
}

__device__ __forceinline__ uint32_t fe_morton24(const double* P, const double* lo,
                                                const double scale) {
  uint32_t code = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const uint32_t q = (uint32_t)fmin(fmax((P[ax] - lo[ax]) * scale, 0.0), 255.0);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) code |= ((q >> bit) & 1u) << (3 * bit + ax);
  }
  return code;
}

// The screen's outputs for candidate e of class L: its pair-order key (class,
// dropped flag, exponent bucket, Morton code of P) and its reference-order
// flag (term count and 1 when kept, else 0).
__device__ __forceinline__ void fe_screen_emit(const ScreenArgs& a, const int64_t e, const int L,
                                               const bool kept, const double* r) {
  uint32_t k = (uint32_t)L << 29;
  if (!kept) {
    k |= 1u << 28;
  } else {
    const int bucket = (int)fmin(fmax((log(r[3]) - a.lpmin[L]) * a.lpinv[L], 0.0), 15.0);
    k |= ((uint32_t)bucket << 24) | fe_morton24(r, a.lo, a.kscale);
  }
  a.key[e] = k;
  a.val[e] = (int32_t)e;
  a.gflag[a.gidx[e]] = kept ? (((unsigned long long)a.ci.nterm[L] << 32) | 1ull) : 0ull;
}

// per-block counters of kept candidates, added to the statistics at the end
__device__ __forceinline__ void fe_flush_counts(int* s_cnt, int* stats) {
  __syncthreads();
  if (threadIdx.x < FE_ST_N && s_cnt[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_cnt[threadIdx.x]);
}

// Every candidate slot: the counters of candidates and K-valid pairs, the
// squared-distance threshold of the test, and either the final decision (no
// candidate, rule 1's G/coefficient test, no points, or a threshold that
// settles every distance) or a screening key that groups candidates with
// alike thresholds and nearby P, so the box walk of a wavefront is coherent.
__global__ __launch_bounds__(256) void k_screen_prep(const ScreenArgs a) {
  __shared__ int s_cnt[FE_ST_N];
  if (threadIdx.x < FE_ST_N) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < a.n_all) {
    const int L = fe_class_of(a.ci, e);
    const double* r = a.cand[L] + (size_t)(e - a.ci.cand_off[L]) * a.ci.stride[L];
    const int flags = (int)r[7];
    const bool valid = flags & 1;
    if (flags & 8) atomicAdd(&s_cnt[FE_ST_KVALID + L], 1);
    if (valid) atomicAdd(&s_cnt[FE_ST_VALID + L], 1);
    const bool any_mode = a.rule != 2;
    int decided = valid ? -1 : 0;  // -1 open, 0 dropped, 1 kept
    double thr = -1.0;
    if (valid) {
      const double kb = r[5], reff = r[6];
      if (a.rule == 1 && !((flags & 2) && (flags & 4))) {
        decided = 0;
      } else if (a.n_pts == 0) {
        // as if dmin = +inf (rules 0, 1) or dmax = 0 (rule 2)
        decided = fe_pass(any_mode ? INFINITY : 0.0, kb, reff, a.tol, a.rule) ? 1 : 0;
      } else {
        thr = fe_pass_threshold(kb, reff, a.tol, a.rule);
        if (thr < 0.0) decided = 0;               // no distance passes
        else if (thr == INFINITY) decided = 1;    // every distance passes
      }
    }
    a.thr[e] = thr;
    if (decided >= 0) {
      fe_screen_emit(a, e, L, decided == 1, r);
      if (decided == 1) atomicAdd(&s_cnt[FE_ST_KEPT + L], 1);
      a.skey[e] = 0xFFFFFFFFu;  // sorts after every open candidate
    } else {
      atomicAdd(a.n_open, 1);
      // 254 at most: the key 0xFFFFFFFF marks a decided candidate
      const int tb = (int)fmin(fmax(floor(log2(thr)) + 128.0, 0.0), 254.0);
      a.skey[e] = ((uint32_t)tb << 24) | fe_morton24(r, a.lo, a.kscale);
    }
    a.sval[e] = (int32_t)e;
  }
  fe_flush_counts(s_cnt, a.stats);
}

// The open candidates, one wavefront each (persistent: wavefront w takes the
// sorted positions w, w + W, ... below the open count).  The rules 0, 1 keep
// a candidate iff some point passes the test (the event is a pass); rule 2 iff
// every point passes (the event is a fail); without an event the decision is
// the opposite.  The lanes take 64 boxes at a time: a box whose distance
// bounds settle the test for all its points decides at once; each lane whose
// box is ambiguous tests that box's points itself.  So no candidate's work
// runs on a single lane, and the slowest candidate costs a few box steps.
__global__ __launch_bounds__(256) void k_screen_boxes(const ScreenArgs a) {
  __shared__ int s_cnt[FE_ST_N];
  if (threadIdx.x < FE_ST_N) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int n_open = *a.n_open;
  const int W = (int)gridDim.x * 4;
  const bool any_mode = a.rule != 2;
  for (int i = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6); i < n_open; i += W) {
    const int64_t e = a.sval[i];
    const int L = fe_class_of(a.ci, e);
    const double* r = a.cand[L] + (size_t)(e - a.ci.cand_off[L]) * a.ci.stride[L];
    const double Px = r[0], Py = r[1], Pz = r[2];
    const double thr = a.thr[e];  // fe_pass(d2) == (d2 <= thr)
    bool keep = !any_mode;
    for (int b0 = 0; b0 < a.n_box; b0 += 64) {
      const int b = b0 + lane;
      const bool vb = b < a.n_box;
      const double* bx = a.box + 6 * (size_t)(vb ? b : 0);
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
      const bool pass_far = hi2 <= thr;   // every point of the box passes
      const bool pass_near = lo2 <= thr;  // some point of the box may pass
      // rules 0/1: a box whose every point passes keeps; rule 2: a box whose
      // every point fails drops
      if (__any(vb && (any_mode ? pass_far : !pass_near))) {
        keep = any_mode;
        break;
      }
      const bool amb = vb && pass_near && !pass_far;
      if (!__any(amb)) continue;
      bool hit = false;  // a point of this lane's box passes (rules 0/1) or fails (rule 2)
      if (amb) {
        const int q0 = b * FE_BOX, q1 = min(a.n_pts, q0 + FE_BOX);
#pragma unroll 8
        for (int q = q0; q < q1; ++q) {
          const double x = a.pts[3 * (size_t)q] - Px;
          const double y = a.pts[3 * (size_t)q + 1] - Py;
          const double z = a.pts[3 * (size_t)q + 2] - Pz;
          hit |= (fma(x, x, fma(y, y, z * z)) <= thr) == any_mode;
        }
      }
      if (__any(hit)) {
        keep = any_mode;
        break;
      }
    }
    if (lane == 0) {
      fe_screen_emit(a, e, L, keep, r);
      if (keep) atomicAdd(&s_cnt[FE_ST_KEPT + L], 1);
    }
  }
  fe_flush_counts(s_cnt, a.stats);
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ a, const int64_t n, const int v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// Sum rule in two fixed-order steps: block b adds the shell-pair partials
// [b * per, (b + 1) * per) (k_srule_part), then one block adds the block sums
// and copies the counters into the table's statistics (k_build_stats).
#define FE_SRULE_BLOCKS 256
__global__ __launch_bounds__(256) void k_srule_part(const double* __restrict__ srule, const int nsp,
                                                    double* __restrict__ part) {
  __shared__ double s_w[4];
  const int per = (nsp + FE_SRULE_BLOCKS - 1) / FE_SRULE_BLOCKS;
  const int i0 = blockIdx.x * per, i1 = min(nsp, i0 + per);
  double s = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) s += srule[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

__global__ __launch_bounds__(256) void k_build_stats(const double* __restrict__ part,
                                                     const int* __restrict__ ws_stats,
                                                     int* __restrict__ stats,
                                                     double* __restrict__ sum_rule) {
  __shared__ double s_w[4];
  double s = threadIdx.x < FE_SRULE_BLOCKS ? part[threadIdx.x] : 0.0;
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
__global__ __launch_bounds__(256) void k_compact(const double* __restrict__ cand, const int CS,
                                                 const int S, const int64_t slot_base,
                                                 const int32_t* __restrict__ sorted,
                                                 const int* __restrict__ kept_p, const int cap,
                                                 const int* __restrict__ gidx,
                                                 const unsigned long long* __restrict__ gscan,
                                                 double* __restrict__ rec, int* __restrict__ flatoff,
                                                 int32_t* __restrict__ orig,
                                                 int* __restrict__ slot2kept) {
  const int n = min(*kept_p, cap);
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * CS) return;
  const int r = (int)(e / CS), c = (int)(e - (int64_t)r * CS);
  const int ge = sorted[r];
  const int slot = (int)(ge - slot_base);
  const double* s = cand + (size_t)slot * CS;
  rec[(size_t)r * S + c] =
      c == 6 ? 0.5 / s[3] : c == 7 ? s[4] * (0.88622692545275801365 / sqrt(s[3])) : s[c];
  if (c == 0) {
    const unsigned long long g = gscan[gidx[ge]];
    flatoff[r] = (int)(g >> 32);
    orig[r] = (int32_t)(g & 0xffffffffull);
    slot2kept[slot] = r;
  }
}

// the derived parts of each kept record of class L (one lane per record): the
// scaled coefficients c g_k and the far-field block
__global__ __launch_bounds__(256) void k_fill_record(double* __restrict__ rec, const int S,
                                                     const int* __restrict__ kept_p, const int cap,
                                                     const int L) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < min(*kept_p, cap)) fe_fill_record_any(L, rec + (size_t)i * S);
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
// Byte layout of a build's workspace (candidates, keys, flags, thresholds,
// boxes, sort and scan scratch); host inputs also stage the density and points.
struct BuildLayout {
  size_t o_dens = 0, o_pts = 0, o_cand[FE_LMAX + 1] = {}, o_gidx = 0, o_key = 0, o_val = 0,
         o_gflag = 0, o_skey = 0, o_sval = 0, o_thr = 0, o_srule = 0, o_stats = 0, o_spart = 0,
         o_box = 0, o_psort = 0, sort_pts = 0, o_sort = 0, total = 0;
};

static BuildLayout build_layout(const fe_table* t, int n_points, bool host_in) {
  BuildLayout y;
  int64_t n_all = 0;
  for (int L = 0; L <= FE_LMAX; ++L) n_all += t->n_cand_L[L];
  const int n_box = (n_points + FE_BOX - 1) / FE_BOX;
  size_t wo = 0;
  y.o_dens = wo;
  if (host_in) wo += al256((size_t)t->n_func * t->n_func * sizeof(double));
  y.o_pts = wo;
  if (host_in) wo += al256((size_t)n_points * 3 * sizeof(double));
  for (int L = 0; L <= FE_LMAX; ++L) {
    y.o_cand[L] = wo;
    wo += al256((size_t)t->n_cand_L[L] * fe_host::CSTRIDE[L] * sizeof(double));
  }
  const size_t na = (size_t)std::max<int64_t>(n_all, 1);
  y.o_gidx = wo;
  wo += al256(na * sizeof(int));
  y.o_key = wo;
  wo += al256(na * sizeof(uint32_t));
  y.o_val = wo;
  wo += al256(na * sizeof(int32_t));
  y.o_gflag = wo;
  wo += al256(na * sizeof(unsigned long long));
  y.o_skey = wo;
  wo += al256(na * sizeof(uint32_t));
  y.o_sval = wo;
  wo += al256(na * sizeof(int32_t));
  y.o_thr = wo;
  wo += al256(na * sizeof(double));
  y.o_srule = wo;
  wo += al256((size_t)std::max(t->n_sp, 1) * sizeof(double));
  y.o_stats = wo;  // FE_ST_N counters, the open count, the sum rule
  wo += al256(FE_ST_SUMRULE_BYTE + sizeof(double));
  y.o_spart = wo;
  wo += al256(FE_SRULE_BLOCKS * sizeof(double));
  y.o_box = wo;
  wo += al256((size_t)std::max(n_box, 1) * 6 * sizeof(double));
  // the point sort (its sorted copy is read by the screen), then the
  // candidate sorts and the scan
  y.o_psort = wo;
  const size_t b_key = al256((size_t)n_points * 4), b_pts = al256((size_t)n_points * 24);
  y.sort_pts = 2 * b_key + b_pts + fe_radix_sort_scratch(n_points);
  wo += al256(y.sort_pts);
  y.o_sort = wo;
  wo += al256(std::max(fe_radix_sort_scratch(n_all), fe_scan_scratch(n_all)));
  y.total = wo;
  return y;
}

// The build proper.  `ws` / `ws_bytes`: a caller-owned workspace of at least
// fe_build_ws_bytes bytes (device inputs only), else the context's.
int build_on_device(fe_table* t, const double* dens, int n_points, const double* pts, double tol,
                    int rule, int ptr_kind, hipStream_t st, bool sized, char* ws = nullptr,
                    size_t ws_bytes = 0) {
  fe_ctx* c = t->ctx;
  const int nsp = t->n_sp;
  ClassInfo ci{};
  for (int L = 0; L <= FE_LMAX; ++L) {
    ci.stride[L] = fe_host::CSTRIDE[L];
    ci.nterm[L] = fe_host::NTERM[L];
    ci.nmono[L] = fe_host::NMONO[L];
    ci.cand_off[L + 1] = ci.cand_off[L] + t->n_cand_L[L];
  }
  const int64_t n_all = ci.cand_off[FE_LMAX + 1];
  const bool host_in = ptr_kind == FE_PTR_HOST;
  const int n_box = (n_points + FE_BOX - 1) / FE_BOX;
  const BuildLayout y = build_layout(t, n_points, host_in);
  const size_t o_dens = y.o_dens, o_pts = y.o_pts, o_gidx = y.o_gidx, o_key = y.o_key,
               o_val = y.o_val, o_gflag = y.o_gflag, o_skey = y.o_skey, o_sval = y.o_sval,
               o_thr = y.o_thr, o_srule = y.o_srule, o_stats = y.o_stats, o_spart = y.o_spart,
               o_box = y.o_box, o_psort = y.o_psort, sort_pts = y.sort_pts, o_sort = y.o_sort;
  const size_t* o_cand = y.o_cand;
  char* wb = ws;
  if (wb) {
    if (host_in || ws_bytes < y.total)
      return fe_set_error(FE_ERR_STATE, "build: caller workspace of %zu bytes, %zu needed",
                          ws_bytes, y.total);
  } else {
    if (int rc = fe_ctx_reserve(c, FE_WS_BUILD, y.total)) return rc;
    wb = (char*)c->ws[FE_WS_BUILD];
  }
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
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, st, ws_stats, (int64_t)FE_ST_N + 1, 0);

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
    if (int rc = fe_sort_points(c, t, d_pts, n_points, st, &sorted, &perm, wb + o_psort, sort_pts))
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
  sa.skey = (uint32_t*)(wb + o_skey);
  sa.sval = (int32_t*)(wb + o_sval);
  sa.thr = (double*)(wb + o_thr);
  sa.pts = sorted;
  sa.n_pts = n_points;
  sa.box = (const double*)(wb + o_box);
  sa.n_box = n_box;
  sa.tol = tol;
  sa.rule = rule;
  for (int a = 0; a < 3; ++a) sa.lo[a] = t->key_lo[a];
  sa.kscale = t->key_scale;
  sa.stats = ws_stats;
  sa.n_open = ws_stats + FE_ST_N;
  sa.n_all = n_all;
  if (n_all > 0) {
    const dim3 g((unsigned)((n_all + 255) / 256));
    hipLaunchKernelGGL(k_screen_prep, g, dim3(256), 0, st, sa);
    if (int rc = fe_radix_sort(sa.skey, sa.sval, (int)n_all, 32, wb + o_sort, st)) return rc;
    // persistent: 16 workgroups (64 wavefronts) per CU walk the open candidates
    hipLaunchKernelGGL(k_screen_boxes, dim3((unsigned)std::max(1, c->n_cu * 16)), dim3(256), 0,
                       st, sa);
    if (int rc = fe_radix_sort(key, val, (int)n_all, 32, wb + o_sort, st)) return rc;
    if (int rc = fe_scan_excl_u64(gflag, gflag, (int)n_all, wb + o_sort, st)) return rc;
  }
  FE_HIPCHK(hipGetLastError());

  double* spart = (double*)(wb + o_spart);
  hipLaunchKernelGGL(k_srule_part, dim3(FE_SRULE_BLOCKS), dim3(256), 0, st, ba.sp_srule, nsp,
                     spart);
  if (!sized) {
    // the one synchronisation of a build: the counts size the table
    hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(256), 0, st, spart, ws_stats, (int*)nullptr,
                       (double*)(wb + o_stats + FE_ST_SUMRULE_BYTE));
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
  hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(256), 0, st, spart, ws_stats, t->d_stats,
                     t->d_sum_rule);

  // ---- compaction into the table, in key order
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!t->n_cand_L[L]) continue;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((t->n_cand_L[L] + 255) / 256)), dim3(256), 0,
                       st, t->d_slot2kept[L], t->n_cand_L[L], -1);
    if (!t->cap_L[L]) continue;
    const int S = fe_host::STRIDE[L], CS = fe_host::CSTRIDE[L];
    const int* kept = t->d_stats + FE_ST_KEPT + L;
    const int64_t ne = t->cap_L[L] * CS;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st,
                       ba.cand[L], CS, S, ci.cand_off[L], val + ci.cand_off[L], kept,
                       (int)t->cap_L[L], ba.gidx, gflag, t->d_rec[L], t->d_flatoff[L],
                       t->d_orig[L], t->d_slot2kept[L]);
    hipLaunchKernelGGL(k_fill_record, dim3((unsigned)((t->cap_L[L] + 255) / 256)), dim3(256), 0,
                       st, t->d_rec[L], S, kept, (int)t->cap_L[L], L);
  }
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
}  // namespace

size_t fe_build_ws_bytes(const fe_table* t, int n_points) {
  return build_layout(t, n_points, false).total;
}

namespace {
// A basis set up for the device: the shells, the shell pairs (si <= sj in
// reference order) with their candidate slots and reference-order indices,
// the pair-order key frame, and one device block holding all of it.  The
// context keeps the last one, so builds that pass the same basis arrays again
// (new densities or points) skip the host set-up and the upload.
struct BasisSetup {
  int device = 0;
  std::vector<int32_t> ikey;  // the input arrays: the cache key
  std::vector<double> dkey;
  int n_func = 0, n_prim = 0, nsh = 0, nsp = 0;
  int64_t ncand[FE_LMAX + 1] = {};
  double lpmin[FE_LMAX + 1] = {}, lpinv[FE_LMAX + 1] = {};
  double lo[3] = {0, 0, 0}, scale = 0.0;
  void* mem = nullptr;
  size_t bytes = 0;
  size_t o_atom = 0, o_l = 0, o_fn0 = 0, o_nfn = 0, o_p0 = 0, o_np = 0, o_c = 0, o_pexp = 0,
         o_pcoef = 0, o_lx = 0, o_ly = 0, o_lz = 0, o_spi = 0, o_spj = 0, o_sps = 0, o_spg = 0,
         o_spc = 0, o_task = 0;
  int sp_cls_off[FE_LMAX + 2] = {};  // shell pairs of class L: sp_by_class[off[L], off[L+1])
  int task_cls_off[FE_LMAX + 2] = {};  // dL/dP tasks of class L: tasks[off[L], off[L+1])
  ~BasisSetup() {
    if (mem) {
      (void)hipSetDevice(device);
      (void)hipFree(mem);
    }
  }
};

void basis_key(int n_func, int n_prim, const int32_t* const* iv, const double* fc,
               const double* pexp, const double* pcoef, std::vector<int32_t>& ik,
               std::vector<double>& dk) {
  ik.clear();
  dk.clear();
  ik.reserve((size_t)7 * n_func + 2);
  ik.push_back(n_func);
  ik.push_back(n_prim);
  for (int k = 0; k < 7; ++k) ik.insert(ik.end(), iv[k], iv[k] + n_func);
  dk.reserve((size_t)3 * n_func + 2 * n_prim);
  dk.insert(dk.end(), fc, fc + 3 * (size_t)n_func);
  dk.insert(dk.end(), pexp, pexp + n_prim);
  dk.insert(dk.end(), pcoef, pcoef + n_prim);
}

// shells (contiguous runs of equal fn_shell, build_pairs.cpp:170-182), shell
// pairs and the device upload; FE_ERR_* on a malformed basis
int basis_setup(fe_ctx* c, int n_func, int n_prim, const int32_t* fn_atom,
                const int32_t* fn_nprim, const int32_t* fn_off, const int32_t* fn_lx,
                const int32_t* fn_ly, const int32_t* fn_lz, const int32_t* fn_shell,
                const double* fc, const double* pexp, const double* pcoef,
                std::shared_ptr<BasisSetup>& out) {
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
  // shell pairs, candidate slots per class, reference-order indices, and the
  // log range of the product exponents of each class (pair-order keys)
  std::vector<double> emin(nsh), emax(nsh);
  for (int i = 0; i < nsh; ++i) {
    emin[i] = *std::min_element(pexp + sh[i].prim0, pexp + sh[i].prim0 + sh[i].nprim);
    emax[i] = *std::max_element(pexp + sh[i].prim0, pexp + sh[i].prim0 + sh[i].nprim);
  }
  const size_t nsp_all = (size_t)nsh * (nsh + 1) / 2;
  std::vector<int> sp_i(nsp_all), sp_j(nsp_all), sp_slot(nsp_all), sp_gidx(nsp_all);
  auto bs = std::make_shared<BasisSetup>();
  int64_t gtot = 0;
  double pmin[FE_LMAX + 1], pmax[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    pmin[L] = INFINITY;
    pmax[L] = 0.0;
  }
  size_t k = 0;
  for (int i = 0; i < nsh; ++i)
    for (int j = i; j < nsh; ++j, ++k) {
      const int L = sh[i].l + sh[j].l;
      const int64_t np = (int64_t)sh[i].nprim * sh[j].nprim;
      sp_i[k] = i;
      sp_j[k] = j;
      sp_slot[k] = (int)bs->ncand[L];
      sp_gidx[k] = (int)gtot;
      bs->ncand[L] += np;
      gtot += np;
      pmin[L] = std::min(pmin[L], emin[i] + emin[j]);
      pmax[L] = std::max(pmax[L], emax[i] + emax[j]);
      if (gtot > INT32_MAX / 2)
        return fe_set_error(FE_ERR_UNSUPPORTED, "too many primitive pairs (%lld)", (long long)gtot);
    }
  const int nsp = (int)nsp_all;
  // the shell pairs grouped by class (reference order within a class): the
  // dL/dP launch of class L needs only class L's dL/dG
  std::vector<int> sp_cls(nsp_all);
  {
    int cnt[FE_LMAX + 2] = {};
    for (size_t q = 0; q < nsp_all; ++q) ++cnt[sh[sp_i[q]].l + sh[sp_j[q]].l + 1];
    for (int L = 0; L <= FE_LMAX; ++L) cnt[L + 1] += cnt[L];
    for (int L = 0; L <= FE_LMAX + 1; ++L) bs->sp_cls_off[L] = cnt[L];
    for (size_t q = 0; q < nsp_all; ++q) sp_cls[cnt[sh[sp_i[q]].l + sh[sp_j[q]].l]++] = (int)q;
  }
  // the dL/dP tasks: one per (shell pair, function of shell i, function of
  // shell j), class by class in the order above, packed as sp << 8 | mi << 4 | mj
  if (nsp_all >= (1u << 24))
    return fe_set_error(FE_ERR_UNSUPPORTED, "too many shell pairs (%zu) for dL/dP", nsp_all);
  std::vector<uint32_t> tasks;
  for (int L = 0; L <= FE_LMAX; ++L) {
    bs->task_cls_off[L] = (int)tasks.size();
    for (int q = bs->sp_cls_off[L]; q < bs->sp_cls_off[L + 1]; ++q) {
      const int sp = sp_cls[q];
      const int ni = sh[sp_i[sp]].nfn, nj = sh[sp_j[sp]].nfn;
      for (int mi = 0; mi < ni; ++mi)
        for (int mj = 0; mj < nj; ++mj)
          tasks.push_back((uint32_t)sp << 8 | (uint32_t)mi << 4 | (uint32_t)mj);
    }
  }
  bs->task_cls_off[FE_LMAX + 1] = (int)tasks.size();
  bs->device = c->device;
  bs->n_func = n_func;
  bs->n_prim = n_prim;
  bs->nsh = nsh;
  bs->nsp = nsp;
  for (int L = 0; L <= FE_LMAX; ++L) {
    const double lo = std::log(std::isfinite(pmin[L]) ? pmin[L] : 1.0);
    const double w = std::log(std::max(pmax[L], 1e-300)) - lo;
    bs->lpmin[L] = lo;
    bs->lpinv[L] = 16.0 / std::max(w, 1e-12);
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const HostShell& s : sh)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], s.c[a]);
      hi[a] = std::max(hi[a], s.c[a]);
    }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6});
  for (int a = 0; a < 3; ++a) bs->lo[a] = lo[a];
  bs->scale = 255.0 / ext;
  // one device block
  std::vector<int> v_atom(nsh), v_l(nsh), v_fn0(nsh), v_nfn(nsh), v_p0(nsh), v_np(nsh);
  std::vector<double> v_c(3 * nsh);
  for (int i = 0; i < nsh; ++i) {
    v_atom[i] = sh[i].atom; v_l[i] = sh[i].l; v_fn0[i] = sh[i].fn0; v_nfn[i] = sh[i].nfn;
    v_p0[i] = sh[i].prim0; v_np[i] = sh[i].nprim;
    for (int q = 0; q < 3; ++q) v_c[3 * i + q] = sh[i].c[q];
  }
  std::vector<char> hb;
  bs->o_atom = push_bytes(hb, v_atom.data(), nsh);
  bs->o_l = push_bytes(hb, v_l.data(), nsh);
  bs->o_fn0 = push_bytes(hb, v_fn0.data(), nsh);
  bs->o_nfn = push_bytes(hb, v_nfn.data(), nsh);
  bs->o_p0 = push_bytes(hb, v_p0.data(), nsh);
  bs->o_np = push_bytes(hb, v_np.data(), nsh);
  bs->o_c = push_bytes(hb, v_c.data(), 3 * (size_t)nsh);
  bs->o_pexp = push_bytes(hb, pexp, n_prim);
  bs->o_pcoef = push_bytes(hb, pcoef, n_prim);
  bs->o_lx = push_bytes(hb, fn_lx, n_func);
  bs->o_ly = push_bytes(hb, fn_ly, n_func);
  bs->o_lz = push_bytes(hb, fn_lz, n_func);
  bs->o_spi = push_bytes(hb, sp_i.data(), nsp);
  bs->o_spj = push_bytes(hb, sp_j.data(), nsp);
  bs->o_sps = push_bytes(hb, sp_slot.data(), nsp);
  bs->o_spg = push_bytes(hb, sp_gidx.data(), nsp);
  bs->o_spc = push_bytes(hb, sp_cls.data(), nsp);
  bs->o_task = push_bytes(hb, tasks.data(), tasks.size());
  if (hipMalloc(&bs->mem, hb.size()) != hipSuccess) {
    (void)hipGetLastError();
    bs->mem = nullptr;
    return fe_set_error(FE_ERR_ALLOC, "fe_table_build: basis allocation failed");
  }
  bs->bytes = hb.size();
  // synchronous: hb is a local buffer
  if (hipMemcpy(bs->mem, hb.data(), hb.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_HIP, "fe_table_build: basis upload (hipMemcpy) failed");
  }
  const int32_t* iv[7] = {fn_atom, fn_nprim, fn_off, fn_lx, fn_ly, fn_lz, fn_shell};
  basis_key(n_func, n_prim, iv, fc, pexp, pcoef, bs->ikey, bs->dkey);
  out = bs;
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

  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  // the basis: the cached set-up when the arrays are the same as last time
  std::shared_ptr<BasisSetup> bs = std::static_pointer_cast<BasisSetup>(c->basis_cache);
  bool same = false;
  if (bs && bs->n_func == n_func && bs->n_prim == n_prim) {
    std::vector<int32_t> ik;
    std::vector<double> dk;
    const int32_t* iv[7] = {fn_atom, fn_nprim, fn_off, fn_lx, fn_ly, fn_lz, fn_shell};
    basis_key(n_func, n_prim, iv, fc, pexp, pcoef, ik, dk);
    same = ik == bs->ikey && dk.size() == bs->dkey.size() &&
           memcmp(dk.data(), bs->dkey.data(), dk.size() * sizeof(double)) == 0;
  }
  if (!same) {
    c->basis_cache.reset();
    if (int rc = basis_setup(c, n_func, n_prim, fn_atom, fn_nprim, fn_off, fn_lx, fn_ly, fn_lz,
                             fn_shell, fc, pexp, pcoef, bs))
      return rc;
    c->basis_cache = bs;
  }
  hipStream_t st = (hipStream_t)stream;
  fe_table* t = new fe_table();
  t->ctx = c;
  t->has_basis = true;
  t->n_func = n_func;
  t->n_shells = bs->nsh;
  t->n_sp = bs->nsp;
  for (int L = 0; L <= FE_LMAX; ++L) {
    t->n_cand_L[L] = bs->ncand[L];
    t->key_lpmin[L] = bs->lpmin[L];
    t->key_lpinv[L] = bs->lpinv[L];
  }
  for (int a = 0; a < 3; ++a) t->key_lo[a] = bs->lo[a];
  t->key_scale = bs->scale;
  t->basis_mem = bs;
  t->dev_bytes += bs->bytes;
  const char* bm = (const char*)bs->mem;
  t->b.atom = (const int*)(bm + bs->o_atom); t->b.l = (const int*)(bm + bs->o_l);
  t->b.fn0 = (const int*)(bm + bs->o_fn0); t->b.nfn = (const int*)(bm + bs->o_nfn);
  t->b.prim0 = (const int*)(bm + bs->o_p0); t->b.nprim = (const int*)(bm + bs->o_np);
  t->b.c = (const double*)(bm + bs->o_c); t->b.pexp = (const double*)(bm + bs->o_pexp);
  t->b.pcoef = (const double*)(bm + bs->o_pcoef); t->b.fn_lx = (const int*)(bm + bs->o_lx);
  t->b.fn_ly = (const int*)(bm + bs->o_ly); t->b.fn_lz = (const int*)(bm + bs->o_lz);
  t->b.sp_i = (const int*)(bm + bs->o_spi); t->b.sp_j = (const int*)(bm + bs->o_spj);
  t->b.sp_slot = (const int*)(bm + bs->o_sps); t->b.sp_gidx = (const int*)(bm + bs->o_spg);
  t->b.sp_by_class = (const int*)(bm + bs->o_spc);
  t->b.tasks = (const uint32_t*)(bm + bs->o_task);
  for (int L = 0; L <= FE_LMAX + 1; ++L) t->sp_cls_off[L] = bs->sp_cls_off[L];
  for (int L = 0; L <= FE_LMAX + 1; ++L) t->task_cls_off[L] = bs->task_cls_off[L];
  auto fail = [&](int rc) {
    (void)hipStreamSynchronize(st);
    fe_table_free(t);
    return rc;
  };
  // nuclei (the Morton frame of the point sort)
  if (int rc = fe_table_set_atoms(t, n_atoms, Z, xyz)) return fail(rc);
  if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, ptr_kind, st, false))
    return fail(rc);
  *out = t;
  return FE_OK;
}

// The rebuild of a table already in its capacity layout (a first
// fe_table_rebuild has run), on device inputs, with a caller-owned workspace:
// nothing here touches the context's workspace, so a graph that records it
// stays valid whatever the context does later.
int fe_table_rebuild_ws(fe_table* t, const double* dens, int n_points, const double* pts,
                        double tol, int rule, hipStream_t st, char* ws, size_t ws_bytes) {
  if (!t || !t->has_basis || !t->capacity_layout)
    return fe_set_error(FE_ERR_STATE, "rebuild: the table has no capacity layout yet");
  if (int rc = check_build_args(n_points, pts, dens, tol, rule, FE_PTR_DEVICE, "rebuild"))
    return rc;
  if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, FE_PTR_DEVICE, st, true, ws,
                               ws_bytes))
    return rc;
  if (int rc = fe_refresh_g32(t, st)) return rc;
  t->counts_stale = true;
  t->count_stream = st;
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
