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
//  kernel : k_build_cands, one lane per candidate, the shell pairs grouped by
//           (l_i, l_j) at basis set-up and compiled per pair of l's (two
//           launches: l_i + l_j >= 4 on the context's second stream, beside
//           the others and the point sort).  Folded density block P_ij
//           (+ P_ji^T), skip if max|blk| < 1e-15; p, P, K (skip K < 1e-30);
//           etab over the |t| <= L simplex from the block's polynomial about P
//           (fe_pair_etab, the closed form of the 1-D Hermite recurrences);
//           G = (2 pi/p)(2p)^m M0 etab (fe_g_from_etab); the screening bound
//           (fe_etab_bound), reff and the candidate's sum-rule term.
//  kernel : the screen (k_screen_boxes), one wavefront per open candidate.
//           The points are Morton-sorted and cut into boxes of 64; the rules
//           rule 0 (NumPy, table.py:190-194) keep = !(K bound/max(dmin,0.05) < tol)
//           rule 1 (C++, build_pairs.cpp:318-328) keep = !(K bound/max(dmin-reff,0.05) < tol)
//                  and G not all zero and c_i c_j != 0
//           rule 2 (ESPEngine, esp.py:265-270)    keep = !(K bound/max(dmax-reff,0.05) < tol)
//           are decided exactly without visiting every point: the test is
//           monotone in the distance, so rules 0/1 keep a candidate iff some
//           point passes it and rule 2 iff every point does, and a box whose
//           distance bounds settle the test for all its points is decided
//           (or skipped) without reading them.
//  device : ordering and compaction.  Each candidate gets a 32-bit key
//           (class, dropped flag, exponent bucket, Morton code of P); the keys
//           are radix-sorted (stable) before the screen, which walks the open
//           candidates in that order and sets the dropped flag of those it
//           drops; one more stable pass over the top byte puts the kept
//           records of class L first in its range.  A prefix sum over the
//           candidates in reference order gives each kept record its flat term
//           offset and its reference-order index.  Nothing goes back to the
//           host but the counts (fe_table_build, one synchronisation to size
//           the table); fe_table_rebuild does not synchronise at all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"

#define FE_BUILD_MAXPRIM 32
#define FE_BOX 64          // sorted points per screening box
#define FE_SENTINELS 32    // points the screen's first pass tests per candidate

struct ClassInfo {
  int stride[FE_LMAX + 1], nterm[FE_LMAX + 1], nmono[FE_LMAX + 1];
  int64_t cand_off[FE_LMAX + 2];  // candidate slots of class L: [cand_off[L], cand_off[L+1])
};

struct BuildArgs {
  const int* sh_l; const int* sh_fn0; const int* sh_prim0; const int* sh_nprim;
  const double* sh_c; const double* pexp; const double* pcoef;
  const int* fn_cperm;
  const int* sp_i; const int* sp_j; const int* sp_slot; const int* sp_gidx;
  const int* cl_sp; const int* cl_s;
  int cmb_off[FE_NCOMBO + 1];  // candidates of shell-l pair c: cl_*[cmb_off[c], cmb_off[c+1])
  const double* dens; int n_func;
  double* cand[FE_LMAX + 1];
  int* gidx;           // per candidate (class-concatenated slot): reference-order index
  double* cand_srule;  // per candidate: its term of the sum rule
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
//
// One lane per candidate (primitive pair) of a shell pair whose shells have
// l = LI and LJ (LI >= LJ; a pair listed the other way round is handled with
// the shells' roles swapped, which leaves the block's polynomial unchanged).
// Everything is straight-line code in registers: the folded block, its
// polynomial about P, etab (fe_pair_etab), G (fe_g_from_etab) and the
// screening bound (fe_etab_bound).
template <int LI, int LJ>
__device__ __forceinline__ void fe_build_cand(const BuildArgs& a, const int q) {
  constexpr int L = LI + LJ;
  constexpr int NI = (LI + 1) * (LI + 2) / 2, NJ = (LJ + 1) * (LJ + 2) / 2;
  constexpr int NT = FeL<L>::NT, NM = FeL<L>::NM, S = FeL<L>::KOFF;
  const int sp = a.cl_sp[q], s = a.cl_s[q];
  const int si = a.sp_i[sp], sj = a.sp_j[sp];
  const int ii = s / a.sh_nprim[sj], jj = s - ii * a.sh_nprim[sj];
  const int slot = a.sp_slot[sp] + s;
  const int64_t e = a.ci.cand_off[L] + slot;
  a.gidx[e] = a.sp_gidx[sp] + s;
  double* rec = a.cand[L] + (size_t)slot * S;
  const double Ax = a.sh_c[3 * si], Ay = a.sh_c[3 * si + 1], Az = a.sh_c[3 * si + 2];
  const double Bx = a.sh_c[3 * sj], By = a.sh_c[3 * sj + 1], Bz = a.sh_c[3 * sj + 2];
  const double dx = Ax - Bx, dy = Ay - By, dz = Az - Bz;
  const double ab2 = dx * dx + dy * dy + dz * dz;
  const double al = a.pexp[a.sh_prim0[si] + ii], be = a.pexp[a.sh_prim0[sj] + jj];
  const double p = al + be;
  const double K = exp(-al * be / p * ab2);
  if (K < 1e-30) {
    rec[7] = 0.0;
    a.cand_srule[e] = 0.0;
    return;
  }
  // the folded block P_ij (+ P_ji^T), shell X (l = LI) by rows; skip the
  // shell pair if max|blk| < 1e-15 (fastesp/esp.py:185-189)
  const bool swap = LI != LJ && a.sh_l[si] != LI;
  const int sx = swap ? sj : si, sy = swap ? si : sj;
  const int fx = a.sh_fn0[sx], fy = a.sh_fn0[sy];
  const size_t nf = (size_t)a.n_func;
  double w[NI * NJ];
  bool nz = false;
#pragma unroll
  for (int x = 0; x < NI; ++x) {
    const int gx = fx + a.fn_cperm[fx + x];
#pragma unroll
    for (int y = 0; y < NJ; ++y) {
      const int gy = fy + a.fn_cperm[fy + y];
      double d = a.dens[(size_t)gx * nf + gy];
      if (si != sj) d += a.dens[(size_t)gy * nf + gx];
      w[x * NJ + y] = d;
      nz |= !(fabs(d) < 1e-15);
    }
  }
  if (!nz) {
    rec[7] = 8.0;  // no candidate; only the capacity flag
    a.cand_srule[e] = 0.0;
    return;
  }
  const double cc = a.pcoef[a.sh_prim0[si] + ii] * a.pcoef[a.sh_prim0[sj] + jj];
  const double Px = (al * Ax + be * Bx) / p;
  const double Py = (al * Ay + be * By) / p;
  const double Pz = (al * Az + be * Bz) / p;
  const double* X = a.sh_c + 3 * sx;
  const double* Y = a.sh_c + 3 * sy;
  double et[NM];
  fe_pair_etab<LI, LJ>(w, Px - X[0], Py - X[1], Pz - X[2], Px - Y[0], Py - Y[1], Pz - Y[2],
                       0.5 / p, cc, et);
  double g[NT];
  fe_g_from_etab<L>(et, p, g);
  bool anyg = false;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    rec[8 + k] = g[k];
    anyg |= g[k] != 0.0;
  }
#pragma unroll
  for (int k = 8 + NT; k < S; ++k) rec[k] = 0.0;
  const double tp = 2.0 * p;
  const double bound = pow(tp, -1.5) * fe_etab_bound<L>(et, sqrt(tp));
  rec[0] = Px;
  rec[1] = Py;
  rec[2] = Pz;
  rec[3] = p;
  rec[4] = K;
  rec[5] = K * bound;
  rec[6] = 1.6 * sqrt((L + 2.0) / p);
  rec[7] = 1.0 + (anyg ? 2.0 : 0.0) + (cc != 0.0 ? 4.0 : 0.0) + 8.0;
  a.cand_srule[e] = K * pow(M_PI / p, 1.5) * et[0];
}

// The shell-l pairs of one launch, in launch order: range r covers the blocks
// [blk_off[r], blk_off[r+1]) and the candidates of shell-l pair combo[r].
struct CandLaunch {
  int n_r;
  int combo[FE_NCOMBO];
  int blk_off[FE_NCOMBO + 1];
};

// The candidates of several shell-l pairs in one launch (the heavy ones first),
// so that the short ranges run side by side rather than one launch after
// another.  HEAVY: the pairs with l_i + l_j >= 4, which need up to 426 VGPRs;
// the others run in a launch of their own at a higher occupancy.
template <bool HEAVY>
__global__ __launch_bounds__(256) void k_build_cands(const BuildArgs a, const CandLaunch pl) {
  const int b = (int)blockIdx.x;
  int r = 0;
  while (r + 1 < pl.n_r && b >= pl.blk_off[r + 1]) ++r;
  const int cb = pl.combo[r];
  const int c0 = a.cmb_off[cb];
  const int i = (b - pl.blk_off[r]) * 256 + (int)threadIdx.x;
  if (i >= a.cmb_off[cb + 1] - c0) return;
  if constexpr (HEAVY) {
    switch (cb) {
      case 5: fe_build_cand<2, 2>(a, c0 + i); break;
      case 7: fe_build_cand<3, 1>(a, c0 + i); break;
      case 8: fe_build_cand<3, 2>(a, c0 + i); break;
      case 9: fe_build_cand<3, 3>(a, c0 + i); break;
      default: break;
    }
  } else {
    switch (cb) {
      case 0: fe_build_cand<0, 0>(a, c0 + i); break;
      case 1: fe_build_cand<1, 0>(a, c0 + i); break;
      case 2: fe_build_cand<1, 1>(a, c0 + i); break;
      case 3: fe_build_cand<2, 0>(a, c0 + i); break;
      case 4: fe_build_cand<2, 1>(a, c0 + i); break;
      case 6: fe_build_cand<3, 0>(a, c0 + i); break;
      default: break;
    }
  }
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
// a bisection over the patterns finds the switch exactly.  The switch lies
// near (kb/tol + reff)^2; when the test confirms a bracket of 1e-13 relative
// around that, the bisection starts there (about 10 steps instead of 63).
__device__ double fe_pass_threshold(const double kb, const double reff, const double tol,
                                    const int rule) {
  if (!fe_pass(0.0, kb, reff, tol, rule)) return -1.0;
  unsigned long long lo = 0ull, hi = 0x7FF0000000000000ull;  // pass(lo), +inf
  if (fe_pass(INFINITY, kb, reff, tol, rule)) return INFINITY;
  const double d = kb / tol + (rule != 0 ? reff : 0.0);
  const double est = d * d;
  if (est > 0.0 && est < 1e300) {
    const double a = est * (1.0 - 1e-13), b = est * (1.0 + 1e-13);
    if (fe_pass(a, kb, reff, tol, rule) && !fe_pass(b, kb, reff, tol, rule)) {
      lo = (unsigned long long)__double_as_longlong(a);
      hi = (unsigned long long)__double_as_longlong(b);
    }
  }
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
  uint32_t* key;              // pair-order keys and candidate indices (sorted in place)
  int32_t* val;
  double* thr;                // per candidate: the squared-distance threshold, < 0 if decided
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

// The pair-order key of candidate e of class L: class, dropped flag, exponent
// bucket, Morton code of P.  The screen sets the dropped flag of the
// candidates it drops, and a stable pass over the top byte then puts each
// class's kept candidates first, in the order of the full key.
__device__ __forceinline__ uint32_t fe_pair_key(const ScreenArgs& a, const int L, const bool dropped,
                                                const double* r) {
  uint32_t k = (uint32_t)L << 29;
  if (dropped) return k | 1u << 28;
  const int bucket = (int)fmin(fmax((log(r[3]) - a.lpmin[L]) * a.lpinv[L], 0.0), 15.0);
  return k | ((uint32_t)bucket << 24) | fe_morton24(r, a.lo, a.kscale);
}

// the reference-order flag of candidate e: term count and 1 when kept, else 0
__device__ __forceinline__ void fe_emit_flag(const ScreenArgs& a, const int64_t e, const int L,
                                             const bool kept) {
  a.gflag[a.gidx[e]] = kept ? (((unsigned long long)a.ci.nterm[L] << 32) | 1ull) : 0ull;
}

// per-block counters of kept candidates, added to the statistics at the end
__device__ __forceinline__ void fe_flush_counts(int* s_cnt, int* stats) {
  __syncthreads();
  if (threadIdx.x < FE_ST_N && s_cnt[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_cnt[threadIdx.x]);
}

// Every candidate slot: the counters of candidates and K-valid pairs, the
// squared-distance threshold of the test, its pair-order key, and, where no
// box walk is needed (no candidate, rule 1's G/coefficient test, no points,
// or a threshold that settles every distance), the final decision; decided
// candidates get a negative threshold.
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
    if (decided < 0) {
      // FE_SENTINELS points spread over the Morton order: one that passes
      // (rules 0, 1) or fails (rule 2) settles the candidate here, which is
      // most of them; the box walk gets the rest
      const double Px = r[0], Py = r[1], Pz = r[2];
      const int ns = min(a.n_pts, FE_SENTINELS);
      for (int k = 0; k < ns; ++k) {
        const int q = (int)(((int64_t)k * a.n_pts) / ns);
        const double x = a.pts[3 * (size_t)q] - Px;
        const double y = a.pts[3 * (size_t)q + 1] - Py;
        const double z = a.pts[3 * (size_t)q + 2] - Pz;
        if ((fma(x, x, fma(y, y, z * z)) <= thr) == any_mode) {
          decided = any_mode ? 1 : 0;
          break;
        }
      }
    }
    a.thr[e] = decided >= 0 ? -1.0 : thr;
    a.key[e] = fe_pair_key(a, L, decided == 0, r);
    a.val[e] = (int32_t)e;
    if (decided >= 0) {
      fe_emit_flag(a, e, L, decided == 1);
      if (decided == 1) atomicAdd(&s_cnt[FE_ST_KEPT + L], 1);
    }
  }
  fe_flush_counts(s_cnt, a.stats);
}

// The open candidates in pair order, one wavefront each (persistent: wavefront
// w takes the sorted positions w, w + W, ..., reading 64 of them at a time and
// skipping the decided ones).  The rules 0, 1 keep a candidate iff some point
// passes the test (the event is a pass); rule 2 iff every point passes (the
// event is a fail); without an event the decision is the opposite.  The lanes
// take 64 boxes at a time: a box whose distance bounds settle the test for all
// its points decides at once; the ambiguous ones are then read one at a time,
// a point per lane.  So no candidate's work runs on a single lane.  A dropped
// candidate gets the dropped flag in its (sorted) key.
__global__ __launch_bounds__(256) void k_screen_boxes(const ScreenArgs a) {
  __shared__ int s_cnt[FE_ST_N];
  if (threadIdx.x < FE_ST_N) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t w = (int64_t)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const bool any_mode = a.rule != 2;
  for (int64_t k0 = 0; w + k0 * W < a.n_all; k0 += 64) {
    const int64_t my_pos = w + (k0 + lane) * W;
    int my_e = 0;
    bool open = false;
    if (my_pos < a.n_all) {
      my_e = a.val[my_pos];
      open = a.thr[my_e] >= 0.0;
    }
    uint64_t todo = __ballot(open);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int64_t e = __builtin_amdgcn_readlane(my_e, j);
      const int64_t pos = w + (k0 + j) * W;
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
        // the ambiguous boxes one at a time, the 64 lanes over its 64 points,
        // until a point passes (rules 0/1) or fails (rule 2)
        uint64_t amb = __ballot(vb && pass_near && !pass_far);
        bool hit = false;
        while (amb && !hit) {
          const int q = (b0 + __builtin_ctzll(amb)) * FE_BOX + lane;
          amb &= amb - 1;
          bool h = false;
          if (q < a.n_pts) {
            const double x = a.pts[3 * (size_t)q] - Px;
            const double y = a.pts[3 * (size_t)q + 1] - Py;
            const double z = a.pts[3 * (size_t)q + 2] - Pz;
            h = (fma(x, x, fma(y, y, z * z)) <= thr) == any_mode;
          }
          hit = __any(h);
        }
        if (hit) {
          keep = any_mode;
          break;
        }
      }
      if (lane == 0) {
        fe_emit_flag(a, e, L, keep);
        if (keep) atomicAdd(&s_cnt[FE_ST_KEPT + L], 1);
        else a.key[pos] |= 1u << 28;
      }
    }
  }
  fe_flush_counts(s_cnt, a.stats);
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ a, const int64_t n, const int v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i] = v;
}

// Sum rule in two fixed-order steps: block b adds the candidates' terms
// [b * per, (b + 1) * per) (k_srule_part), then one block adds the block sums
// and copies the counters into the table's statistics (k_build_stats).
#define FE_SRULE_BLOCKS 256
__global__ __launch_bounds__(256) void k_srule_part(const double* __restrict__ srule,
                                                    const int64_t n, double* __restrict__ part) {
  __shared__ double s_w[4];
  const int64_t per = (n + FE_SRULE_BLOCKS - 1) / FE_SRULE_BLOCKS;
  const int64_t i0 = blockIdx.x * per, i1 = std::min(n, i0 + per);
  double s = 0.0;
  for (int64_t i = i0 + (int)threadIdx.x; i < i1; i += 256) s += srule[i];
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

// The compaction of every class in three launches (slot map cleared, records
// copied, derived parts filled), each class a range of the launch's grid.
struct CompactArgs {
  const double* cand[FE_LMAX + 1];
  double* rec[FE_LMAX + 1];
  int* flatoff[FE_LMAX + 1];
  int32_t* orig[FE_LMAX + 1];
  int* slot2kept[FE_LMAX + 1];
  int64_t cap[FE_LMAX + 1], ncand[FE_LMAX + 1], cand_off[FE_LMAX + 1];
  int e_blk[FE_LMAX + 2];      // copy blocks of class L: [e_blk[L], e_blk[L+1])
  int s_blk[FE_LMAX + 2];      // slot-map blocks of class L: [s_blk[L], s_blk[L+1])
  int r_blk[FE_LMAX + 2];      // record blocks of class FE_LMAX - k: [r_blk[k], r_blk[k+1])
  const int32_t* sorted;       // candidate indices, each class's kept ones first
  const int* gidx;
  const unsigned long long* gscan;
  const int* stats;            // the kept counts
};

__device__ __forceinline__ int fe_range_of(const int* off, const int b) {
  int L = 0;
  while (L < FE_LMAX && b >= off[L + 1]) ++L;
  return L;
}

__global__ __launch_bounds__(256) void k_slotmap_clear(const CompactArgs a) {
  const int L = fe_range_of(a.s_blk, (int)blockIdx.x);
  const int64_t i = (int64_t)(blockIdx.x - a.s_blk[L]) * 256 + threadIdx.x;
  if (i < a.ncand[L]) a.slot2kept[L][i] = -1;
}

// Kept record k of class L: candidate sorted[k] (k < the kept count), with
// slots 6/7 holding the far-field constants 1/(2p) and K sqrt(pi/p)/2 instead
// of reff and the flags; column 0's thread also writes the flat term offset,
// the reference-order index and the slot -> record map.  One thread per
// copied double, the class's strides known at compile time.
template <int L>
__device__ __forceinline__ void fe_compact_cls(const CompactArgs& a, const int el) {
  constexpr int CS = FeL<L>::KOFF, S = FeL<L>::STRIDE;
  const int n = min(a.stats[FE_ST_KEPT + L], (int)a.cap[L]);
  const int r = el / CS, c = el - r * CS;
  if (r >= n) return;
  const int ge = a.sorted[a.cand_off[L] + r];
  const int slot = (int)(ge - a.cand_off[L]);
  const double* s = a.cand[L] + (size_t)slot * CS;
  double v = s[c];
  if (c == 6) v = 0.5 / s[3];
  else if (c == 7) v = s[4] * (0.88622692545275801365 / sqrt(s[3]));
  a.rec[L][(size_t)r * S + c] = v;
  if (c == 0) {
    const unsigned long long g = a.gscan[a.gidx[ge]];
    a.flatoff[L][r] = (int)(g >> 32);
    a.orig[L][r] = (int32_t)(g & 0xffffffffull);
    a.slot2kept[L][slot] = r;
  }
}

__global__ __launch_bounds__(256) void k_compact(const CompactArgs a) {
  const int L = fe_range_of(a.e_blk, (int)blockIdx.x);
  const int el = (int)(blockIdx.x - a.e_blk[L]) * 256 + (int)threadIdx.x;
  switch (L) {
    case 0: fe_compact_cls<0>(a, el); break;
    case 1: fe_compact_cls<1>(a, el); break;
    case 2: fe_compact_cls<2>(a, el); break;
    case 3: fe_compact_cls<3>(a, el); break;
    case 4: fe_compact_cls<4>(a, el); break;
    case 5: fe_compact_cls<5>(a, el); break;
    case 6: fe_compact_cls<6>(a, el); break;
    default: break;
  }
}

// the derived parts of each kept record (one lane per record): the scaled
// coefficients c g_k and the far-field block; the highest class first, whose
// lanes run longest
__global__ __launch_bounds__(256) void k_fill_record(const CompactArgs a) {
  const int k = fe_range_of(a.r_blk, (int)blockIdx.x);
  const int L = FE_LMAX - k;
  const int i = (int)(blockIdx.x - a.r_blk[k]) * 256 + (int)threadIdx.x;
  if (i < min(a.stats[FE_ST_KEPT + L], (int)a.cap[L]))
    fe_fill_record_any(L, a.rec[L] + (size_t)i * fe_stride(L));
}

// The per-shell-pair lists of a basis, one lane per shell pair: its dL/dP
// tasks (sp << 8 | mi << 4 | mj, from toff[sp]) and its build candidates
// (shell pair and primitive pair i-major, from coff[sp]).
__global__ __launch_bounds__(256) void k_expand_lists(const int* __restrict__ sp_i,
                                                      const int* __restrict__ sp_j,
                                                      const int* __restrict__ nfn,
                                                      const int* __restrict__ nprim,
                                                      const int* __restrict__ toff,
                                                      const int* __restrict__ coff, const int nsp,
                                                      uint32_t* __restrict__ tasks,
                                                      int* __restrict__ cl_sp,
                                                      int* __restrict__ cl_s) {
  const int sp = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (sp >= nsp) return;
  const int si = sp_i[sp], sj = sp_j[sp];
  const int ni = nfn[si], nj = nfn[sj];
  uint32_t* t = tasks + toff[sp];
  for (int mi = 0; mi < ni; ++mi)
    for (int mj = 0; mj < nj; ++mj)
      t[mi * nj + mj] = (uint32_t)sp << 8 | (uint32_t)mi << 4 | (uint32_t)mj;
  const int np = nprim[si] * nprim[sj], c0 = coff[sp];
  for (int q = 0; q < np; ++q) {
    cl_sp[c0 + q] = sp;
    cl_s[c0 + q] = q;
  }
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
         o_gflag = 0, o_thr = 0, o_srule = 0, o_stats = 0, o_spart = 0,
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
  y.o_thr = wo;
  wo += al256(na * sizeof(double));
  y.o_srule = wo;  // per candidate
  wo += al256(na * sizeof(double));
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

// The two candidate launches, each over its shell-l pairs (c = LI (LI + 1) / 2
// + LJ) with candidates, heavy-first
CandLaunch cand_launch(const fe_table* t, const bool heavy) {
  static const int kHeavy[] = {9, 8, 5, 7}, kLight[] = {6, 4, 3, 2, 1, 0};
  CandLaunch pl{};
  const int* order = heavy ? kHeavy : kLight;
  const int n = heavy ? 4 : 6;
  for (int k = 0; k < n; ++k) {
    const int cb = order[k], m = t->cmb_off[cb + 1] - t->cmb_off[cb];
    if (m <= 0) continue;
    pl.combo[pl.n_r] = cb;
    pl.blk_off[pl.n_r + 1] = pl.blk_off[pl.n_r] + (m + 255) / 256;
    ++pl.n_r;
  }
  return pl;
}

// The build proper.  `ws` / `ws_bytes`: a caller-owned workspace of at least
// fe_build_ws_bytes bytes (device inputs only), else the context's.
int build_on_device(fe_table* t, const double* dens, int n_points, const double* pts, double tol,
                    int rule, int ptr_kind, hipStream_t st, bool sized, char* ws = nullptr,
                    size_t ws_bytes = 0, bool fork = true) {
  fe_ctx* c = t->ctx;
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
               o_val = y.o_val, o_gflag = y.o_gflag,
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
  ba.sh_l = t->b.l; ba.sh_fn0 = t->b.fn0; ba.sh_prim0 = t->b.prim0; ba.sh_nprim = t->b.nprim;
  ba.sh_c = t->b.c; ba.pexp = t->b.pexp; ba.pcoef = t->b.pcoef; ba.fn_cperm = t->b.fn_cperm;
  ba.sp_i = t->b.sp_i; ba.sp_j = t->b.sp_j; ba.sp_slot = t->b.sp_slot; ba.sp_gidx = t->b.sp_gidx;
  ba.cl_sp = t->b.cl_sp; ba.cl_s = t->b.cl_s;
  for (int q = 0; q <= FE_NCOMBO; ++q) ba.cmb_off[q] = t->cmb_off[q];
  ba.dens = d_dens;
  ba.n_func = t->n_func;
  for (int L = 0; L <= FE_LMAX; ++L) ba.cand[L] = (double*)(wb + o_cand[L]);
  ba.gidx = (int*)(wb + o_gidx);
  ba.cand_srule = (double*)(wb + o_srule);
  ba.ci = ci;
  // the heavy shell-l pairs on the context's second stream, beside the light
  // ones and the point sort on the caller's stream (`fork` false: all on the
  // caller's stream, for callers that run several builds side by side)
  const CandLaunch pl_heavy = cand_launch(t, true), pl_light = cand_launch(t, false);
  hipStream_t side = st;
  fork = fork && pl_heavy.n_r > 0;
  if (pl_heavy.n_r) {
    if (fork)
      if (int rc = fe_fork(c, st, &side)) return rc;
    hipLaunchKernelGGL(k_build_cands<true>, dim3((unsigned)pl_heavy.blk_off[pl_heavy.n_r]),
                       dim3(256), 0, side, ba, pl_heavy);
  }
  if (pl_light.n_r)
    hipLaunchKernelGGL(k_build_cands<false>, dim3((unsigned)pl_light.blk_off[pl_light.n_r]),
                       dim3(256), 0, st, ba, pl_light);
  FE_HIPCHK(hipGetLastError());

  // ---- points: Morton order, boxes of FE_BOX
  const double* sorted = nullptr;
  if (n_points > 0) {
    const int32_t* perm = nullptr;
    if (int rc = fe_sort_points(c, t, d_pts, n_points, st, &sorted, &perm, wb + o_psort, sort_pts))
      return rc;
    hipLaunchKernelGGL(k_boxes, dim3(n_box), dim3(64), 0, st, sorted, n_points,
                       (double*)(wb + o_box));
  }
  if (fork) {
    if (int rc = fe_join(c, st)) return rc;
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
  sa.n_all = n_all;
  if (n_all > 0) {
    const dim3 g((unsigned)((n_all + 255) / 256));
    hipLaunchKernelGGL(k_screen_prep, g, dim3(256), 0, st, sa);
    // pair order; then the box walk in that order, which flags the keys of the
    // candidates it drops; then a stable pass over the class and flag byte
    if (int rc = fe_radix_sort(key, val, (int)n_all, 32, wb + o_sort, st)) return rc;
    // persistent: 16 workgroups (64 wavefronts) per CU walk the open candidates
    hipLaunchKernelGGL(k_screen_boxes, dim3((unsigned)std::max(1, c->n_cu * 16)), dim3(256), 0,
                       st, sa);
    if (int rc = fe_radix_sort_bits(key, val, (int)n_all, 24, 32, wb + o_sort, st)) return rc;
    if (int rc = fe_scan_excl_u64(gflag, gflag, (int)n_all, wb + o_sort, st)) return rc;
  }
  FE_HIPCHK(hipGetLastError());

  double* spart = (double*)(wb + o_spart);
  hipLaunchKernelGGL(k_srule_part, dim3(FE_SRULE_BLOCKS), dim3(256), 0, st, ba.cand_srule,
                     n_all, spart);
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
  CompactArgs ca{};
  for (int L = 0; L <= FE_LMAX; ++L) {
    const int64_t cap = t->n_cand_L[L] ? t->cap_L[L] : 0;
    ca.cand[L] = ba.cand[L];
    ca.rec[L] = t->d_rec[L];
    ca.flatoff[L] = t->d_flatoff[L];
    ca.orig[L] = t->d_orig[L];
    ca.slot2kept[L] = t->d_slot2kept[L];
    ca.cap[L] = cap;
    ca.ncand[L] = t->n_cand_L[L];
    ca.cand_off[L] = ci.cand_off[L];
    if (cap * fe_host::CSTRIDE[L] > INT32_MAX)
      return fe_set_error(FE_ERR_UNSUPPORTED, "table: %lld records of class %d",
                          (long long)cap, L);
    ca.e_blk[L + 1] = ca.e_blk[L] + (int)((cap * fe_host::CSTRIDE[L] + 255) / 256);
    ca.s_blk[L + 1] = ca.s_blk[L] + (int)((t->n_cand_L[L] + 255) / 256);
  }
  for (int k = 0; k <= FE_LMAX; ++k) {
    const int L = FE_LMAX - k;
    const int64_t cap = t->n_cand_L[L] ? t->cap_L[L] : 0;
    ca.r_blk[k + 1] = ca.r_blk[k] + (int)((cap + 255) / 256);
  }
  ca.sorted = val;
  ca.gidx = ba.gidx;
  ca.gscan = gflag;
  ca.stats = t->d_stats;
  if (ca.s_blk[FE_LMAX + 1] > 0)
    hipLaunchKernelGGL(k_slotmap_clear, dim3((unsigned)ca.s_blk[FE_LMAX + 1]), dim3(256), 0, st, ca);
  if (ca.e_blk[FE_LMAX + 1] > 0) {
    hipLaunchKernelGGL(k_compact, dim3((unsigned)ca.e_blk[FE_LMAX + 1]), dim3(256), 0, st, ca);
    hipLaunchKernelGGL(k_fill_record, dim3((unsigned)ca.r_blk[FE_LMAX + 1]), dim3(256), 0, st, ca);
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
         o_spc = 0, o_task = 0, o_cperm = 0, o_clsp = 0, o_cls = 0;
  int cmb_off[FE_NCOMBO + 1] = {};  // build candidates of shell-l pair c: cl_*[off[c], off[c+1])
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
  std::vector<int> cperm(n_func, 0);
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
    // every Cartesian component of the shell exactly once (in any order)
    {
      unsigned seen = 0;
      for (int q = i; q < j; ++q) {
        // canonical index: x-major, (l,0,0), (l-1,1,0), (l-1,0,1), ...
        const bool neg = fn_lx[q] < 0 || fn_ly[q] < 0 || fn_lz[q] < 0;
        const int lyz = fn_ly[q] + fn_lz[q];
        const int cidx = neg ? 0 : lyz * (lyz + 1) / 2 + fn_lz[q];
        if (neg || cidx >= s.nfn || (seen >> cidx & 1u))
          return fe_set_error(FE_ERR_FORMAT,
                              "shell at function %d: components are not the %d of l=%d", i, s.nfn,
                              s.l);
        seen |= 1u << cidx;
        cperm[i + cidx] = q - i;
      }
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
  // (expanded on the device from per-shell-pair offsets, k_expand_lists)
  std::vector<int> sp_toff(nsp_all), sp_coff(nsp_all);
  {
    int64_t nt = 0;
    for (int L = 0; L <= FE_LMAX; ++L) {
      bs->task_cls_off[L] = (int)nt;
      for (int q = bs->sp_cls_off[L]; q < bs->sp_cls_off[L + 1]; ++q) {
        const int sp = sp_cls[q];
        sp_toff[sp] = (int)nt;
        nt += (int64_t)sh[sp_i[sp]].nfn * sh[sp_j[sp]].nfn;
      }
    }
    if (nt > INT32_MAX) return fe_set_error(FE_ERR_UNSUPPORTED, "too many dL/dP tasks");
    bs->task_cls_off[FE_LMAX + 1] = (int)nt;
  }
  // the build candidates grouped by shell-l pair (l_i >= l_j, either order of
  // the two shells), reference order within a group: sp_coff[sp] is where the
  // shell pair's candidates start
  {
    auto combo = [&](int sp) {
      const int la = sh[sp_i[sp]].l, lb = sh[sp_j[sp]].l;
      const int hi = std::max(la, lb), lo = std::min(la, lb);
      return hi * (hi + 1) / 2 + lo;
    };
    int64_t at[FE_NCOMBO + 1] = {};
    for (int sp = 0; sp < nsp; ++sp)
      at[combo(sp) + 1] += (int64_t)sh[sp_i[sp]].nprim * sh[sp_j[sp]].nprim;
    for (int cb = 0; cb < FE_NCOMBO; ++cb) at[cb + 1] += at[cb];
    for (int cb = 0; cb <= FE_NCOMBO; ++cb) bs->cmb_off[cb] = (int)at[cb];
    for (int sp = 0; sp < nsp; ++sp) {
      int64_t& o = at[combo(sp)];
      sp_coff[sp] = (int)o;
      o += (int64_t)sh[sp_i[sp]].nprim * sh[sp_j[sp]].nprim;
    }
  }
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
  hb.reserve(256 * 24 + 4 * (6 * (size_t)nsh + 4 * (size_t)n_func + 7 * nsp_all) +
             8 * (3 * (size_t)nsh + 2 * (size_t)n_prim));
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
  bs->o_cperm = push_bytes(hb, cperm.data(), cperm.size());
  const size_t o_toff = push_bytes(hb, sp_toff.data(), nsp_all);
  const size_t o_coff = push_bytes(hb, sp_coff.data(), nsp_all);
  // device-only tail: the task and candidate lists, filled by k_expand_lists
  size_t tail = al256(hb.size());
  bs->o_task = tail;
  tail += al256((size_t)bs->task_cls_off[FE_LMAX + 1] * sizeof(uint32_t));
  bs->o_clsp = tail;
  tail += al256((size_t)gtot * sizeof(int));
  bs->o_cls = tail;
  tail += al256((size_t)gtot * sizeof(int));
  if (hipMalloc(&bs->mem, tail) != hipSuccess) {
    (void)hipGetLastError();
    bs->mem = nullptr;
    return fe_set_error(FE_ERR_ALLOC, "fe_table_build: basis allocation failed");
  }
  bs->bytes = tail;
  // synchronous: hb is a local buffer
  if (hipMemcpy(bs->mem, hb.data(), hb.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_HIP, "fe_table_build: basis upload (hipMemcpy) failed");
  }
  {
    char* m = (char*)bs->mem;
    if (nsp > 0)
      hipLaunchKernelGGL(k_expand_lists, dim3((unsigned)((nsp + 255) / 256)), dim3(256), 0, nullptr,
                         (const int*)(m + bs->o_spi), (const int*)(m + bs->o_spj),
                         (const int*)(m + bs->o_nfn), (const int*)(m + bs->o_np),
                         (const int*)(m + o_toff), (const int*)(m + o_coff), nsp,
                         (uint32_t*)(m + bs->o_task), (int*)(m + bs->o_clsp),
                         (int*)(m + bs->o_cls));
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
      (void)hipGetLastError();
      return fe_set_error(FE_ERR_HIP, "fe_table_build: task and candidate lists (k_expand_lists)");
    }
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
  t->b.fn_cperm = (const int*)(bm + bs->o_cperm);
  t->b.cl_sp = (const int*)(bm + bs->o_clsp);
  t->b.cl_s = (const int*)(bm + bs->o_cls);
  for (int q = 0; q <= FE_NCOMBO; ++q) t->cmb_off[q] = bs->cmb_off[q];
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
                        double tol, int rule, hipStream_t st, char* ws, size_t ws_bytes,
                        bool fork) {
  if (!t || !t->has_basis || !t->capacity_layout)
    return fe_set_error(FE_ERR_STATE, "rebuild: the table has no capacity layout yet");
  if (int rc = check_build_args(n_points, pts, dens, tol, rule, FE_PTR_DEVICE, "rebuild"))
    return rc;
  if (int rc = build_on_device(t, dens, n_points, pts, tol, rule, FE_PTR_DEVICE, st, true, ws,
                               ws_bytes, fork))
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
