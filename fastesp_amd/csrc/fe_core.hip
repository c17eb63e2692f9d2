// libfastesp_hip: context, device pair tables and the ESP forward evaluation.
//
// Device table layout ("per-L records"): the pairs of each class
// L = l_A + l_B (0..FE_LMAX) are stored contiguously, one 64-byte aligned
// record of FeL<L>::STRIDE doubles per primitive pair:
//
//     [0] Px [1] Py [2] Pz [3] p [4] K [5] K*bound [6] 1/(2p) [7] K sqrt(pi/p)/2
//     [8 .. 8+NT)  g_k in canonical key order
//     [8+NT .. 8+NT+ND)  the integer multiples c g_k the field contraction uses
//     (fe_fill_dcoef<L>), zero padded to a multiple of 8
//     [KOFF .. KOFF+NK)  the far-field block (fe_fill_kelvin<L>): the pair's
//     potential where every Boys order is asymptotic, as |u| H(u) with
//     u = R/|R|^2 and H a polynomial of degree L, zero padded to a multiple of 8
//
// A wavefront evaluates 64 points (x points-per-lane) against the same pair,
// so the whole record is wave-uniform and is fetched with scalar loads into
// SGPRs; g_k then feeds v_fma_f64 directly as the scalar operand.  The
// per-point state (R, Boys vector, monomials) lives in VGPRs, the Boys grid of
// the kernel's top order is staged once per workgroup in LDS.
//
// Work decomposition: grid = (point blocks) x (pair chunks of one class).
// Every (class, chunk) writes one row of a partial-sum buffer; a second
// kernel adds the rows in a fixed order plus the nuclear term, so results are
// bitwise reproducible run to run (the reference's two-kernel split,
// fastesp/native/esp_gpu.cu:161-229, with the chunk count sized to the chip
// instead of a fixed 4096).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/fastesp_hip.h"
#include "fe_boys.h"
#include "fe_lclass.inc"
#include "fe_internal.h"
#include "fe_reduce.h"

#define FE_VERSION 101


namespace {
thread_local std::string g_err;
}

int fe_set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" int fe_version(void) { return FE_VERSION; }
extern "C" const char* fe_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Boys grid (host, long double)
// ---------------------------------------------------------------------------
static void boys_series(long double T, int jmax, long double* F) {
  const long double e = expl(-T);
  for (int j = 0; j <= jmax; ++j) {
    long double term = 1.0L / (2 * j + 1), sum = term;
    for (int i = 1; i < 4000; ++i) {
      term *= (2.0L * T) / (2 * j + 2 * i + 1);
      sum += term;
      if (term < 1e-24L * sum) break;
    }
    F[j] = e * sum;
  }
}

// Degree-(n-1) polynomial c[0..n-1] in dt economized to degree 7 over
// |dt| <= h: the dt^9 and dt^8 terms are rewritten through the Chebyshev
// polynomials (x = dt/h)
//   x^9 = T_9/256 + 9/4 x^7 - 27/16 x^5 + 15/32 x^3 - 9/256 x,
//   x^8 = T_8/128 + 2 x^6 - 5/4 x^4 + 1/4 x^2 - 1/128,
// and the T_9, T_8 parts (at most h^9/256 |c_9|, h^8/128 |c_8|) dropped.
static void economize_to_7(long double* c, const long double h) {
  const long double h2 = h * h, h4 = h2 * h2, h6 = h4 * h2, h8 = h4 * h4;
  const long double a9 = c[9];
  c[7] += a9 * 2.25L * h2;
  c[5] -= a9 * 1.6875L * h4;
  c[3] += a9 * 0.46875L * h6;
  c[1] -= a9 * 0.03515625L * h8;
  const long double a8 = c[8];
  c[6] += a8 * 2.0L * h2;
  c[4] -= a8 * 1.25L * h4;
  c[2] += a8 * 0.25L * h6;
  c[0] -= a8 * h8 / 128.0L;
  c[8] = c[9] = 0.0L;
}

static std::vector<double> make_boys_tables() {
  const int jmax = FE_BOYS_MMAX + FE_BOYS_NTAYLOR - 1;
  std::vector<double> tab((size_t)(FE_BOYS_MMAX + 1) * FE_BOYS_NR * FE_BOYS_COLS);
  std::vector<long double> F(jmax + 1);
  long double fact[FE_BOYS_NTAYLOR];
  fact[0] = 1.0L;
  for (int k = 1; k < FE_BOYS_NTAYLOR; ++k) fact[k] = fact[k - 1] * k;
  for (int r = 0; r < FE_BOYS_NR; ++r) {
    const long double T = (long double)r / 8.0L;
    boys_series(T, jmax, F.data());
    for (int M = 0; M <= FE_BOYS_MMAX; ++M) {
      // Taylor coefficients of F_M(T_r + dt): d^k F_M / dT^k = (-1)^k F_{M+k}
      long double c[FE_BOYS_NTAYLOR];
      for (int k = 0; k < FE_BOYS_NTAYLOR; ++k)
        c[k] = ((k & 1) ? -1.0L : 1.0L) * F[M + k] / fact[k];
      economize_to_7(c, 0.5L * (long double)FE_BOYS_DT);
      double* row = &tab[((size_t)M * FE_BOYS_NR + r) * FE_BOYS_COLS];
      for (int k = 0; k < FE_BOYS_NK; ++k) row[k] = (double)c[k];
      row[FE_BOYS_NK] = (double)expl(-T);
    }
  }
  return tab;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
int fe_ctx_reserve(fe_ctx* c, int slot, size_t bytes) {
  if (c->ws_bytes[slot] >= bytes) return FE_OK;
  if (c->ws[slot]) FE_HIPCHK(hipFree(c->ws[slot]));
  c->ws[slot] = nullptr;
  c->ws_bytes[slot] = 0;
  size_t b = std::max<size_t>(bytes, 4096);
  b = (b + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  if (hipMalloc(&c->ws[slot], b) != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_ALLOC, "device allocation of %zu bytes failed", b);
  }
  c->ws_bytes[slot] = b;
  return FE_OK;
}

int fe_prof_begin(fe_ctx* c, int phase, hipStream_t st) {
  if (!c->prof_on) return -1;
  if (c->prof_n == c->prof.size()) {
    FeProfEv e{};
    if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
    c->prof.push_back(e);
  }
  FeProfEv& e = c->prof[c->prof_n];
  e.phase = phase;
  if (hipEventRecord(e.a, st) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  return (int)c->prof_n++;
}

void fe_prof_end(fe_ctx* c, int slot, hipStream_t st) {
  if (slot < 0) return;
  if (hipEventRecord(c->prof[slot].b, st) != hipSuccess) {
    (void)hipGetLastError();
    c->prof[slot].phase = -1;
  }
}

extern "C" int fe_profile(fe_ctx* c, int op, double* ms, int64_t* count, int n_phase) {
  if (!c) return fe_set_error(FE_ERR_ARG, "fe_profile: null context");
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  if (op == 1) {
    c->prof_n = 0;
    c->prof_on = true;
    return FE_OK;
  }
  if (op != 0) return fe_set_error(FE_ERR_ARG, "fe_profile: op must be 0 (stop) or 1 (start)");
  c->prof_on = false;
  if (ms)
    for (int i = 0; i < n_phase; ++i) ms[i] = 0.0;
  if (count)
    for (int i = 0; i < n_phase; ++i) count[i] = 0;
  for (size_t k = 0; k < c->prof_n; ++k) {
    const FeProfEv& e = c->prof[k];
    if (e.phase < 0 || e.phase >= n_phase) continue;
    FE_HIPCHK(hipEventSynchronize(e.b));
    float t = 0.f;
    FE_HIPCHK(hipEventElapsedTime(&t, e.a, e.b));
    if (ms) ms[e.phase] += t;
    if (count) count[e.phase] += 1;
  }
  c->prof_n = 0;
  return FE_OK;
}

extern "C" int fe_init(int device, fe_ctx** out) {
  if (!out) return fe_set_error(FE_ERR_ARG, "fe_init: null output pointer");
  *out = nullptr;
  int ndev = 0;
  FE_HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fe_set_error(FE_ERR_ARG, "fe_init: device %d out of range (%d devices)", device, ndev);
  FE_HIPCHK(hipSetDevice(device));
  fe_ctx* c = new fe_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return fe_set_error(FE_ERR_HIP, "hipGetDeviceProperties failed");
  }
  c->n_cu = prop.multiProcessorCount;
  std::vector<double> tab = make_boys_tables();
  if (hipMalloc(&c->d_boys, tab.size() * sizeof(double)) != hipSuccess) {
    delete c;
    return fe_set_error(FE_ERR_ALLOC, "fe_init: Boys table allocation failed");
  }
  if (hipMemcpy(c->d_boys, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(c->d_boys);
    delete c;
    return fe_set_error(FE_ERR_HIP, "fe_init: Boys table upload failed");
  }
  c->target_wg = 32 * c->n_cu;
  // the second stream of the context (table build, dL/dG classes), created
  // here so that no later call creates one (say, while a graph is captured)
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->side2) (void)hipStreamDestroy(c->side2);
    if (c->side) (void)hipStreamDestroy(c->side);
    (void)hipFree(c->d_boys);
    delete c;
    return fe_set_error(FE_ERR_HIP, "fe_init: second stream creation failed");
  }
  if (int rc = fe_noalloc_init(c)) {
    (void)hipEventDestroy(c->ev_join2);
    (void)hipEventDestroy(c->ev_join);
    (void)hipEventDestroy(c->ev_fork);
    (void)hipStreamDestroy(c->side2);
    (void)hipStreamDestroy(c->side);
    (void)hipFree(c->d_boys);
    delete c;
    return rc;
  }
  *out = c;
  return FE_OK;
}

extern "C" int fe_finalize(fe_ctx* c) {
  if (!c) return fe_set_error(FE_ERR_ARG, "fe_finalize: null context");
  FE_HIPCHK(hipSetDevice(c->device));
  for (hipStream_t s : {c->side, c->side2})
    if (s) {
      (void)hipStreamSynchronize(s);  // its launches read the workspace slots
      (void)hipStreamDestroy(s);
    }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
  for (int s = 0; s < FE_WS_SLOTS; ++s)
    if (c->ws[s]) (void)hipFree(c->ws[s]);
  c->basis_cache.reset();
  if (c->d_boys) (void)hipFree(c->d_boys);
  for (FeProfEv& e : c->prof) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  delete c;
  return FE_OK;
}

extern "C" int fe_ctx_trim(fe_ctx* c) {
  if (!c) return fe_set_error(FE_ERR_ARG, "fe_ctx_trim: null context");
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  // work still queued on any stream may read the slots
  FE_HIPCHK(hipDeviceSynchronize());
  for (int s = 0; s < FE_WS_SLOTS; ++s) {
    if (c->ws[s]) FE_HIPCHK(hipFree(c->ws[s]));
    c->ws[s] = nullptr;
    c->ws_bytes[s] = 0;
  }
  c->basis_cache.reset();
  return FE_OK;
}

int fe_fork(fe_ctx* c, hipStream_t st, hipStream_t* side) {
  FE_HIPCHK(hipEventRecord(c->ev_fork, st));
  FE_HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  *side = c->side;
  return FE_OK;
}

int fe_join(fe_ctx* c, hipStream_t st) {
  FE_HIPCHK(hipEventRecord(c->ev_join, c->side));
  FE_HIPCHK(hipStreamWaitEvent(st, c->ev_join, 0));
  return FE_OK;
}

int fe_fork_n(fe_ctx* c, hipStream_t st, int n, hipStream_t* s) {
  s[0] = s[1] = s[2] = st;
  if (n < 2) return FE_OK;
  FE_HIPCHK(hipEventRecord(c->ev_fork, st));
  FE_HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  s[1] = c->side;
  if (n >= 3) {
    FE_HIPCHK(hipStreamWaitEvent(c->side2, c->ev_fork, 0));
    s[2] = c->side2;
  }
  return FE_OK;
}

int fe_join_n(fe_ctx* c, hipStream_t st, int n) {
  if (n < 2) return FE_OK;
  FE_HIPCHK(hipEventRecord(c->ev_join, c->side));
  FE_HIPCHK(hipStreamWaitEvent(st, c->ev_join, 0));
  if (n >= 3) {
    FE_HIPCHK(hipEventRecord(c->ev_join2, c->side2));
    FE_HIPCHK(hipStreamWaitEvent(st, c->ev_join2, 0));
  }
  return FE_OK;
}

void fe_lpt(const double* work, int n, int* order, int* q) {
  for (int L = 0; L <= FE_LMAX; ++L) order[L] = L;
  std::stable_sort(order, order + FE_LMAX + 1, [&](int a, int b) { return work[a] > work[b]; });
  double load[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k <= FE_LMAX; ++k) {
    int best = 0;
    for (int j = 1; j < n; ++j)
      if (load[j] < load[best]) best = j;
    load[best] += work[order[k]];
    q[k] = best;
  }
}

extern "C" int fe_set_tuning(fe_ctx* c, int key, int64_t value) {
  if (!c) return fe_set_error(FE_ERR_ARG, "fe_set_tuning: null context");
  FE_CTX_LOCK(c);
  switch (key) {
    case 0:  // points per lane of every class
      if (value != 1 && value != 2 && value != 4)
        return fe_set_error(FE_ERR_ARG, "points per lane must be 1, 2 or 4");
      for (int L = 0; L <= FE_LMAX; ++L) c->ppl_L[L] = c->ppl_grad_L[L] = (int)value;
      return FE_OK;
    case 10: case 11: case 12: case 13: case 14: case 15: case 16:  // per class L = key-10
      if (key - 10 > FE_LMAX) return fe_set_error(FE_ERR_ARG, "class out of range");
      if (value != 1 && value != 2 && value != 4)
        return fe_set_error(FE_ERR_ARG, "points per lane must be 1, 2 or 4");
      c->ppl_L[key - 10] = c->ppl_grad_L[key - 10] = (int)value;
      return FE_OK;
    case 1:
      if (value < 1 || value > (1 << 24)) return fe_set_error(FE_ERR_ARG, "bad target workgroups");
      c->target_wg = (int)value;
      return FE_OK;
    case 3:
      c->sort_points = value != 0;
      return FE_OK;
    case 5:
      c->pair_order = value != 0;
      return FE_OK;
    case 6:
      c->dg_streams = value <= 0 ? 1 : value == 1 ? 2 : (int)std::min<int64_t>(value, 3);
      return FE_OK;
    case 7:
      c->split_eval = value != 0 ? 1 : 0;
      c->split_eval_grad = value == 1 ? 1 : value >= 2 ? 2 : 0;
      return FE_OK;
    case 20: case 21: case 22: case 23: case 24: case 25: case 26:  // FP32 contraction of class L
      c->mixed_L[key - 20] = value != 0;
      return FE_OK;
    case 4:  // mixed-precision routing threshold, in units of 1e-12 a.u.
      if (value < 0) return fe_set_error(FE_ERR_ARG, "mixed threshold must be >= 0");
      c->mixed_tau = (double)value * 1e-12;
      return FE_OK;
    default:
      return fe_set_error(FE_ERR_ARG, "unknown tuning key %d", key);
  }
}

// ---------------------------------------------------------------------------
// table upload (pairs.bin fields -> per-L records)
// ---------------------------------------------------------------------------
static int class_of_nterm(int nt) {
  for (int L = 0; L <= FE_LMAX; ++L)
    if (fe_host::NTERM[L] == nt) return L;
  return -1;
}

int fe_table_layout(fe_table* t, const int64_t* cap, int64_t n_slots) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t o_stats = off;
  off += al(FE_ST_SUMRULE_BYTE + sizeof(double));
  size_t o_rec[FE_LMAX + 1], o_off[FE_LMAX + 1], o_orig[FE_LMAX + 1], o_s2k[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_rec[L] = off;
    off += al((size_t)cap[L] * fe_host::STRIDE[L] * sizeof(double));
    o_off[L] = off;
    off += al((size_t)cap[L] * sizeof(int));
    o_orig[L] = off;
    off += al((size_t)cap[L] * sizeof(int32_t));
  }
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_s2k[L] = off;
    off += n_slots ? al((size_t)t->n_cand_L[L] * sizeof(int)) : 0;
  }
  void* m = nullptr;
  if (hipMalloc(&m, off) != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_ALLOC, "table: allocation of %zu bytes failed", off);
  }
  if (t->d_mem) {
    (void)hipFree(t->d_mem);
    t->dev_bytes -= t->mem_bytes;
  }
  t->d_mem = m;
  t->mem_bytes = off;
  t->dev_bytes += off;
  char* base = (char*)m;
  t->d_stats = (int*)(base + o_stats);
  t->d_sum_rule = (double*)(base + o_stats + FE_ST_SUMRULE_BYTE);
  for (int L = 0; L <= FE_LMAX; ++L) {
    t->cap_L[L] = cap[L];
    t->d_rec[L] = (double*)(base + o_rec[L]);
    t->d_flatoff[L] = (int*)(base + o_off[L]);
    t->d_orig[L] = (int32_t*)(base + o_orig[L]);
    t->d_slot2kept[L] = n_slots ? (int*)(base + o_s2k[L]) : nullptr;
  }
  return FE_OK;
}

int fe_table_sync_counts(fe_table* t) {
  if (!t->counts_stale) return FE_OK;
  FE_HIPCHK(hipSetDevice(t->ctx->device));
  FE_HIPCHK(hipStreamSynchronize(t->count_stream));
  int st[FE_ST_N];
  double sr = 0.0;
  FE_HIPCHK(hipMemcpy(st, t->d_stats, sizeof st, hipMemcpyDeviceToHost));
  FE_HIPCHK(hipMemcpy(&sr, t->d_sum_rule, sizeof sr, hipMemcpyDeviceToHost));
  int64_t np = 0, nt = 0, nc = 0;
  for (int L = 0; L <= FE_LMAX; ++L) {
    t->n_L[L] = st[FE_ST_KEPT + L];
    np += st[FE_ST_KEPT + L];
    nt += (int64_t)st[FE_ST_KEPT + L] * fe_host::NTERM[L];
    nc += st[FE_ST_VALID + L];
  }
  t->n_pairs = (int)np;
  t->n_terms = (int)nt;
  t->n_candidates = nc;
  t->sum_rule = sr;
  t->counts_stale = false;
  return FE_OK;
}

std::vector<int> fe_pair_order(const double* rec4, size_t stride, int64_t n) {
  std::vector<int> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int)i;
  if (n < 2) return order;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double pmin = INFINITY, pmax = -INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    const double* r = rec4 + i * stride;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], r[a]);
      hi[a] = std::max(hi[a], r[a]);
    }
    pmin = std::min(pmin, r[3]);
    pmax = std::max(pmax, r[3]);
  }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-12});
  // exponent bucket b = floor(16 (ln p - ln pmin) / (ln pmax - ln pmin)), clamped
  // to 15, as the count of the 15 bucket edges p_k = pmin (pmax/pmin)^(k/16) <= p
  const double lpmin = std::log(pmin), lpw = std::max(std::log(pmax) - lpmin, 1e-12);
  double edge[15];
  for (int k = 0; k < 15; ++k) edge[k] = std::exp(lpmin + lpw * (k + 1) / 16.0);
  auto spread10 = [](uint64_t x) {  // bit b of x to bit 3b
    x &= 0x3ff;
    x = (x | (x << 16)) & 0x030000ffull;
    x = (x | (x << 8)) & 0x0300f00full;
    x = (x | (x << 4)) & 0x030c30c3ull;
    x = (x | (x << 2)) & 0x09249249ull;
    return x;
  };
  std::vector<uint64_t> key(n);
  for (int64_t i = 0; i < n; ++i) {
    const double* r = rec4 + i * stride;
    uint64_t code = 0;
    for (int a = 0; a < 3; ++a)
      code |= spread10((uint64_t)std::min(1023.0, std::max(0.0, (r[a] - lo[a]) / ext * 1023.0)))
              << a;
    const uint64_t bucket = (uint64_t)(std::upper_bound(edge, edge + 15, r[3]) - edge);
    key[i] = (bucket << 30) | code;
  }
  // stable LSD radix sort of the 34-bit keys, 12 bits a pass
  std::vector<int> tmp(n);
  std::vector<int64_t> cnt(1 << 12);
  for (int shift = 0; shift < 34; shift += 12) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < n; ++i) ++cnt[(key[i] >> shift) & 0xfff];
    int64_t sum = 0;
    for (auto& c : cnt) {
      const int64_t v = c;
      c = sum;
      sum += v;
    }
    for (int64_t i = 0; i < n; ++i) {
      const int o = order[i];
      tmp[cnt[(key[o] >> shift) & 0xfff]++] = o;
    }
    order.swap(tmp);
  }
  return order;
}

int fe_table_set_atoms(fe_table* t, int n_atoms, const double* Z, const double* xyz) {
  t->n_atoms = n_atoms;
  if (n_atoms <= 0) return FE_OK;
  std::vector<double> a((size_t)4 * n_atoms);
  for (int i = 0; i < n_atoms; ++i) {
    a[4 * i + 0] = xyz[3 * i + 0];
    a[4 * i + 1] = xyz[3 * i + 1];
    a[4 * i + 2] = xyz[3 * i + 2];
    a[4 * i + 3] = Z[i];
  }
  t->h_atoms = a;
  if (hipMalloc(&t->d_atoms, a.size() * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_ALLOC, "table: atom allocation failed");
  }
  FE_HIPCHK(hipMemcpy(t->d_atoms, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
  t->dev_bytes += a.size() * sizeof(double);
  return FE_OK;
}

extern "C" int fe_table_from_pairs(fe_ctx* c, int n_pairs, int n_terms, const double* pp_exp,
                                   const double* pp_cx, const double* pp_cy, const double* pp_cz,
                                   const double* pp_K, const int32_t* pp_nt,
                                   const int32_t* pp_off, const int32_t* pp_nmono,
                                   const int32_t* pp_monooff, const int32_t* mo_parent,
                                   const int32_t* mo_axis, const int32_t* tm_aidx,
                                   const int32_t* tm_m, const double* tm_g, int n_atoms,
                                   const double* Z, const double* xyz, fe_table** out) {
  if (!c || !out) return fe_set_error(FE_ERR_ARG, "fe_table_from_pairs: null argument");
  *out = nullptr;
  if (n_pairs < 0 || n_terms < 0 || n_atoms < 0)
    return fe_set_error(FE_ERR_ARG, "fe_table_from_pairs: negative size");
  if (n_pairs > 0 && (!pp_exp || !pp_cx || !pp_cy || !pp_cz || !pp_K || !pp_nt || !pp_off ||
                      !pp_nmono || !pp_monooff || !mo_parent || !mo_axis || !tm_aidx || !tm_m ||
                      !tm_g))
    return fe_set_error(FE_ERR_ARG, "fe_table_from_pairs: null array");
  if (n_atoms > 0 && (!Z || !xyz)) return fe_set_error(FE_ERR_ARG, "fe_table_from_pairs: null atoms");
  std::vector<double> rec[FE_LMAX + 1];
  std::vector<int32_t> orig[FE_LMAX + 1];
  for (int i = 0; i < n_pairs; ++i) {
    const int nt = pp_nt[i];
    const int L = class_of_nterm(nt);
    if (L < 0)
      return fe_set_error(FE_ERR_UNSUPPORTED,
                          "pair %d: %d terms matches no class L <= %d", i, nt, FE_LMAX);
    const int off = pp_off[i], mo = pp_monooff[i];
    if (off < 0 || (int64_t)off + nt > n_terms)
      return fe_set_error(FE_ERR_FORMAT, "pair %d: term range out of bounds", i);
    if (pp_nmono[i] != fe_host::NMONO[L] || mo < 0)
      return fe_set_error(FE_ERR_FORMAT, "pair %d: %d monomials, class %d needs %d", i,
                          pp_nmono[i], L, fe_host::NMONO[L]);
    for (int j = 0; j < nt; ++j) {
      const int* k = fe_host::KEYS[fe_host::KEY_OFF[L] + j];
      if (tm_m[off + j] != k[3] || tm_aidx[off + j] != k[4])
        return fe_set_error(FE_ERR_FORMAT, "pair %d term %d: (aidx,m)=(%d,%d), canonical (%d,%d)",
                            i, j, tm_aidx[off + j], tm_m[off + j], k[4], k[3]);
    }
    for (int j = 1; j < fe_host::NMONO[L]; ++j) {
      const int* m = fe_host::MONOS[fe_host::MONO_OFF[L] + j];
      if (mo_parent[mo + j] != m[3] || mo_axis[mo + j] != m[4])
        return fe_set_error(FE_ERR_FORMAT, "pair %d monomial %d: non-canonical schedule", i, j);
    }
    const int S = fe_host::STRIDE[L];
    const size_t base = rec[L].size();
    rec[L].resize(base + S, 0.0);
    double* r = &rec[L][base];
    r[0] = pp_cx[i];
    r[1] = pp_cy[i];
    r[2] = pp_cz[i];
    r[3] = pp_exp[i];
    r[4] = pp_K[i];
    r[5] = INFINITY;  // K*bound unknown for uploaded tables
    r[6] = 0.5 / pp_exp[i];                        // 1/(2p)
    r[7] = pp_K[i] * (0.88622692545275801365 / sqrt(pp_exp[i]));  // K sqrt(pi/p)/2
    for (int j = 0; j < nt; ++j) r[8 + j] = tm_g[off + j];
    fe_fill_record_any(L, r);
    orig[L].push_back(i);
  }
  if (c->pair_order)
    for (int L = 0; L <= FE_LMAX; ++L) {
      const int S = fe_host::STRIDE[L];
      const int64_t nL = (int64_t)orig[L].size();
      const std::vector<int> ord = fe_pair_order(rec[L].data(), S, nL);
      std::vector<double> r2(rec[L].size());
      std::vector<int32_t> o2(nL);
      for (int64_t k = 0; k < nL; ++k) {
        std::copy(rec[L].begin() + (size_t)ord[k] * S, rec[L].begin() + (size_t)(ord[k] + 1) * S,
                  r2.begin() + (size_t)k * S);
        o2[k] = orig[L][ord[k]];
      }
      rec[L].swap(r2);
      orig[L].swap(o2);
    }
  FE_HIPCHK(hipSetDevice(c->device));
  fe_table* t = new fe_table();
  t->ctx = c;
  t->n_pairs = n_pairs;
  t->n_terms = n_terms;
  auto fail = [&](int rc) {
    fe_table_free(t);
    return rc;
  };
  // flat term offset of every record: term counts in reference order, scanned
  std::vector<int32_t> toff(n_pairs + 1, 0);
  for (int i = 0; i < n_pairs; ++i) toff[i + 1] = toff[i] + pp_nt[i];
  int64_t cap[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) cap[L] = t->n_L[L] = (int64_t)orig[L].size();
  int rc = fe_table_layout(t, cap, 0);
  if (rc != FE_OK) return fail(rc);
  std::vector<int> st(FE_ST_N, 0);
  for (int L = 0; L <= FE_LMAX; ++L) {
    st[FE_ST_KEPT + L] = st[FE_ST_VALID + L] = st[FE_ST_KVALID + L] = (int)t->n_L[L];
    if (!t->n_L[L]) continue;
    std::vector<int> fo(t->n_L[L]);
    for (int64_t k = 0; k < t->n_L[L]; ++k) fo[k] = toff[orig[L][k]];
    FE_HIPCHK(hipMemcpy(t->d_rec[L], rec[L].data(), rec[L].size() * sizeof(double),
                        hipMemcpyHostToDevice));
    FE_HIPCHK(hipMemcpy(t->d_flatoff[L], fo.data(), fo.size() * sizeof(int),
                        hipMemcpyHostToDevice));
    FE_HIPCHK(hipMemcpy(t->d_orig[L], orig[L].data(), orig[L].size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  }
  const double sr = 0.0;
  FE_HIPCHK(hipMemcpy(t->d_stats, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
  FE_HIPCHK(hipMemcpy(t->d_sum_rule, &sr, sizeof sr, hipMemcpyHostToDevice));
  rc = fe_table_set_atoms(t, n_atoms, Z, xyz);
  if (rc != FE_OK) return fail(rc);
  *out = t;
  return FE_OK;
}

extern "C" int fe_table_free(fe_table* t) {
  if (!t) return fe_set_error(FE_ERR_ARG, "fe_table_free: null table");
  (void)hipSetDevice(t->ctx->device);
  // a rebuild may still be running on its stream
  if (t->counts_stale) (void)hipStreamSynchronize(t->count_stream);
  for (int L = 0; L <= FE_LMAX; ++L)
    if (t->d_rec32[L]) (void)hipFree(t->d_rec32[L]);
  if (t->d_mem) (void)hipFree(t->d_mem);
  if (t->d_atoms) (void)hipFree(t->d_atoms);
  t->basis_mem.reset();
  delete t;
  return FE_OK;
}

extern "C" int fe_table_info(const fe_table* t, int64_t* info, int n_info) {
  if (!t || !info) return fe_set_error(FE_ERR_ARG, "fe_table_info: null argument");
  if (int rc = fe_table_sync_counts(const_cast<fe_table*>(t))) return rc;
  int64_t v[16] = {0};
  v[0] = t->n_pairs;
  v[1] = t->n_terms;
  v[2] = t->n_atoms;
  for (int L = 0; L <= FE_LMAX; ++L) v[3 + L] = t->n_L[L];
  v[10] = (int64_t)t->dev_bytes;
  v[11] = t->n_candidates;
  for (int i = 0; i < n_info && i < 16; ++i) info[i] = v[i];
  return FE_OK;
}

extern "C" int fe_table_export(const fe_table* t, double* pp_exp, double* pp_cx, double* pp_cy,
                               double* pp_cz, double* pp_K, int32_t* pp_nt, int32_t* pp_off,
                               int32_t* pp_nmono, int32_t* pp_monooff, int32_t* mo_parent,
                               int32_t* mo_axis, int32_t* tm_aidx, int32_t* tm_m, double* tm_g,
                               double* pp_kbound, double* pp_reff) {
  if (!t || !pp_exp || !pp_cx || !pp_cy || !pp_cz || !pp_K || !pp_nt || !pp_off || !pp_nmono ||
      !pp_monooff || !mo_parent || !mo_axis || !tm_aidx || !tm_m || !tm_g)
    return fe_set_error(FE_ERR_ARG, "fe_table_export: null argument");
  if (int rc = fe_table_sync_counts(const_cast<fe_table*>(t))) return rc;
  FE_HIPCHK(hipSetDevice(t->ctx->device));
  // (orig position, class, index within class), sorted by original position
  std::vector<std::tuple<int32_t, int, int>> order;
  order.reserve(t->n_pairs);
  std::vector<std::vector<double>> rec(FE_LMAX + 1);
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!t->n_L[L]) continue;
    rec[L].resize((size_t)t->n_L[L] * fe_host::STRIDE[L]);
    FE_HIPCHK(hipMemcpy(rec[L].data(), t->d_rec[L], rec[L].size() * sizeof(double),
                        hipMemcpyDeviceToHost));
    std::vector<int32_t> orig(t->n_L[L]);
    FE_HIPCHK(hipMemcpy(orig.data(), t->d_orig[L], orig.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < t->n_L[L]; ++k) order.emplace_back(orig[k], L, (int)k);
  }
  std::sort(order.begin(), order.end());
  int32_t toff = 0, moff = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    const int L = std::get<1>(order[i]);
    const int k = std::get<2>(order[i]);
    const double* r = &rec[L][(size_t)k * fe_host::STRIDE[L]];
    pp_cx[i] = r[0];
    pp_cy[i] = r[1];
    pp_cz[i] = r[2];
    pp_exp[i] = r[3];
    pp_K[i] = r[4];
    if (pp_kbound) pp_kbound[i] = r[5];
    if (pp_reff) pp_reff[i] = 1.6 * sqrt((L + 2.0) / r[3]);  // fastesp/esp.py:254
    const int nt = fe_host::NTERM[L], nm = fe_host::NMONO[L];
    pp_nt[i] = nt;
    pp_off[i] = toff;
    pp_nmono[i] = nm;
    pp_monooff[i] = moff;
    for (int j = 0; j < nt; ++j) {
      const int* key = fe_host::KEYS[fe_host::KEY_OFF[L] + j];
      tm_aidx[toff + j] = key[4];
      tm_m[toff + j] = key[3];
      tm_g[toff + j] = r[8 + j];
    }
    for (int j = 0; j < nm; ++j) {
      const int* m = fe_host::MONOS[fe_host::MONO_OFF[L] + j];
      mo_parent[moff + j] = m[3];
      mo_axis[moff + j] = m[4];
    }
    toff += nt;
    moff += nm;
  }
  return FE_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// V = [sum_A Z_A/|C-R_A|] - [sum of partial rows], rows added in a fixed order.
// 64 points per 256-thread block, each point's rows split over the 4 waves
__global__ __launch_bounds__(256) void k_esp_reduce(const double* __restrict__ partial,
                                                     const int n_rows, const int ld,
                                                     const double* __restrict__ pts,
                                                     const int n_pts,
                                                     const double* __restrict__ atoms,
                                                     const int n_atoms, const int parts,
                                                     const int* __restrict__ perm,
                                                     double* __restrict__ V) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  double sum[1];
  fe_sum_rows_split<1>(partial + i, i < n_pts, n_rows, ld, 0, sum);
  if (threadIdx.x >= 64 || i >= n_pts) return;
  const int o = perm ? perm[i] : i;
  const double e = sum[0];
  const double nuc = (parts & FE_PART_NUCLEAR)
                         ? fe_nuclear(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], atoms, n_atoms)
                         : 0.0;
  V[o] = (parts & FE_PART_ELECTRON) ? nuc - e : nuc;
}

void fe_launch_esp_reduce(const double* partial, int n_rows, int ld, const double* pts, int n,
                          const fe_table* t, int parts, const int32_t* perm, double* V,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_esp_reduce, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, partial,
                     n_rows, ld, pts, n, t->d_atoms, t->n_atoms, parts, perm, V);
}

template <int M>
__global__ __launch_bounds__(256) void k_boys_eval(const double* __restrict__ T, const int64_t n,
                                                   const double* __restrict__ tab,
                                                   double* __restrict__ F) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double f[M + 1];
  fe_boys<M>(T[i], f, tab);
#pragma unroll
  for (int m = 0; m <= M; ++m) F[i * (M + 1) + m] = f[m];
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
FwdPlan fe_plan(const fe_table* t, int64_t n, const int* ppl, int target_wg, int split) {
  FwdPlan pl{};
  pl.ld = (int)((n + 1023) / 1024 * 1024);
  pl.split = split;
  for (int L = 0; L <= FE_LMAX; ++L) {
    const int64_t nL = fe_plan_n(t, L);
    pl.ppl[L] = (ppl[L] == 4 || ppl[L] == 2) ? ppl[L] : 1;
    const int per_block = FE_WG * pl.ppl[L];
    pl.n_pb[L] = (int)((n + per_block - 1) / per_block);
    if (nL == 0) {
      pl.chunks[L] = 0;
      continue;
    }
    int64_t c = std::max<int64_t>(1, target_wg / std::max(pl.n_pb[L], 1));
    c = std::min<int64_t>(c, nL);
    c = std::min<int64_t>(c, 65535);
    const int64_t ppc = (nL + c - 1) / c;
    c = (nL + ppc - 1) / ppc;
    pl.ppc[L] = (int)ppc;
    pl.chunks[L] = (int)c;
  }
  // rows: chunk c of a class writes row c of its set if no class launched
  // before it on the same set had a chunk c, else adds into it
  int rows = 0;
  const int n_sets = split ? 2 : 1;
  for (int g = 0; g < n_sets; ++g) {
    const int* order = split == 2 ? FE_ORDER_SMALL[g] : split ? FE_ORDER_SPLIT[g] : FE_ORDER_ONE;
    int set_rows = 0;
    for (int k = 0; order[k] >= 0; ++k) {
      const int L = order[k];
      if (L > FE_LMAX) continue;
      pl.row0[L] = rows;
      pl.nadd[L] = set_rows;
      set_rows = std::max(set_rows, pl.chunks[L]);
    }
    rows += set_rows;
  }
  pl.rows = rows;
  return pl;
}

int fe_stage_points(fe_ctx* c, const double* pts, int64_t n, int ptr_kind, hipStream_t st,
                    const double** d_pts) {
  if (ptr_kind == FE_PTR_DEVICE) {
    *d_pts = pts;
    return FE_OK;
  }
  int rc = fe_ctx_reserve(c, FE_WS_PTS, (size_t)n * 3 * sizeof(double));
  if (rc) return rc;
  FE_HIPCHK(hipMemcpyAsync(c->ws[FE_WS_PTS], pts, (size_t)n * 3 * sizeof(double),
                           hipMemcpyHostToDevice, st));
  *d_pts = (const double*)c->ws[FE_WS_PTS];
  return FE_OK;
}

extern "C" int fe_esp_forward(fe_table* t, const double* pts, int64_t n, int precision, int parts,
                              int ptr_kind, void* stream, double* V) {
  if (!t || (!pts && n > 0) || (!V && n > 0))
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward: null argument");
  if (n < 0 || n > (int64_t)INT32_MAX / 4)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward: point count %lld out of range", (long long)n);
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward: bad ptr_kind %d", ptr_kind);
  if (precision != FE_PREC_FP64 && precision != FE_PREC_MIXED)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward: bad precision %d", precision);
  if (parts < 1 || parts > 3) return fe_set_error(FE_ERR_ARG, "fe_esp_forward: bad parts %d", parts);
  if (n == 0) return FE_OK;
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const double* d_pts = nullptr;
  int rc = fe_stage_points(c, pts, n, ptr_kind, st, &d_pts);
  if (rc) return rc;
  double* d_V = V;
  if (ptr_kind == FE_PTR_HOST) {
    rc = fe_ctx_reserve(c, FE_WS_OUT, (size_t)n * sizeof(double));
    if (rc) return rc;
    d_V = (double*)c->ws[FE_WS_OUT];
  }
  const int32_t* perm = nullptr;
  if (c->sort_points && n >= 2 * FE_WG && (parts & FE_PART_ELECTRON)) {
    const int ps = fe_prof_begin(c, FE_PH_SORT, st);
    rc = fe_sort_points(c, t, d_pts, (int)n, st, &d_pts, &perm);
    if (rc) return rc;
    fe_prof_end(c, ps, st);
  }
  FwdPlan pl = fe_plan(t, n, c->ppl_L, c->target_wg, c->split_eval);
  double* partial = nullptr;
  const int groups = (parts & FE_PART_ELECTRON) ? fe_row_groups((int)n, pl.rows) : 0;
  if ((parts & FE_PART_ELECTRON) && pl.rows > 0) {
    rc = fe_ctx_reserve(c, FE_WS_PARTIAL, (size_t)(pl.rows + groups) * pl.ld * sizeof(double));
    if (rc) return rc;
    partial = (double*)c->ws[FE_WS_PARTIAL];
    rc = fe_launch_eval(t, pl, d_pts, (int)n, partial, 0, false, precision == FE_PREC_MIXED, st);
    if (rc) return rc;
  }
  int nrows = (parts & FE_PART_ELECTRON) ? pl.rows : 0;
  const int pr = fe_prof_begin(c, FE_PH_REDUCE, st);
  if (groups) {
    double* staged = partial + (size_t)pl.rows * pl.ld;
    const int rpg = (pl.rows + groups - 1) / groups;
    hipLaunchKernelGGL(k_rows_stage<1>, dim3((unsigned)((n + 63) / 64), (unsigned)groups),
                       dim3(256), 0, st, partial, pl.rows, pl.ld, (size_t)0, (int)n, rpg, staged,
                       (size_t)0);
    partial = staged;
    nrows = groups;
  }
  hipLaunchKernelGGL(k_esp_reduce, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, partial,
                     nrows, pl.ld, d_pts, (int)n, t->d_atoms, t->n_atoms, parts, perm, d_V);
  fe_prof_end(c, pr, st);
  FE_HIPCHK(hipGetLastError());
  if (ptr_kind == FE_PTR_HOST) {
    FE_HIPCHK(hipMemcpyAsync(V, d_V, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
  }
  return FE_OK;
}

extern "C" int fe_boys_eval(fe_ctx* c, const double* T, int64_t n, int M, int ptr_kind,
                            void* stream, double* F) {
  if (!c || (n > 0 && (!T || !F))) return fe_set_error(FE_ERR_ARG, "fe_boys_eval: null argument");
  if (M < 0 || M > FE_BOYS_MMAX) return fe_set_error(FE_ERR_ARG, "fe_boys_eval: M=%d out of range", M);
  if (n < 0) return fe_set_error(FE_ERR_ARG, "fe_boys_eval: negative n");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_boys_eval: bad ptr_kind");
  if (n == 0) return FE_OK;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  const double* dT = T;
  double* dF = F;
  if (ptr_kind == FE_PTR_HOST) {
    int rc = fe_ctx_reserve(c, FE_WS_PTS, (size_t)n * sizeof(double));
    if (rc) return rc;
    rc = fe_ctx_reserve(c, FE_WS_OUT, (size_t)n * (M + 1) * sizeof(double));
    if (rc) return rc;
    FE_HIPCHK(hipMemcpyAsync(c->ws[FE_WS_PTS], T, (size_t)n * sizeof(double),
                             hipMemcpyHostToDevice, st));
    dT = (const double*)c->ws[FE_WS_PTS];
    dF = (double*)c->ws[FE_WS_OUT];
  }
  const dim3 grid((unsigned)((n + 255) / 256));
  const double* tab = c->d_boys + (size_t)M * FE_BOYS_NR * FE_BOYS_COLS;
  switch (M) {
#define FE_BOYS_CASE(m) \
  case m: hipLaunchKernelGGL(k_boys_eval<m>, grid, dim3(256), 0, st, dT, n, tab, dF); break;
    FE_BOYS_CASE(0) FE_BOYS_CASE(1) FE_BOYS_CASE(2) FE_BOYS_CASE(3) FE_BOYS_CASE(4)
    FE_BOYS_CASE(5) FE_BOYS_CASE(6) FE_BOYS_CASE(7) FE_BOYS_CASE(8)
#undef FE_BOYS_CASE
  }
  FE_HIPCHK(hipGetLastError());
  if (ptr_kind == FE_PTR_HOST) {
    FE_HIPCHK(hipMemcpyAsync(F, dF, (size_t)n * (M + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
  }
  return FE_OK;
}
