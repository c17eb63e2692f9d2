// Internal types shared by the translation units of libfastesp_hip.
#pragma once

#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/fastesp_hip.h"

#ifndef FE_LMAX
#define FE_LMAX 6
#endif

#ifndef FE_WG
#define FE_WG 256  // threads per workgroup of the evaluation kernels (4 waves)
#endif

// Minimum resident waves per SIMD requested from the compiler for the
// evaluation (k_eval, k_batch_eval; 2 points per lane) and dL/dG (k_dG,
// k_dG_batch) kernels of class L: the register budget becomes 512 / waves
// VGPRs.  1 leaves the allocation to the compiler.  Overridable at build time (-DFE_EVAL_MINW_L=...,
// -DFE_DG_MINW_L=..., comma lists of 7 values).
#ifndef FE_EVAL_MINW_L
#define FE_EVAL_MINW_L 1, 1, 1, 1, 3, 1, 1
#endif
#ifndef FE_DG_MINW_L
#define FE_DG_MINW_L 1, 1, 1, 1, 1, 2, 1
#endif
// Highest class whose value + field evaluation folds the weight into F before
// the contraction (fe_contract_grad_acc) instead of weighting its four results.
#ifndef FE_GRAD_ACC_LMAX
#define FE_GRAD_ACC_LMAX 2
#endif
// Lowest class whose far-field wavefronts (value, and value + field) use the
// record's far-field block (|u| H(u), u = R/|R|^2) instead of the asymptotic
// Boys vector and the keyed contraction; FE_LMAX + 1 turns it off.
#ifndef FE_KELVIN_LMIN
#define FE_KELVIN_LMIN 0
#endif
// Highest class whose far and near evaluation paths hand the factors of their
// accumulation to one shared multiply-add (fewest operations); the classes
// above it form the products inside each path (fewest live registers).
#ifndef FE_TAIL_LMAX
#define FE_TAIL_LMAX 4
#endif
// Classes (bit L) whose evaluation loop reads the next record's header one
// trip ahead.
#ifndef FE_HDR_PREFETCH
#define FE_HDR_PREFETCH 0x6f
#endif
// The same for one point per lane.
#ifndef FE_EVAL_MINW1_L
#define FE_EVAL_MINW1_L 1, 1, 1, 1, 1, 3, 3
#endif
__host__ __device__ constexpr int fe_eval_minw(int L, int ppl) {
  constexpr int t[7] = {FE_EVAL_MINW_L};
  constexpr int t1[7] = {FE_EVAL_MINW1_L};
  return ppl == 2 ? t[L] : ppl == 1 ? t1[L] : 1;
}
__host__ __device__ constexpr int fe_dg_minw(int L) {
  constexpr int t[7] = {FE_DG_MINW_L};
  return t[L];
}

int fe_set_error(int code, const char* fmt, ...);

// Device primitives (fe_prim.hip): deterministic, no memsets, no host
// synchronisation.  Stable LSD radix sort of (key, value) pairs on the low
// key_bits bits (or on the bits [lo_bit, hi_bit), the others left as they
// are), in place, with fe_radix_sort_scratch(n) bytes of scratch; exclusive
// prefix sums (out may equal in) with fe_scan_scratch(n) bytes.
size_t fe_radix_sort_scratch(int64_t n);
int fe_radix_sort(uint32_t* keys, int32_t* vals, int n, int key_bits, void* scratch,
                  hipStream_t st);
int fe_radix_sort_bits(uint32_t* keys, int32_t* vals, int n, int lo_bit, int hi_bit,
                       void* scratch, hipStream_t st);
size_t fe_scan_scratch(int64_t n);
int fe_scan_excl_u64(const unsigned long long* in, unsigned long long* out, int n,
                     void* scratch, hipStream_t st);

#define FE_HIPCHK(x)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      (void)hipGetLastError();                                                               \
      return fe_set_error(FE_ERR_HIP, "%s:%d: %s: %s", __FILE__, __LINE__, #x,               \
                          hipGetErrorString(e_));                                            \
    }                                                                                        \
  } while (0)

// workspace slots of a context (grown on demand, freed at fe_finalize)
enum {
  FE_WS_PTS = 0,
  FE_WS_OUT,
  FE_WS_PARTIAL,
  FE_WS_AUX0,
  FE_WS_AUX1,
  FE_WS_AUX2,
  FE_WS_AUX3,
  FE_WS_SORT,
  FE_WS_BUILD,  // scratch of fe_table_build (kept for the next build)
  FE_WS_SLOTS
};

// phases timed by fe_profile (HIP events on the launch stream)
enum {
  FE_PH_SORT = 0,
  FE_PH_REDUCE = 1,
  FE_PH_PACK = 2,
  FE_PH_DP = 3,
  FE_PH_DENS_PHI = 4,  // basis values at the points (density)
  FE_PH_DENS_MM = 5,   // density quadratic form on the matrix cores
  FE_PH_DENS_RED = 6,  // density block reduction
  FE_PH_EVAL0 = 8,    // + L: evaluation launch of class L
  FE_PH_DG0 = 16,     // + L: dL/dG launch of class L
  FE_PH_DGRED0 = 24,  // + L: dL/dG chunk reduction of class L
  FE_PH_N = 32
};

struct FeProfEv {
  int phase;
  hipEvent_t a, b;
};

struct fe_ctx {
  // serialises the entry points that use the workspace slots below (calls on
  // one context from several host threads queue up here; device-pointer calls
  // on different streams must still be ordered by the caller, because the
  // slots are shared by every stream)
  std::recursive_mutex mu;
  int device = 0;
  int n_cu = 256;
  double* d_boys = nullptr;  // (FE_BOYS_MMAX+1) x NR x COLS
  void* ws[FE_WS_SLOTS] = {};
  size_t ws_bytes[FE_WS_SLOTS] = {};
  // points per lane, per class, of the value kernels and of the value + field
  // kernels (the shapes the autotuner picks at config 1)
  int ppl_L[FE_LMAX + 1] = {2, 2, 2, 2, 2, 2, 4};
  int ppl_grad_L[FE_LMAX + 1] = {2, 2, 2, 2, 2, 1, 1};
  int target_wg = 4096; // workgroups per class launch
  bool sort_points = true;  // Morton-order the points before evaluation
  bool pair_order = true;   // order each class by exponent bucket, then Morton code of P
  double mixed_tau = 1e-4;  // FP32 routing threshold of the mixed path (a.u.)
  // classes that take the FP32 contraction under FE_PREC_MIXED (the others run
  // the FP64 kernel: for small L the conversion and routing cost more than the
  // FP32 contraction saves)
  bool mixed_L[FE_LMAX + 1] = {false, false, false, true, true, true, true};
  bool prof_on = false;     // record phase events (fe_profile)
  std::vector<FeProfEv> prof;
  size_t prof_n = 0;
  // device address of this device's copy of the file-scope scratch array of
  // fe_esp_forward_noalloc (from hipGetSymbolAddress at fe_init)
  char* noalloc_scratch = nullptr;
  // the last basis fe_table_build set up (host layout + device copy), reused
  // when the next build passes the same basis arrays
  std::shared_ptr<void> basis_cache;
  // Two more streams for launches that are independent of each other (the
  // dL/dG classes, the heavy candidates of a table build): fe_fork makes
  // them wait for the caller's stream, fe_join makes the caller's stream
  // wait for them.  Default priority, non-blocking; created by fe_init.
  // Under graph capture the fork and join become graph edges.
  hipStream_t side = nullptr, side2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
  int dg_streams = 2;       // streams of the dL/dG classes, 1-3 (tuning key 6)
  // the evaluation classes on two streams, each with its own partial rows
  // (tuning key 7, FwdPlan::split): the value kernels in two balanced halves
  // (forward 4.31 -> 4.16 ms at config 1), the value + field kernels with
  // classes 5 and 6 beside the first large ones
  int split_eval = 1;
  int split_eval_grad = 2;
};

// fork / join of the context's second stream with the caller's stream st
int fe_fork(fe_ctx* c, hipStream_t st, hipStream_t* side);
int fe_join(fe_ctx* c, hipStream_t st);
// the same for n streams in all (1-3): s[0] = st, s[1..n-1] the context's
int fe_fork_n(fe_ctx* c, hipStream_t st, int n, hipStream_t* s);
int fe_join_n(fe_ctx* c, hipStream_t st, int n);
// The order and stream of the class launches of a backward over n streams:
// largest estimated work first, each to the stream with the least so far.
// work[L] is the class's estimated cost; q[k] receives the stream index of
// the k-th class of order.
void fe_lpt(const double* work, int n, int* order, int* q);

#define FE_CTX_LOCK(c) std::lock_guard<std::recursive_mutex> fe_ctx_lock_((c)->mu)

// start / end of a timed phase on stream st (no-ops unless profiling)
int fe_prof_begin(fe_ctx* c, int phase, hipStream_t st);
void fe_prof_end(fe_ctx* c, int slot, hipStream_t st);

// Table statistics on the device (FE_ST_* ints, then the sum rule as a double)
enum {
  FE_ST_KEPT = 0,                    // + L: kept pairs of class L (what the kernels read)
  FE_ST_VALID = FE_LMAX + 1,         // + L: candidates of the reference's pair list
  FE_ST_KVALID = 2 * (FE_LMAX + 1),  // + L: primitive pairs with K >= 1e-30 (rebuild capacity)
  FE_ST_N = 3 * (FE_LMAX + 1),
  FE_ST_SUMRULE_BYTE = 128           // byte offset of the sum rule (a double)
};

// shell-l pairs (LI >= LJ, LI <= 3) of the table build: c = LI (LI + 1) / 2 + LJ
#define FE_NCOMBO 10

struct fe_table {
  fe_ctx* ctx = nullptr;
  int n_pairs = 0;
  int n_terms = 0;
  int n_atoms = 0;
  int64_t n_candidates = 0;
  // One allocation holds the statistics, the per-class records, their flat
  // term offsets and reference-order indices, and (built tables) the slot map.
  void* d_mem = nullptr;
  size_t mem_bytes = 0;
  int* d_stats = nullptr;             // FE_ST_N ints (kernels read the kept counts here)
  double* d_sum_rule = nullptr;
  double* d_rec[FE_LMAX + 1] = {};    // per-L records (FP64), cap_L[L] of them
  float* d_rec32[FE_LMAX + 1] = {};   // per-L g in FP32 (mixed precision), lazily built
  int64_t n_L[FE_LMAX + 1] = {};      // kept pairs per class (host copy, see counts_stale)
  int64_t cap_L[FE_LMAX + 1] = {};    // records allocated per class
  int32_t* d_orig[FE_LMAX + 1] = {};  // reference-order (flat-table) index of every record
  int* d_flatoff[FE_LMAX + 1] = {};   // term offset of every record in the flat layout
  // A rebuild (fe_table_rebuild) leaves the new counts on the device only:
  // host-side plans then size launches by cap_L and the kernels read the true
  // counts; host queries read them back first (fe_table_sync_counts).
  bool counts_stale = false;
  bool capacity_layout = false;       // records for every K-valid candidate
  hipStream_t count_stream = nullptr;
  double* d_atoms = nullptr;          // (x, y, z, Z) per atom
  std::vector<double> h_atoms;        // host copy of d_atoms
  size_t dev_bytes = 0;
  double sum_rule = 0.0;              // sum over candidates of K (pi/p)^1.5 etab_0
  // --- present when built on the device from a basis (fe_table_build) ---
  bool has_basis = false;
  int n_func = 0;
  int n_shells = 0;
  int n_sp = 0;
  // device copy of the shells, primitives, components and shell pairs, shared
  // with the context's basis cache and every table built from the same basis
  std::shared_ptr<void> basis_mem;
  struct {
    const int* atom; const int* l; const int* fn0; const int* nfn; const int* prim0;
    const int* nprim; const double* c; const double* pexp; const double* pcoef;
    const int* fn_lx; const int* fn_ly; const int* fn_lz;
    const int* sp_i; const int* sp_j; const int* sp_slot; const int* sp_gidx;
    const int* sp_by_class;  // shell-pair indices grouped by class L = l_i + l_j
    const uint32_t* tasks;   // dL/dP tasks (sp << 8 | mi << 4 | mj) grouped by class
    const int* fn_cperm;     // function fn0 + c: offset in its shell of canonical component c
    const int* cl_sp;        // build candidates grouped by shell-l pair: shell pair,
    const int* cl_s;         //   primitive pair within it (i-major)
  } b{};
  int cmb_off[FE_NCOMBO + 1] = {};  // shell-l pair (l_i >= l_j) c: cl_*[cmb_off[c], cmb_off[c+1])
  int sp_cls_off[FE_LMAX + 2] = {};  // class L: sp_by_class[sp_cls_off[L], sp_cls_off[L+1])
  int task_cls_off[FE_LMAX + 2] = {};  // class L: tasks[task_cls_off[L], task_cls_off[L+1])
  int64_t n_cand_L[FE_LMAX + 1] = {}; // candidate slots (every primitive pair) per class
  int64_t kvalid_L[FE_LMAX + 1] = {}; // K-valid slots per class
  int* d_slot2kept[FE_LMAX + 1] = {}; // candidate slot -> compact record (-1 dropped)
  // pair-order key frame: bounding box of the shell centres, and per class the
  // log range of the candidate exponents p
  double key_lo[3] = {0, 0, 0};
  double key_scale = 0.0;
  double key_lpmin[FE_LMAX + 1] = {}, key_lpinv[FE_LMAX + 1] = {};
};

// pairs of class L that launches are sized for: the kept count when the host
// knows it, else the allocated capacity (the kernels read the kept count)
inline int64_t fe_plan_n(const fe_table* t, int L) {
  return t->counts_stale ? t->cap_L[L] : t->n_L[L];
}
// the pairs of class L a launch shape is chosen for (work split, stream
// balance): after an in-place rebuild the last kept count the host saw (the
// kept set seldom moves much), else fe_plan_n; buffers and grids still cover
// fe_plan_n
inline int64_t fe_plan_est(const fe_table* t, int L) {
  return t->counts_stale && t->n_L[L] > 0 ? std::min(t->n_L[L], t->cap_L[L]) : fe_plan_n(t, L);
}
// one device allocation for a table: statistics, records (cap[L] per class),
// flat offsets and reference-order indices, plus n_slots ints of slot map
int fe_table_layout(fe_table* t, const int64_t* cap, int64_t n_slots);
// read the device statistics back after a rebuild (synchronises its stream)
int fe_table_sync_counts(fe_table* t);
// the FP32 coefficients again after the records changed (no-op without them)
int fe_refresh_g32(fe_table* t, hipStream_t st);
// workspace bytes of a build of table t on n_points device points, and the
// rebuild on device inputs into a caller-owned workspace (the table must be in
// its capacity layout: a first fe_table_rebuild has run); `fork` false keeps
// every launch on st (no use of the context's second stream)
size_t fe_build_ws_bytes(const fe_table* t, int n_points);
int fe_table_rebuild_ws(fe_table* t, const double* dens, int n_points, const double* pts,
                        double tol, int rule, hipStream_t st, char* ws, size_t ws_bytes,
                        bool fork = true);


// Partial sums: pair chunk c of every class adds into row c.  The classes run
// in the order 0..FE_LMAX on one stream; class L writes rows [nadd[L],
// chunks[L]) and adds into rows [0, nadd[L]), which earlier classes wrote, so
// each row holds its classes' sums in a fixed order and the reduction reads
// max(chunks) rows rather than one set per class.
// Launch order of the class launches: one stream (0..6), or two streams
// (split), the odd classes on the caller's stream and the even ones on the
// context's second stream, each set with rows of its own.  -1 ends a list.
static constexpr int FE_ORDER_ONE[FE_LMAX + 2] = {0, 1, 2, 3, 4, 5, 6, -1};
static constexpr int FE_ORDER_SPLIT[2][FE_LMAX + 2] = {{1, 3, 5, -1}, {2, 0, 4, 6, -1}};
// the two small, latency-bound classes beside classes 0 and 1 (the row sets
// of FwdPlan::split 2; fe_launch_eval joins the streams before class 2)
static constexpr int FE_ORDER_SMALL[2][FE_LMAX + 2] = {{0, 1, 2, 3, 4, -1}, {6, 5, -1}};

struct FwdPlan {
  int ld = 0;     // leading dimension of a partial row (>= n, multiple of 1024)
  int rows = 0;   // partial rows (all row sets)
  int split = 0;  // 0 one stream; two streams, two row sets: 1 FE_ORDER_SPLIT, 2 FE_ORDER_SMALL
  int n_pb[FE_LMAX + 1] = {};   // point blocks of the class launch
  int ppl[FE_LMAX + 1] = {};    // points per lane of the class launch (1, 2, 4)
  int chunks[FE_LMAX + 1] = {};
  int ppc[FE_LMAX + 1] = {};
  int nadd[FE_LMAX + 1] = {};   // rows of its set already holding sums when class L runs
  int row0[FE_LMAX + 1] = {};   // first row of class L's row set
};

int fe_ctx_reserve(fe_ctx* c, int slot, size_t bytes);
int fe_table_set_atoms(fe_table* t, int n_atoms, const double* Z, const double* xyz);
// Record order within a class: 16 buckets of log p over the class's range, then
// the 30-bit Morton code of P in the bounding box of the records.  Pairs whose
// near-field radii are alike then share a wavefront in the pair-per-lane
// adjoint, whose far-field test is per wavefront.  rec4 holds (Px, Py, Pz, p)
// at a stride of `stride` doubles; returns a stable permutation.
std::vector<int> fe_pair_order(const double* rec4, size_t stride, int64_t n);
FwdPlan fe_plan(const fe_table* t, int64_t n, const int* ppl, int target_wg, int split = 0);

// dispatch a templated class launcher on the plan's points-per-lane of class L
#define FE_DISPATCH_PPL(pl, L, FN, ...)                  \
  do {                                                   \
    if ((pl).chunks[L] == 0) break;                      \
    switch ((pl).ppl[L]) {                               \
      case 4: FN<L, 4>(__VA_ARGS__); break;              \
      case 2: FN<L, 2>(__VA_ARGS__); break;              \
      default: FN<L, 1>(__VA_ARGS__); break;             \
    }                                                    \
  } while (0)

#define FE_FOR_EACH_CLASS(pl, FN, ...)                   \
  do {                                                   \
    FE_DISPATCH_PPL(pl, 0, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 1, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 2, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 3, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 4, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 5, FN, __VA_ARGS__);             \
    FE_DISPATCH_PPL(pl, 6, FN, __VA_ARGS__);             \
  } while (0)
int fe_stage_points(fe_ctx* c, const double* pts, int64_t n, int ptr_kind, hipStream_t st,
                    const double** d_pts);
// Morton-sorted copy of the points and the permutation (sorted -> caller
// index), in the context's FE_WS_SORT slot, or in `buf` (buf_bytes long) when
// it is given: then nothing is allocated and FE_ERR_ARG comes back if it is too
// small.
int fe_sort_points(fe_ctx* c, const fe_table* t, const double* d_pts, int n, hipStream_t st,
                   const double** d_sorted, const int32_t** d_perm, void* buf = nullptr,
                   size_t buf_bytes = 0);
// one work item of the batched evaluation: pairs [p0, p1) of class L of
// conformer `conf` against the sorted points [base, min(base + PPL*FE_WG, end));
// the partial sums go to partial[out + ...] (planes `plane` apart for the field)
struct FeBatchItem {
  int64_t out;
  int conf, p0, p1, base, end, pad;
};

struct FeBatchEvalArgs {
  const FeBatchItem* items;
  int n_items;
  int* counter;                 // work-queue head, zero before the launch
  const double* const* recs;    // per conformer: class-L records
  const int* const* kept;       // per conformer: its class-L kept count (device)
  const float* const* g32s;     // per conformer: FP32 coefficients (mixed)
  const double* pts;            // sorted points of all conformers
  const double* boys;           // Boys grids of the context (all orders)
  double* partial;
  size_t plane;
  float tau;
  int grid;
};

int fe_ensure_g32(fe_table* t, hipStream_t st);
// V = [nuclear] - [sum of the first n_rows partial rows] (k_esp_reduce)
void fe_launch_esp_reduce(const double* partial, int n_rows, int ld, const double* pts, int n,
                          const fe_table* t, int parts, const int32_t* perm, double* V,
                          hipStream_t st);
// takes this device's address of the no-allocation scratch array (fe_init)
int fe_noalloc_init(fe_ctx* c);
int fe_launch_eval(fe_table* t, const FwdPlan& pl, const double* pts, int n, double* partial,
                   size_t plane, bool grad, bool mixed, hipStream_t st);
// launches the batched evaluation of class L (a != nullptr), or returns its
// resident workgroups per CU (a == nullptr)
int fe_batch_eval(int L, int ppl, bool grad, bool mixed, const FeBatchEvalArgs* a,
                  hipStream_t st);
