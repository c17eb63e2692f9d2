// Batched evaluation of many conformers (tables of uneven size, uneven point
// counts) with the forward, the loss gradient and the backward of one step
// recorded as a single hipGraph.
//
// Layout.  The points of all conformers are concatenated (conformer c owns
// [pt_off[c], pt_off[c+1])) and Morton-ordered within each conformer; perm maps
// a sorted slot to its caller-order index.  Tables stay where fe_table_build /
// fe_table_from_pairs put them: per class L a device array holds the record
// pointer of every conformer.
//
// Work.  Per class L one persistent launch covers every conformer: a work
// item is (conformer, pair range of class L, block of FE_WG*PPL sorted
// points), and workgroups pull items from an atomic queue head until it is
// drained.  Items are sized from the class total so that each workgroup takes
// several, and ordered largest first.  Each item writes its own partial row
// slot; the per-point reduction adds a conformer's rows in a fixed order, so a
// result does not depend on which workgroup took which item.  One kernel per
// class rather than one for all classes: a fused kernel would run every class
// at the register allocation (and occupancy) of the largest one.
//
// Backward.  dL/dG items are (conformer, block of 256 pairs, point chunk), one
// lane per pair as in the single-table adjoint; a per-class reduction over the
// chunks in fixed order; then dL/dP for every shell pair of every conformer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/fastesp_hip.h"
#include "fe_internal.h"
#include "fe_adjoint.h"
#include "fe_reduce.h"

#define FE_RB_LANES 4  // streams of the in-graph table rebuilds

namespace {

struct FeDgItem {
  int64_t out;  // offset of the item's chunk slot in the dG partial buffer
  int conf, pb; // conformer, block of 256 pairs
  int q0, q1;   // sorted point range
  int n_pairs;  // class-L pairs of the conformer (row stride of the slot)
  int pad;
};

struct DgSeg {  // one (conformer, class) of the dG reduction, non-empty only
  int64_t part, dg, e0;
  int n_pairs, n_chunks;  // n_pairs: the planned count (the slot's row stride)
  int conf, pad;
};

struct ConfDev {
  int64_t part;  // offset of the conformer's first partial row
  const double* atoms;
  int rows, ld, pt0, n, n_atoms, pad;
};

enum { CNT_EVAL = 0, CNT_DG = FE_LMAX + 1, CNT_N = 2 * (FE_LMAX + 1) };

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// carve aligned sub-buffers out of one allocation
struct Carver {
  size_t off = 0;
  template <class T>
  size_t take(size_t count) {
    const size_t o = off;
    off += up256(std::max<size_t>(count, 1) * sizeof(T));
    return o;
  }
};

}  // namespace

struct fe_batch {
  fe_ctx* ctx = nullptr;
  int n_conf = 0;
  std::vector<fe_table*> tabs;
  std::vector<int64_t> pt_off;  // n_conf + 1
  int64_t n_tot = 0;
  bool all_basis = true;
  // fixed buffers
  void* mem = nullptr;
  double *pts_user = nullptr, *pts = nullptr, *V = nullptr, *field = nullptr, *w = nullptr,
         *target = nullptr, *gV = nullptr, *loss = nullptr, *pw = nullptr;
  int *perm = nullptr, *pt_conf = nullptr, *counters = nullptr;
  int64_t* d_pt_off = nullptr;
  const double** recs[FE_LMAX + 1] = {};
  const int** kept[FE_LMAX + 1] = {};  // per class: each conformer's device kept count
  const float** g32s[FE_LMAX + 1] = {};
  // Tables rebuilt inside the step (fe_batch_set_density): the densities of
  // every conformer and a build workspace of the batch's own, so a recorded
  // step depends on no context workspace.  Plans are then cut from the
  // tables' capacities and the kernels read the kept counts on the device.
  bool rb = false;
  double rb_tol = 0.0;
  int rb_rule = 0;
  void* rbmem = nullptr;
  double* dens = nullptr;
  char* bws = nullptr;      // FE_RB_LANES build workspaces of bws_bytes each
  size_t bws_bytes = 0;
  // the rebuilds run on FE_RB_LANES streams of the batch's own (table i on
  // lane i mod FE_RB_LANES), forked from and joined into the step's stream
  hipStream_t rb_st[FE_RB_LANES] = {};
  hipEvent_t rb_fork = nullptr, rb_join[FE_RB_LANES] = {};
  std::vector<int64_t> dens_off;  // n_conf + 1
  // forward plan
  bool fplan = false, f_grad = false, f_mixed = false;
  void* fmem = nullptr;
  FeBatchItem* items[FE_LMAX + 1] = {};
  int n_items[FE_LMAX + 1] = {}, grid[FE_LMAX + 1] = {};
  ConfDev* conf = nullptr;
  double* partial = nullptr;
  size_t plane = 0;
  bool has_field = false;
  // backward plan
  bool bplan = false;
  void* bmem = nullptr;
  FeDgItem* dgitems[FE_LMAX + 1] = {};
  int n_dgitems[FE_LMAX + 1] = {}, dggrid[FE_LMAX + 1] = {};
  DgSeg* segs[FE_LMAX + 1] = {};
  int n_segs[FE_LMAX + 1] = {};
  int64_t n_elem[FE_LMAX + 1] = {};
  double *dg_part = nullptr, *dG = nullptr, *dP = nullptr;
  DPArgs* dpargs = nullptr;
  int64_t* sp_off = nullptr;  // per class L: (n_conf + 1) offsets of its shell pairs
  int64_t n_sp[FE_LMAX + 1] = {}, n_sp_tot = 0;
  int64_t* rec_off = nullptr;   // per class L: (n_conf + 1) offsets of the planned records
  int64_t* task_off = nullptr;  // per class L: (n_conf + 1) offsets of the dL/dP tasks
  int64_t n_rec[FE_LMAX + 1] = {}, n_task[FE_LMAX + 1] = {};
  std::vector<int64_t> dp_off;  // n_conf + 1
  // tuning
  int ppl[FE_LMAX + 1] = {2, 2, 2, 2, 2, 1, 1};
  bool mixed_L[FE_LMAX + 1] = {false, false, false, true, true, true, true};
  int k_eval = 8, k_dg = 8, mask = 0x7f;
  // graph
  hipStream_t cs = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_prec = 0, g_flags = 0;
  bool capturing = false;
};

// phase events only for eager launches (never recorded into the graph)
static int prof_begin(fe_batch* b, int phase, hipStream_t st) {
  return b->capturing ? -1 : fe_prof_begin(b->ctx, phase, st);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_zero_i32(int* __restrict__ x, const int n) {
  if ((int)threadIdx.x < n) x[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_fill_batch(double* __restrict__ x, const int64_t n,
                                                    const double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}

// perm[g] (global caller-order index of sorted slot g) and pt_conf[g] of one
// conformer; local = nullptr means identity order
__global__ __launch_bounds__(256) void k_perm_fill(const int* __restrict__ local, const int n,
                                                   const int pt0, const int conf,
                                                   int* __restrict__ perm,
                                                   int* __restrict__ pt_conf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  perm[pt0 + i] = pt0 + (local ? local[i] : i);
  pt_conf[pt0 + i] = conf;
}

__global__ __launch_bounds__(256) void k_gather_sorted(const double* __restrict__ user,
                                                       const int* __restrict__ perm, const int n,
                                                       double* __restrict__ pts) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  const int o = perm[g];
  pts[3 * g] = user[3 * (size_t)o];
  pts[3 * g + 1] = user[3 * (size_t)o + 1];
  pts[3 * g + 2] = user[3 * (size_t)o + 2];
}

template <bool GRAD>
__global__ __launch_bounds__(64) void k_batch_reduce(const double* __restrict__ partial,
                                                     const size_t plane,
                                                     const ConfDev* __restrict__ conf,
                                                     const int* __restrict__ pt_conf,
                                                     const double* __restrict__ pts,
                                                     const int* __restrict__ perm, const int n,
                                                     double* __restrict__ V,
                                                     double* __restrict__ G) {
  const int g = blockIdx.x * 64 + threadIdx.x;
  if (g >= n) return;
  const ConfDev cf = conf[pt_conf[g]];
  const double* p = partial + cf.part + (g - cf.pt0);
  const double x = pts[3 * g], y = pts[3 * g + 1], z = pts[3 * g + 2];
  const int o = perm[g];
  if constexpr (GRAD) {
    double e, gx, gy, gz, nuc, nx, ny, nz;
    fe_sum_rows4(p, cf.rows, cf.ld, plane, e, gx, gy, gz);
    fe_nuclear_grad(x, y, z, cf.atoms, cf.n_atoms, nuc, nx, ny, nz);
    V[o] = nuc - e;
    G[3 * (size_t)o] = nx - gx;
    G[3 * (size_t)o + 1] = ny - gy;
    G[3 * (size_t)o + 2] = nz - gz;
  } else {
    V[o] = fe_nuclear(x, y, z, cf.atoms, cf.n_atoms) - fe_sum_rows(p, cf.rows, cf.ld);
  }
}

// loss_c = sum_i w_i (V_i - T_i)^2 over the caller-order points of conformer c,
// gV_i = 2 w_i (V_i - T_i); one workgroup per conformer, fixed-order tree
__global__ __launch_bounds__(256) void k_loss(const int64_t* __restrict__ pt_off,
                                              const double* __restrict__ V,
                                              const double* __restrict__ w,
                                              const double* __restrict__ T,
                                              double* __restrict__ gV, double* __restrict__ loss) {
  __shared__ double s[256];
  const int c = blockIdx.x;
  const int64_t i0 = pt_off[c], i1 = pt_off[c + 1];
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const double d = V[i] - T[i];
    gV[i] = 2.0 * w[i] * d;
    acc = fma(w[i] * d, d, acc);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[c] = s[0];
}

__global__ __launch_bounds__(256) void k_pack_pw_perm(const double* __restrict__ pts,
                                                      const double* __restrict__ gV,
                                                      const int* __restrict__ perm, const int n,
                                                      double* __restrict__ pw) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  pw[4 * (size_t)g] = pts[3 * (size_t)g];
  pw[4 * (size_t)g + 1] = pts[3 * (size_t)g + 1];
  pw[4 * (size_t)g + 2] = pts[3 * (size_t)g + 2];
  pw[4 * (size_t)g + 3] = gV[perm[g]];
}

// persistent dL/dG: items from a queue, one lane per pair
template <int L>
__global__ __launch_bounds__(256, fe_dg_minw(L)) void k_dG_batch(const FeDgItem* __restrict__ items,
                                                  const int n_items, int* __restrict__ counter,
                                                  const double* const* __restrict__ recs,
                                                  const int* const* __restrict__ kept,
                                                  const double* __restrict__ pw,
                                                  const double* __restrict__ boys_tab,
                                                  double* __restrict__ part) {
  constexpr int NTAB = fe_boys_rows(L) * FE_BOYS_COLS;
  __shared__ double s_tab[NTAB];
  __shared__ int s_it;
  for (int i = threadIdx.x; i < NTAB; i += 256) s_tab[i] = boys_tab[i];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_it = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_it);
    if (it >= n_items) break;
    const FeDgItem d = items[it];
    // pairs past the conformer's kept count (plans cut from capacities) are
    // skipped; the test is block-uniform
    const int np = min(*kept[d.conf], d.n_pairs);
    if (d.pb * 256 >= np) continue;
    fe_dg_block<L>(recs[d.conf], np, d.pb * 256 + (int)threadIdx.x, pw, d.q0, d.q1, s_tab,
                   part + d.out, d.n_pairs);
  }
}

// dG[pair * NT + k] = sum over the point chunks, fixed order
template <int L>
__global__ __launch_bounds__(256) void k_dG_reduce_batch(const DgSeg* __restrict__ segs,
                                                         const int n_segs, const int64_t n_elem,
                                                         const int* const* __restrict__ kept,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ dG) {
  constexpr int NT = FeL<L>::NT;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n_elem) return;
  int lo = 0, hi = n_segs - 1;  // last segment with e0 <= e
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].e0 <= e) lo = mid;
    else hi = mid - 1;
  }
  const DgSeg sg = segs[lo];
  const int64_t el = e - sg.e0;
  const int k = (int)(el / sg.n_pairs), pair = (int)(el - (int64_t)k * sg.n_pairs);
  if (pair >= *kept[sg.conf]) return;  // no record there (dL/dP reads kept pairs only)
  double s = 0.0;
  for (int c = 0; c < sg.n_chunks; ++c)
    s += part[sg.part + ((int64_t)c * NT + k) * sg.n_pairs + pair];
  dG[sg.dg + (int64_t)pair * NT + k] = s;
}

// the last conformer c with off[c] <= x (off: n_conf + 1 ascending offsets)
__device__ __forceinline__ int fe_conf_of(const int64_t* __restrict__ off, const int n_conf,
                                          const int64_t x) {
  int lo = 0, hi = n_conf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// d etab of every conformer's kept class-L records (one lane each; the
// records of conformer c are [rec_off[c], rec_off[c+1]) of the launch)
template <int L>
__global__ __launch_bounds__(256) void k_det_batch(const DPArgs* __restrict__ args,
                                                   const int64_t* __restrict__ rec_off,
                                                   const int n_conf, const int64_t n_rec,
                                                   const double* const* __restrict__ recs,
                                                   const int* const* __restrict__ kept) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n_rec) return;
  const int cf = fe_conf_of(rec_off, n_conf, g);
  const int r = (int)(g - rec_off[cf]);
  if (r >= *kept[cf]) return;
  const DPArgs& a = args[cf];
  fe_det_from_dg<L>(a.dG[L] + (size_t)r * FeL<L>::NT, recs[cf][(size_t)r * FeL<L>::STRIDE + 3],
                    (double*)a.det[L] + (size_t)r * FeL<L>::NM);
}

// the class-L dL/dP tasks of every conformer (their d etab is ready)
template <int L>
__global__ __launch_bounds__(256) void k_dp_tasks_batch(const DPArgs* __restrict__ args,
                                                        const int64_t* __restrict__ task_off,
                                                        const int n_conf, const int64_t n_task) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= n_task) return;
  const int cf = fe_conf_of(task_off, n_conf, g);
  const DPArgs& a = args[cf];
  fe_dp_task<L>(a, a.tasks[a.task_cls_off[L] + (int)(g - task_off[cf])]);
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static void drop_graph(fe_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void drop_fplan(fe_batch* b) {
  drop_graph(b);
  if (b->fmem) (void)hipFree(b->fmem);
  b->fmem = nullptr;
  b->fplan = false;
}

static void drop_bplan(fe_batch* b) {
  drop_graph(b);
  if (b->bmem) (void)hipFree(b->bmem);
  b->bmem = nullptr;
  b->bplan = false;
}

static int batch_alloc(void** p, size_t bytes, const char* what) {
  if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fe_set_error(FE_ERR_ALLOC, "batch: allocation of %zu bytes for %s failed", bytes, what);
  }
  return FE_OK;
}

// Morton order within each conformer, then the sorted copy of the points
static int sort_points(fe_batch* b, hipStream_t st) {
  fe_ctx* c = b->ctx;
  std::vector<int> order(b->n_conf);
  std::iota(order.begin(), order.end(), 0);
  // largest first: the sort workspace then never grows while earlier work is queued
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b->pt_off[x + 1] - b->pt_off[x] > b->pt_off[y + 1] - b->pt_off[y];
  });
  for (int ci : order) {
    const int64_t pt0 = b->pt_off[ci], n = b->pt_off[ci + 1] - pt0;
    if (n == 0) continue;
    const int32_t* local = nullptr;
    if (c->sort_points && n >= 2 * FE_WG) {
      const double* sorted = nullptr;
      int rc = fe_sort_points(c, b->tabs[ci], b->pts_user + 3 * pt0, (int)n, st, &sorted, &local);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_perm_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, local,
                       (int)n, (int)pt0, ci, b->perm, b->pt_conf);
  }
  if (b->n_tot)
    hipLaunchKernelGGL(k_gather_sorted, dim3((unsigned)((b->n_tot + 255) / 256)), dim3(256), 0, st,
                       b->pts_user, b->perm, (int)b->n_tot, b->pts);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

// per class, each conformer's record pointer and kept-count pointer (the
// first rebuild of a table moves both)
static int upload_table_ptrs(fe_batch* b) {
  std::vector<const void*> rp(b->n_conf), kp(b->n_conf);
  for (int L = 0; L <= FE_LMAX; ++L) {
    for (int i = 0; i < b->n_conf; ++i) {
      rp[i] = b->tabs[i]->d_rec[L];
      kp[i] = b->tabs[i]->d_stats + FE_ST_KEPT + L;
    }
    if (hipMemcpy(b->recs[L], rp.data(), b->n_conf * sizeof(void*), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(b->kept[L], kp.data(), b->n_conf * sizeof(void*), hipMemcpyHostToDevice) !=
            hipSuccess)
      return fe_set_error(FE_ERR_HIP, "batch: upload of table pointers failed");
  }
  return FE_OK;
}

// pairs of class L a plan is cut for: the kept count, or the capacity when
// the batch rebuilds its tables
static int64_t plan_pairs(const fe_batch* b, const fe_table* t, int L) {
  return b->rb ? t->cap_L[L] : t->n_L[L];
}

extern "C" int fe_batch_create(fe_ctx* ctx, int n_conf, fe_table* const* tables,
                               const double* pts, const int64_t* n_pts, int ptr_kind,
                               void* stream, fe_batch** out) {
  if (!out) return fe_set_error(FE_ERR_ARG, "fe_batch_create: null output pointer");
  *out = nullptr;
  if (!ctx || !tables || !n_pts || n_conf < 1)
    return fe_set_error(FE_ERR_ARG, "fe_batch_create: null argument or n_conf < 1");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_create: bad ptr_kind");
  int64_t tot = 0;
  for (int i = 0; i < n_conf; ++i) {
    if (!tables[i]) return fe_set_error(FE_ERR_ARG, "fe_batch_create: table %d is null", i);
    if (tables[i]->ctx != ctx)
      return fe_set_error(FE_ERR_ARG, "fe_batch_create: table %d belongs to another context", i);
    if (n_pts[i] < 0) return fe_set_error(FE_ERR_ARG, "fe_batch_create: n_pts[%d] < 0", i);
    // work items are cut from the host-side counts
    if (int rc = fe_table_sync_counts(tables[i])) return rc;
    tot += n_pts[i];
  }
  if (tot > (int64_t)INT32_MAX / 8)
    return fe_set_error(FE_ERR_ARG, "fe_batch_create: %lld points in total is too many",
                        (long long)tot);
  if (tot > 0 && !pts) return fe_set_error(FE_ERR_ARG, "fe_batch_create: null points");
  FE_CTX_LOCK(ctx);
  FE_HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  fe_batch* b = new fe_batch();
  b->ctx = ctx;
  b->n_conf = n_conf;
  b->tabs.assign(tables, tables + n_conf);
  b->pt_off.resize(n_conf + 1, 0);
  for (int i = 0; i < n_conf; ++i) {
    b->pt_off[i + 1] = b->pt_off[i] + n_pts[i];
    b->all_basis = b->all_basis && tables[i]->has_basis;
  }
  b->n_tot = tot;
  auto fail = [&](int rc) {
    if (b->mem) (void)hipFree(b->mem);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    delete b;
    return rc;
  };
  const size_t N = (size_t)tot;
  Carver cv;
  const size_t o_pu = cv.take<double>(3 * N), o_ps = cv.take<double>(3 * N);
  const size_t o_V = cv.take<double>(N), o_F = cv.take<double>(3 * N);
  const size_t o_w = cv.take<double>(N), o_T = cv.take<double>(N), o_gV = cv.take<double>(N);
  const size_t o_loss = cv.take<double>(n_conf), o_pw = cv.take<double>(4 * N);
  const size_t o_perm = cv.take<int>(N), o_pc = cv.take<int>(N), o_cnt = cv.take<int>(CNT_N);
  const size_t o_off = cv.take<int64_t>(n_conf + 1);
  size_t o_rec[FE_LMAX + 1], o_g32[FE_LMAX + 1], o_kept[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_rec[L] = cv.take<const double*>(n_conf);
    o_g32[L] = cv.take<const float*>(n_conf);
    o_kept[L] = cv.take<const int*>(n_conf);
  }
  if (int rc = batch_alloc(&b->mem, cv.off, "points and outputs")) return fail(rc);
  char* m = (char*)b->mem;
  b->pts_user = (double*)(m + o_pu);
  b->pts = (double*)(m + o_ps);
  b->V = (double*)(m + o_V);
  b->field = (double*)(m + o_F);
  b->w = (double*)(m + o_w);
  b->target = (double*)(m + o_T);
  b->gV = (double*)(m + o_gV);
  b->loss = (double*)(m + o_loss);
  b->pw = (double*)(m + o_pw);
  b->perm = (int*)(m + o_perm);
  b->pt_conf = (int*)(m + o_pc);
  b->counters = (int*)(m + o_cnt);
  b->d_pt_off = (int64_t*)(m + o_off);
  for (int L = 0; L <= FE_LMAX; ++L) {
    b->recs[L] = (const double**)(m + o_rec[L]);
    b->g32s[L] = (const float**)(m + o_g32[L]);
    b->kept[L] = (const int**)(m + o_kept[L]);
  }
  if (int rc = upload_table_ptrs(b)) return fail(rc);
  if (hipMemcpy(b->d_pt_off, b->pt_off.data(), (n_conf + 1) * sizeof(int64_t),
                hipMemcpyHostToDevice) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_batch_create: upload of point offsets failed"));
  if (hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_batch_create: stream creation failed"));
  if (N) {
    if (hipMemcpyAsync(b->pts_user, pts, 3 * N * sizeof(double),
                       ptr_kind == FE_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
      return fail(fe_set_error(FE_ERR_HIP, "fe_batch_create: point upload failed"));
    const dim3 g((unsigned)((N + 255) / 256));
    hipLaunchKernelGGL(k_fill_batch, g, dim3(256), 0, st, b->w, (int64_t)N, 1.0);
    hipLaunchKernelGGL(k_fill_batch, g, dim3(256), 0, st, b->target, (int64_t)N, 0.0);
    hipLaunchKernelGGL(k_fill_batch, g, dim3(256), 0, st, b->V, (int64_t)N, 0.0);
    hipLaunchKernelGGL(k_fill_batch, g, dim3(256), 0, st, b->gV, (int64_t)N, 0.0);
  }
  if (int rc = sort_points(b, st)) return fail(rc);
  if (hipStreamSynchronize(st) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_batch_create: synchronisation failed"));
  *out = b;
  return FE_OK;
}

extern "C" int fe_batch_free(fe_batch* b) {
  if (!b) return FE_OK;
  (void)hipSetDevice(b->ctx->device);
  drop_fplan(b);
  drop_bplan(b);
  if (b->rbmem) (void)hipFree(b->rbmem);
  for (int k = 0; k < FE_RB_LANES; ++k) {
    if (b->rb_st[k]) (void)hipStreamDestroy(b->rb_st[k]);
    if (b->rb_join[k]) (void)hipEventDestroy(b->rb_join[k]);
  }
  if (b->rb_fork) (void)hipEventDestroy(b->rb_fork);
  if (b->mem) (void)hipFree(b->mem);
  if (b->cs) (void)hipStreamDestroy(b->cs);
  delete b;
  return FE_OK;
}

// forward plan: work items of every class, partial row slots of every conformer
static int plan_fwd(fe_batch* b, bool grad, bool mixed, hipStream_t st) {
  if (b->fplan && b->f_grad == grad && b->f_mixed == mixed) return FE_OK;
  drop_fplan(b);
  fe_ctx* c = b->ctx;
  const int nc = b->n_conf;
  if (mixed) {
    for (int i = 0; i < nc; ++i)
      if (int rc = fe_ensure_g32(b->tabs[i], st)) return rc;
    std::vector<const float*> gp(nc);
    for (int L = 0; L <= FE_LMAX; ++L) {
      for (int i = 0; i < nc; ++i) gp[i] = b->tabs[i]->d_rec32[L];
      FE_HIPCHK(hipMemcpy(b->g32s[L], gp.data(), nc * sizeof(void*), hipMemcpyHostToDevice));
    }
  }
  int resident[FE_LMAX + 1];
  double total[FE_LMAX + 1] = {};
  for (int L = 0; L <= FE_LMAX; ++L) {
    const int occ = fe_batch_eval(L, b->ppl[L], grad, mixed && b->mixed_L[L], nullptr, nullptr);
    if (occ < 1) return fe_set_error(FE_ERR_HIP, "batch: class %d kernel cannot be resident", L);
    resident[L] = occ * c->n_cu;
    const int blk = FE_WG * b->ppl[L];
    for (int i = 0; i < nc; ++i) {
      const int64_t n = b->pt_off[i + 1] - b->pt_off[i];
      total[L] += (double)((n + blk - 1) / blk) * blk * (double)plan_pairs(b, b->tabs[i], L);
    }
  }
  std::vector<ConfDev> conf(nc);
  std::vector<FeBatchItem> items[FE_LMAX + 1];
  size_t plane = 0;
  for (int i = 0; i < nc; ++i) {
    const fe_table* t = b->tabs[i];
    const int pt0 = (int)b->pt_off[i], n = (int)(b->pt_off[i + 1] - b->pt_off[i]);
    const int ld = (n + 63) / 64 * 64;
    int rows = 0;
    for (int L = 0; L <= FE_LMAX; ++L) {
      const int64_t nL = plan_pairs(b, t, L);
      if (nL == 0 || n == 0) continue;
      const int blk = FE_WG * b->ppl[L];
      const double cap = std::max(1.0, total[L] / ((double)resident[L] * b->k_eval));
      int64_t ch = (int64_t)std::ceil((double)blk * (double)nL / cap);
      ch = std::max<int64_t>(1, std::min<int64_t>(ch, nL));
      const int64_t ppc = (nL + ch - 1) / ch;
      ch = (nL + ppc - 1) / ppc;
      const int n_pb = (n + blk - 1) / blk;
      // chunk-major: workgroups that run together read the same pair records
      // (scalar loads served by the shared scalar cache)
      for (int64_t k = 0; k < ch; ++k)
        for (int q = 0; q < n_pb; ++q) {
          FeBatchItem it{};
          it.out = (int64_t)plane + (int64_t)(rows + k) * ld + (int64_t)q * blk;
          it.conf = i;
          it.p0 = (int)(k * ppc);
          it.p1 = (int)std::min<int64_t>(nL, (k + 1) * ppc);
          it.base = pt0 + q * blk;
          it.end = pt0 + n;
          items[L].push_back(it);
        }
      rows += (int)ch;
    }
    conf[i].part = (int64_t)plane;
    conf[i].atoms = t->d_atoms;
    conf[i].rows = rows;
    conf[i].ld = ld;
    conf[i].pt0 = pt0;
    conf[i].n = n;
    conf[i].n_atoms = t->n_atoms;
    plane += (size_t)rows * ld;
  }
  // host-side bounds check of every slot the kernels will write or read
  for (int L = 0; L <= FE_LMAX; ++L)
    for (const FeBatchItem& it : items[L]) {
      const int blk = FE_WG * b->ppl[L];
      const int span = std::min(blk, it.end - it.base);
      if (it.p0 < 0 || it.p1 > plan_pairs(b, b->tabs[it.conf], L) || it.p0 >= it.p1 || span <= 0 ||
          it.end > b->n_tot || it.out < 0 || (size_t)(it.out + span) > plane)
        return fe_set_error(FE_ERR_STATE, "batch: internal plan error (class %d)", L);
    }
  for (int L = 0; L <= FE_LMAX; ++L) {
    auto cost = [](const FeBatchItem& x) { return (int64_t)(x.p1 - x.p0) * (x.end - x.base); };
    std::stable_sort(items[L].begin(), items[L].end(),
                     [&](const FeBatchItem& x, const FeBatchItem& y) { return cost(x) > cost(y); });
  }
  Carver cv;
  size_t o_items[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) o_items[L] = cv.take<FeBatchItem>(items[L].size());
  const size_t o_conf = cv.take<ConfDev>(nc);
  const size_t o_part = cv.take<double>(std::max<size_t>(plane, 1) * (grad ? 4 : 1));
  FE_HIPCHK(hipSetDevice(c->device));
  if (int rc = batch_alloc(&b->fmem, cv.off, "forward partial sums")) return rc;
  char* m = (char*)b->fmem;
  for (int L = 0; L <= FE_LMAX; ++L) {
    b->items[L] = (FeBatchItem*)(m + o_items[L]);
    b->n_items[L] = (int)items[L].size();
    b->grid[L] = std::min<int>(resident[L], std::max<int>(b->n_items[L], 1));
    if (!items[L].empty())
      FE_HIPCHK(hipMemcpy(b->items[L], items[L].data(), items[L].size() * sizeof(FeBatchItem),
                          hipMemcpyHostToDevice));
  }
  b->conf = (ConfDev*)(m + o_conf);
  FE_HIPCHK(hipMemcpy(b->conf, conf.data(), nc * sizeof(ConfDev), hipMemcpyHostToDevice));
  b->partial = (double*)(m + o_part);
  b->plane = std::max<size_t>(plane, 1);
  b->f_grad = grad;
  b->f_mixed = mixed;
  b->fplan = true;
  return FE_OK;
}

template <int L>
static int dg_occupancy() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_dG_batch<L>, 256, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return nb;
}

static int plan_bwd(fe_batch* b) {
  if (b->bplan) return FE_OK;
  if (!b->all_basis)
    return fe_set_error(FE_ERR_STATE,
                        "batch backward: dL/dP needs every table built from a basis (fe_table_build)");
  drop_bplan(b);
  fe_ctx* c = b->ctx;
  const int nc = b->n_conf;
  const int occ[FE_LMAX + 1] = {dg_occupancy<0>(), dg_occupancy<1>(), dg_occupancy<2>(),
                                dg_occupancy<3>(), dg_occupancy<4>(), dg_occupancy<5>(),
                                dg_occupancy<6>()};
  double total[FE_LMAX + 1] = {};
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (occ[L] < 1) return fe_set_error(FE_ERR_HIP, "batch: dG kernel %d cannot be resident", L);
    for (int i = 0; i < nc; ++i)
      total[L] += (double)((plan_pairs(b, b->tabs[i], L) + 255) / 256) * 256.0 *
                  (double)(b->pt_off[i + 1] - b->pt_off[i]);
  }
  std::vector<FeDgItem> items[FE_LMAX + 1];
  std::vector<DgSeg> segs[FE_LMAX + 1];
  int64_t e_run[FE_LMAX + 1] = {};
  size_t part_total = 0, dg_total = 0;
  std::vector<DPArgs> args(nc);
  std::vector<int64_t> sp_off((size_t)(FE_LMAX + 1) * (nc + 1), 0);
  // per class: (n_conf + 1) offsets of the conformers' planned records and
  // of their dL/dP tasks; each (conformer, class)'s d etab goes to its
  // dL/dG partial slots once the class's reduction has read them
  std::vector<int64_t> rec_off((size_t)(FE_LMAX + 1) * (nc + 1), 0);
  std::vector<int64_t> task_off((size_t)(FE_LMAX + 1) * (nc + 1), 0);
  std::vector<std::array<size_t, FE_LMAX + 1>> det_off(nc);
  std::vector<size_t> dgbase(nc);
  std::vector<std::array<size_t, FE_LMAX + 1>> dg_off(nc);
  b->dp_off.assign(nc + 1, 0);
  for (int i = 0; i < nc; ++i) {
    const fe_table* t = b->tabs[i];
    const int pt0 = (int)b->pt_off[i], n = (int)(b->pt_off[i + 1] - b->pt_off[i]);
    dgbase[i] = dg_total;
    for (int L = 0; L <= FE_LMAX; ++L) {
      const int64_t nL = plan_pairs(b, t, L);
      const int NT = fe_host::NTERM[L];
      dg_off[i][L] = dg_total - dgbase[i];
      if (nL == 0) continue;
      int64_t ch = 0, ppc = 0;
      if (n > 0) {
        const double cap = std::max(1.0, total[L] / ((double)occ[L] * c->n_cu * b->k_dg));
        ch = (int64_t)std::ceil(256.0 * n / cap);
        ch = std::min<int64_t>(ch, std::max(1, n / 64));
        ch = std::min<int64_t>(ch, std::max<int64_t>(1, (16ll << 20) / (nL * NT)));
        ch = std::max<int64_t>(ch, 1);
        ppc = (n + ch - 1) / ch;
        ch = (n + ppc - 1) / ppc;
      }
      det_off[i][L] = part_total;
      DgSeg sg{};
      sg.part = (int64_t)part_total;
      sg.dg = (int64_t)dg_total;
      sg.e0 = e_run[L];
      sg.n_pairs = (int)nL;
      sg.n_chunks = (int)ch;
      sg.conf = i;
      segs[L].push_back(sg);
      const int nb = (int)((nL + 255) / 256);
      // point-chunk-major: workgroups that run together read the same points
      for (int64_t k = 0; k < ch; ++k)
        for (int pb = 0; pb < nb; ++pb) {
          FeDgItem d{};
          d.out = (int64_t)part_total + k * NT * nL;
          d.conf = i;
          d.pb = pb;
          d.q0 = pt0 + (int)(k * ppc);
          d.q1 = pt0 + (int)std::min<int64_t>(n, (k + 1) * ppc);
          d.n_pairs = (int)nL;
          items[L].push_back(d);
        }
      // the slots also hold the class's d etab later (NM <= NT per record),
      // also for a conformer without points (no chunks)
      part_total += std::max((size_t)ch * NT * nL, (size_t)nL * fe_host::NMONO[L]);
      dg_total += (size_t)nL * NT;
      e_run[L] += nL * NT;
    }
    for (int L = 0; L <= FE_LMAX; ++L) {
      const size_t o = (size_t)L * (nc + 1) + i;
      sp_off[o + 1] = sp_off[o] + (t->sp_cls_off[L + 1] - t->sp_cls_off[L]);
      rec_off[o + 1] = rec_off[o] + plan_pairs(b, t, L);
      task_off[o + 1] = task_off[o] + (t->task_cls_off[L + 1] - t->task_cls_off[L]);
    }
    b->dp_off[i + 1] = b->dp_off[i] + (int64_t)t->n_func * t->n_func;
  }
  for (int L = 0; L <= FE_LMAX; ++L) {
    const int NT = fe_host::NTERM[L];
    for (const FeDgItem& d : items[L])
      if (d.q0 > d.q1 || d.q1 > b->n_tot || d.out < 0 ||
          (size_t)(d.out + (int64_t)NT * d.n_pairs) > part_total ||
          d.pb * 256 >= d.n_pairs)
        return fe_set_error(FE_ERR_STATE, "batch: internal dG plan error (class %d)", L);
    auto cost = [](const FeDgItem& x) {
      return (int64_t)std::min(256, x.n_pairs - x.pb * 256) * (x.q1 - x.q0);
    };
    std::stable_sort(items[L].begin(), items[L].end(),
                     [&](const FeDgItem& x, const FeDgItem& y) { return cost(x) > cost(y); });
  }
  Carver cv;
  size_t o_it[FE_LMAX + 1], o_sg[FE_LMAX + 1];
  for (int L = 0; L <= FE_LMAX; ++L) {
    o_it[L] = cv.take<FeDgItem>(items[L].size());
    o_sg[L] = cv.take<DgSeg>(segs[L].size());
  }
  const size_t o_part = cv.take<double>(part_total), o_dG = cv.take<double>(dg_total);
  const size_t o_dP = cv.take<double>((size_t)b->dp_off[nc]);
  const size_t o_args = cv.take<DPArgs>(nc), o_spo = cv.take<int64_t>(sp_off.size());
  const size_t o_reco = cv.take<int64_t>(rec_off.size()), o_tko = cv.take<int64_t>(task_off.size());
  FE_HIPCHK(hipSetDevice(c->device));
  if (int rc = batch_alloc(&b->bmem, cv.off, "backward buffers")) return rc;
  char* m = (char*)b->bmem;
  b->dg_part = (double*)(m + o_part);
  b->dG = (double*)(m + o_dG);
  b->dP = (double*)(m + o_dP);
  b->dpargs = (DPArgs*)(m + o_args);
  b->sp_off = (int64_t*)(m + o_spo);
  b->rec_off = (int64_t*)(m + o_reco);
  b->task_off = (int64_t*)(m + o_tko);
  for (int L = 0; L <= FE_LMAX; ++L) {
    b->dgitems[L] = (FeDgItem*)(m + o_it[L]);
    b->n_dgitems[L] = (int)items[L].size();
    b->dggrid[L] = std::min<int>(occ[L] * c->n_cu, std::max<int>(b->n_dgitems[L], 1));
    b->segs[L] = (DgSeg*)(m + o_sg[L]);
    b->n_segs[L] = (int)segs[L].size();
    b->n_elem[L] = e_run[L];
    if (!items[L].empty())
      FE_HIPCHK(hipMemcpy(b->dgitems[L], items[L].data(), items[L].size() * sizeof(FeDgItem),
                          hipMemcpyHostToDevice));
    if (!segs[L].empty())
      FE_HIPCHK(hipMemcpy(b->segs[L], segs[L].data(), segs[L].size() * sizeof(DgSeg),
                          hipMemcpyHostToDevice));
  }
  for (int i = 0; i < nc; ++i) {
    args[i] = fe_dp_args(b->tabs[i], b->dG + dgbase[i], dg_off[i].data(), b->dP + b->dp_off[i]);
    for (int L = 0; L <= FE_LMAX; ++L) args[i].det[L] = b->dg_part + det_off[i][L];
  }
  FE_HIPCHK(hipMemcpy(b->dpargs, args.data(), nc * sizeof(DPArgs), hipMemcpyHostToDevice));
  FE_HIPCHK(hipMemcpy(b->sp_off, sp_off.data(), sp_off.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  FE_HIPCHK(hipMemcpy(b->rec_off, rec_off.data(), rec_off.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  FE_HIPCHK(hipMemcpy(b->task_off, task_off.data(), task_off.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  for (int L = 0; L <= FE_LMAX; ++L) {
    b->n_rec[L] = rec_off[(size_t)L * (nc + 1) + nc];
    b->n_task[L] = task_off[(size_t)L * (nc + 1) + nc];
  }
  b->n_sp_tot = 0;
  for (int L = 0; L <= FE_LMAX; ++L) {
    b->n_sp[L] = sp_off[(size_t)L * (nc + 1) + nc];
    b->n_sp_tot += b->n_sp[L];
  }
  b->bplan = true;
  return FE_OK;
}

// --- launch sequences (no allocation, no synchronisation: also recorded) ---
static void launch_forward(fe_batch* b, hipStream_t st) {
  fe_ctx* c = b->ctx;
  const bool grad = b->f_grad, mixed = b->f_mixed;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, b->counters + CNT_EVAL, FE_LMAX + 1);
  for (int L = 0; L <= FE_LMAX; ++L) {
    if (!b->n_items[L] || !(b->mask & (1 << L))) continue;
    FeBatchEvalArgs a{};
    a.items = b->items[L];
    a.n_items = b->n_items[L];
    a.counter = b->counters + CNT_EVAL + L;
    a.recs = b->recs[L];
    a.kept = b->kept[L];
    a.g32s = b->g32s[L];
    a.pts = b->pts;
    a.boys = c->d_boys;
    a.partial = b->partial;
    a.plane = b->plane;
    a.tau = (float)c->mixed_tau;
    a.grid = b->grid[L];
    const int ph = prof_begin(b, FE_PH_EVAL0 + L, st);
    fe_batch_eval(L, b->ppl[L], grad, mixed && b->mixed_L[L], &a, st);
    fe_prof_end(c, ph, st);
  }
  if (b->n_tot) {
    const int pr = prof_begin(b, FE_PH_REDUCE, st);
    const dim3 g((unsigned)((b->n_tot + 63) / 64));
    if (grad)
      hipLaunchKernelGGL(k_batch_reduce<true>, g, dim3(64), 0, st, b->partial, b->plane, b->conf,
                         b->pt_conf, b->pts, b->perm, (int)b->n_tot, b->V, b->field);
    else
      hipLaunchKernelGGL(k_batch_reduce<false>, g, dim3(64), 0, st, b->partial, b->plane, b->conf,
                         b->pt_conf, b->pts, b->perm, (int)b->n_tot, b->V, b->field);
    fe_prof_end(c, pr, st);
  }
  b->has_field = grad;
}

static void launch_loss(fe_batch* b, hipStream_t st) {
  hipLaunchKernelGGL(k_loss, dim3(b->n_conf), dim3(256), 0, st, b->d_pt_off, b->V, b->w, b->target,
                     b->gV, b->loss);
}

template <int L>
static void launch_dG_L(fe_batch* b, hipStream_t st) {
  if (b->n_dgitems[L] && (b->mask & (1 << L))) {
    const int pg = prof_begin(b, FE_PH_DG0 + L, st);
    hipLaunchKernelGGL(k_dG_batch<L>, dim3(b->dggrid[L]), dim3(256), 0, st, b->dgitems[L],
                       b->n_dgitems[L], b->counters + CNT_DG + L, b->recs[L], b->kept[L], b->pw,
                       b->ctx->d_boys + (size_t)L * FE_BOYS_NR * FE_BOYS_COLS, b->dg_part);
    fe_prof_end(b->ctx, pg, st);
  }
  if (b->n_elem[L]) {
    const int pr = prof_begin(b, FE_PH_DGRED0 + L, st);
    hipLaunchKernelGGL(k_dG_reduce_batch<L>, dim3((unsigned)((b->n_elem[L] + 255) / 256)),
                       dim3(256), 0, st, b->segs[L], b->n_segs[L], b->n_elem[L], b->kept[L],
                       b->dg_part, b->dG);
    fe_prof_end(b->ctx, pr, st);
  }
}

template <int L>
static void launch_dp_T(fe_batch* b, hipStream_t st) {
  const size_t o = (size_t)L * (b->n_conf + 1);
  if (b->n_rec[L])
    hipLaunchKernelGGL(k_det_batch<L>, dim3((unsigned)((b->n_rec[L] + 255) / 256)), dim3(256), 0,
                       st, b->dpargs, b->rec_off + o, b->n_conf, b->n_rec[L],
                       (const double* const*)b->recs[L], (const int* const*)b->kept[L]);
  hipLaunchKernelGGL(k_dp_tasks_batch<L>, dim3((unsigned)((b->n_task[L] + 255) / 256)),
                     dim3(256), 0, st, b->dpargs, b->task_off + o, b->n_conf, b->n_task[L]);
}

static void launch_dp_L(fe_batch* b, int L, hipStream_t st) {
  if (!b->n_task[L]) return;
  const int pd = prof_begin(b, FE_PH_DP, st);
  switch (L) {
    case 0: launch_dp_T<0>(b, st); break;
    case 1: launch_dp_T<1>(b, st); break;
    case 2: launch_dp_T<2>(b, st); break;
    case 3: launch_dp_T<3>(b, st); break;
    case 4: launch_dp_T<4>(b, st); break;
    case 5: launch_dp_T<5>(b, st); break;
    case 6: launch_dp_T<6>(b, st); break;
    default: break;
  }
  fe_prof_end(b->ctx, pd, st);
}

static int launch_backward(fe_batch* b, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, b->counters + CNT_DG, FE_LMAX + 1);
  if (b->n_tot) {
    const int pp = prof_begin(b, FE_PH_PACK, st);
    hipLaunchKernelGGL(k_pack_pw_perm, dim3((unsigned)((b->n_tot + 255) / 256)), dim3(256), 0, st,
                       b->pts, b->gV, b->perm, (int)b->n_tot, b->pw);
    fe_prof_end(b->ctx, pp, st);
  }
  // the classes on the context's dL/dG streams (as in the single-table
  // backward: largest first, each to the least loaded stream), each class's
  // dL/dP right after its dL/dG
  hipStream_t ss[3];
  const int ns = b->ctx->dg_streams;
  if (int rc = fe_fork_n(b->ctx, st, ns, ss)) return rc;
  double work[FE_LMAX + 1] = {};
  for (int i = 0; i < b->n_conf; ++i)
    for (int L = 0; L <= FE_LMAX; ++L)
      work[L] += (double)fe_plan_est(b->tabs[i], L) * (double)(b->pt_off[i + 1] - b->pt_off[i]);
  for (int L = 0; L <= FE_LMAX; ++L) work[L] *= 20.0 + 2.2 * fe_host::NTERM[L];
  int order[FE_LMAX + 1], q[FE_LMAX + 1];
  fe_lpt(work, ns, order, q);
  for (int k = 0; k <= FE_LMAX; ++k) {
    const int L = order[k];
    hipStream_t s = ss[q[k]];
    switch (L) {
      case 0: launch_dG_L<0>(b, s); break;
      case 1: launch_dG_L<1>(b, s); break;
      case 2: launch_dG_L<2>(b, s); break;
      case 3: launch_dG_L<3>(b, s); break;
      case 4: launch_dG_L<4>(b, s); break;
      case 5: launch_dG_L<5>(b, s); break;
      case 6: launch_dG_L<6>(b, s); break;
      default: break;
    }
    launch_dp_L(b, L, s);
  }
  if (int rc = fe_join_n(b->ctx, st, ns)) return rc;
  return FE_OK;
}

static int check_prec(int precision, const char* fn) {
  if (precision != FE_PREC_FP64 && precision != FE_PREC_MIXED)
    return fe_set_error(FE_ERR_ARG, "%s: bad precision %d", fn, precision);
  return FE_OK;
}

extern "C" int fe_batch_forward(fe_batch* b, int precision, int flags, void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_forward: null batch");
  if (int rc = check_prec(precision, "fe_batch_forward")) return rc;
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (int rc = plan_fwd(b, flags & FE_BATCH_FIELD, precision == FE_PREC_MIXED, st)) return rc;
  launch_forward(b, st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

extern "C" int fe_batch_loss(fe_batch* b, void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_loss: null batch");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  launch_loss(b, (hipStream_t)stream);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

extern "C" int fe_batch_backward(fe_batch* b, const double* gV, int ptr_kind, void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_backward: null batch");
  if (gV && ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_backward: bad ptr_kind");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (int rc = plan_bwd(b)) return rc;
  if (gV && b->n_tot)
    FE_HIPCHK(hipMemcpyAsync(b->gV, gV, (size_t)b->n_tot * sizeof(double),
                             ptr_kind == FE_PTR_HOST ? hipMemcpyHostToDevice
                                                     : hipMemcpyDeviceToDevice,
                             st));
  if (int rc = launch_backward(b, st)) return rc;
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

static int prepare_step(fe_batch* b, int precision, int flags, hipStream_t st) {
  if (int rc = check_prec(precision, "fe_batch_step")) return rc;
  if ((flags & FE_BATCH_REBUILD) && !b->rb)
    return fe_set_error(FE_ERR_STATE,
                        "fe_batch_step: FE_BATCH_REBUILD needs the densities (fe_batch_set_density)");
  if (int rc = plan_fwd(b, flags & FE_BATCH_FIELD, precision == FE_PREC_MIXED, st)) return rc;
  if (flags & FE_BATCH_DLDP)
    if (int rc = plan_bwd(b)) return rc;
  return FE_OK;
}

// every table built again from the batch's densities and current points (the
// screen reads the sorted copy; it does not depend on the order)
static int launch_rebuilds(fe_batch* b, hipStream_t st) {
  const int K = std::min(FE_RB_LANES, b->n_conf);
  FE_HIPCHK(hipEventRecord(b->rb_fork, st));
  for (int k = 0; k < K; ++k) FE_HIPCHK(hipStreamWaitEvent(b->rb_st[k], b->rb_fork, 0));
  for (int i = 0; i < b->n_conf; ++i) {
    const int k = i % K;
    const int64_t pt0 = b->pt_off[i], n = b->pt_off[i + 1] - pt0;
    if (int rc = fe_table_rebuild_ws(b->tabs[i], b->dens + b->dens_off[i], (int)n,
                                     b->pts + 3 * pt0, b->rb_tol, b->rb_rule, b->rb_st[k],
                                     b->bws + (size_t)k * b->bws_bytes, b->bws_bytes, false))
      return rc;
  }
  for (int k = 0; k < K; ++k) {
    FE_HIPCHK(hipEventRecord(b->rb_join[k], b->rb_st[k]));
    FE_HIPCHK(hipStreamWaitEvent(st, b->rb_join[k], 0));
  }
  // the counts are complete once st is: host queries wait on st, not on a
  // lane stream of this batch
  for (int i = 0; i < b->n_conf; ++i) b->tabs[i]->count_stream = st;
  return FE_OK;
}

static int launch_step(fe_batch* b, int flags, hipStream_t st) {
  if (flags & FE_BATCH_REBUILD)
    if (int rc = launch_rebuilds(b, st)) return rc;
  launch_forward(b, st);
  launch_loss(b, st);
  if (flags & FE_BATCH_DLDP)
    if (int rc = launch_backward(b, st)) return rc;
  return FE_OK;
}

extern "C" int fe_batch_step(fe_batch* b, int precision, int flags, void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_step: null batch");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (int rc = prepare_step(b, precision, flags, st)) return rc;
  if (int rc = launch_step(b, flags, st)) return rc;
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

extern "C" int fe_batch_capture(fe_batch* b, int precision, int flags) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_capture: null batch");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  if (int rc = prepare_step(b, precision, flags, b->cs)) return rc;
  FE_HIPCHK(hipStreamSynchronize(b->cs));
  drop_graph(b);
  FE_HIPCHK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
  b->capturing = true;
  // no phase events in the graph (the table rebuilds would record them)
  const bool prof = b->ctx->prof_on;
  b->ctx->prof_on = false;
  const int rc_step = launch_step(b, flags, b->cs);
  b->ctx->prof_on = prof;
  b->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cs, &g);
  if (rc_step != FE_OK || e != hipSuccess || !g) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    if (rc_step != FE_OK) return rc_step;
    return fe_set_error(FE_ERR_HIP, "fe_batch_capture: capture failed: %s", hipGetErrorString(e));
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    drop_graph(b);
    return fe_set_error(FE_ERR_HIP, "fe_batch_capture: graph instantiation failed");
  }
  b->g_prec = precision;
  b->g_flags = flags;
  return FE_OK;
}

extern "C" int fe_batch_launch(fe_batch* b, void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_launch: null batch");
  if (!b->exec)
    return fe_set_error(FE_ERR_STATE, "fe_batch_launch: no captured step (fe_batch_capture)");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  FE_HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  b->has_field = b->f_grad;
  return FE_OK;
}

extern "C" int fe_batch_set_points(fe_batch* b, const double* pts, int ptr_kind, void* stream) {
  if (!b || (b->n_tot && !pts)) return fe_set_error(FE_ERR_ARG, "fe_batch_set_points: null argument");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_set_points: bad ptr_kind");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (b->n_tot)
    FE_HIPCHK(hipMemcpyAsync(b->pts_user, pts, (size_t)b->n_tot * 3 * sizeof(double),
                             ptr_kind == FE_PTR_HOST ? hipMemcpyHostToDevice
                                                     : hipMemcpyDeviceToDevice,
                             st));
  if (int rc = sort_points(b, st)) return rc;
  if (ptr_kind == FE_PTR_HOST) FE_HIPCHK(hipStreamSynchronize(st));
  return FE_OK;
}

extern "C" int fe_batch_set_loss(fe_batch* b, const double* w, const double* target, int ptr_kind,
                                 void* stream) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_set_loss: null batch");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_set_loss: bad ptr_kind");
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const hipMemcpyKind k = ptr_kind == FE_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const size_t bytes = (size_t)b->n_tot * sizeof(double);
  if (w && bytes) FE_HIPCHK(hipMemcpyAsync(b->w, w, bytes, k, st));
  if (target && bytes) FE_HIPCHK(hipMemcpyAsync(b->target, target, bytes, k, st));
  if (ptr_kind == FE_PTR_HOST) FE_HIPCHK(hipStreamSynchronize(st));
  return FE_OK;
}

extern "C" int fe_batch_set_density(fe_batch* b, const double* dens, double tol, int rule,
                                    int ptr_kind, void* stream) {
  if (!b || !dens) return fe_set_error(FE_ERR_ARG, "fe_batch_set_density: null argument");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_set_density: bad ptr_kind");
  if (!(tol >= 0.0) || !std::isfinite(tol))
    return fe_set_error(FE_ERR_ARG, "fe_batch_set_density: tol must be finite and >= 0");
  if (rule < 0 || rule > 2) return fe_set_error(FE_ERR_ARG, "fe_batch_set_density: bad rule %d", rule);
  if (!b->all_basis)
    return fe_set_error(FE_ERR_STATE,
                        "fe_batch_set_density: every table must be built from a basis (fe_table_build)");
  FE_CTX_LOCK(b->ctx);
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const int nc = b->n_conf;
  if (!b->rbmem) {
    b->dens_off.assign(nc + 1, 0);
    size_t ws = 0;
    for (int i = 0; i < nc; ++i) {
      const fe_table* t = b->tabs[i];
      b->dens_off[i + 1] = b->dens_off[i] + (int64_t)t->n_func * t->n_func;
      ws = std::max(ws, fe_build_ws_bytes(t, (int)(b->pt_off[i + 1] - b->pt_off[i])));
    }
    ws = up256(ws);
    const size_t db = up256((size_t)b->dens_off[nc] * sizeof(double));
    if (int rc = batch_alloc(&b->rbmem, db + FE_RB_LANES * ws, "densities and build workspaces"))
      return rc;
    b->dens = (double*)b->rbmem;
    b->bws = (char*)b->rbmem + db;
    b->bws_bytes = ws;
  }
  for (int k = 0; k < FE_RB_LANES; ++k) {
    if (!b->rb_st[k]) FE_HIPCHK(hipStreamCreateWithFlags(&b->rb_st[k], hipStreamNonBlocking));
    if (!b->rb_join[k]) FE_HIPCHK(hipEventCreateWithFlags(&b->rb_join[k], hipEventDisableTiming));
  }
  if (!b->rb_fork) FE_HIPCHK(hipEventCreateWithFlags(&b->rb_fork, hipEventDisableTiming));
  if (b->dens_off[nc])
    FE_HIPCHK(hipMemcpyAsync(b->dens, dens, (size_t)b->dens_off[nc] * sizeof(double),
                             ptr_kind == FE_PTR_HOST ? hipMemcpyHostToDevice
                                                     : hipMemcpyDeviceToDevice,
                             st));
  if (!b->rb) {
    // first time: every table's first rebuild (its capacity layout; this
    // allocates and synchronises), then the plans are cut from capacities
    for (int i = 0; i < nc; ++i) {
      fe_table* t = b->tabs[i];
      const int64_t pt0 = b->pt_off[i], n = b->pt_off[i + 1] - pt0;
      if (int rc = fe_table_rebuild(t, b->dens + b->dens_off[i], (int)n, b->pts + 3 * pt0, tol,
                                    rule, FE_PTR_DEVICE, st))
        return rc;
    }
    FE_HIPCHK(hipStreamSynchronize(st));
    if (int rc = upload_table_ptrs(b)) return rc;
    drop_fplan(b);
    drop_bplan(b);
    b->rb = true;
  } else if (tol != b->rb_tol || rule != b->rb_rule) {
    drop_graph(b);  // the recorded rebuilds carry tol and rule
  }
  b->rb_tol = tol;
  b->rb_rule = rule;
  if (ptr_kind == FE_PTR_HOST) FE_HIPCHK(hipStreamSynchronize(st));
  return FE_OK;
}

extern "C" int fe_batch_read(fe_batch* b, int what, int ptr_kind, void* stream, double* dst) {
  if (!b || !dst) return fe_set_error(FE_ERR_ARG, "fe_batch_read: null argument");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_batch_read: bad ptr_kind");
  const double* src = nullptr;
  size_t count = 0;
  switch (what) {
    case FE_BOUT_V: src = b->V; count = (size_t)b->n_tot; break;
    case FE_BOUT_FIELD:
      if (!b->has_field)
        return fe_set_error(FE_ERR_STATE, "fe_batch_read: the last forward computed no field");
      src = b->field; count = 3 * (size_t)b->n_tot; break;
    case FE_BOUT_DLDP:
      if (!b->bplan) return fe_set_error(FE_ERR_STATE, "fe_batch_read: no backward has run");
      src = b->dP; count = (size_t)b->dp_off[b->n_conf]; break;
    case FE_BOUT_LOSS: src = b->loss; count = (size_t)b->n_conf; break;
    case FE_BOUT_GV: src = b->gV; count = (size_t)b->n_tot; break;
    default: return fe_set_error(FE_ERR_ARG, "fe_batch_read: unknown output %d", what);
  }
  FE_HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  if (count)
    FE_HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double),
                             ptr_kind == FE_PTR_HOST ? hipMemcpyDeviceToHost
                                                     : hipMemcpyDeviceToDevice,
                             st));
  if (ptr_kind == FE_PTR_HOST) FE_HIPCHK(hipStreamSynchronize(st));
  return FE_OK;
}

extern "C" int fe_batch_set_tuning(fe_batch* b, int key, int64_t value) {
  if (!b) return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: null batch");
  if (key >= 20 && key <= 20 + FE_LMAX) {
    if (b->mixed_L[key - 20] != (value != 0)) drop_fplan(b);
    b->mixed_L[key - 20] = value != 0;
    return FE_OK;
  }
  if (key >= 10 && key <= 10 + FE_LMAX) {
    if (value != 1 && value != 2 && value != 4)
      return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: points per lane must be 1, 2 or 4");
    if (b->ppl[key - 10] != value) drop_fplan(b);
    b->ppl[key - 10] = (int)value;
    return FE_OK;
  }
  switch (key) {
    case FE_BT_ITEMS_PER_WG:
      if (value < 1 || value > 4096)
        return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: items per workgroup out of range");
      if (b->k_eval != value) drop_fplan(b);
      b->k_eval = (int)value;
      return FE_OK;
    case FE_BT_DG_ITEMS_PER_WG:
      if (value < 1 || value > 4096)
        return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: items per workgroup out of range");
      if (b->k_dg != value) drop_bplan(b);
      b->k_dg = (int)value;
      return FE_OK;
    case FE_BT_CLASS_MASK:
      if (value < 0 || value > 0x7f)
        return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: class mask out of range");
      if (b->mask != value) drop_graph(b);
      b->mask = (int)value;
      return FE_OK;
    default:
      return fe_set_error(FE_ERR_ARG, "fe_batch_set_tuning: unknown key %d", key);
  }
}

extern "C" int fe_batch_info(const fe_batch* b, int64_t* out, int n_out) {
  if (!b || !out) return fe_set_error(FE_ERR_ARG, "fe_batch_info: null argument");
  int64_t v[12] = {};
  v[0] = b->n_conf;
  v[1] = b->n_tot;
  for (int i = 0; i < b->n_conf; ++i) v[2] += b->tabs[i]->n_pairs;
  for (int L = 0; L <= FE_LMAX; ++L) {
    v[3] += b->n_items[L];
    v[4] += b->n_dgitems[L];
  }
  v[5] = (int64_t)b->plane * (b->f_grad ? 4 : 1) * (int64_t)sizeof(double);
  v[6] = b->bplan ? b->dp_off[b->n_conf] : 0;
  v[7] = b->n_sp_tot;
  v[8] = b->exec ? 1 : 0;
  v[9] = b->k_eval;
  v[10] = b->k_dg;
  v[11] = b->mask;
  for (int i = 0; i < n_out && i < 12; ++i) out[i] = v[i];
  return FE_OK;
}
