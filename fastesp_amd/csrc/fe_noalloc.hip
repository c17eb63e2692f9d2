// fe_esp_forward_noalloc: the ESP forward on device buffers with no
// allocation.  Its scratch is a fixed file-scope device array: the first
// FE_NOALLOC_SORT_BYTES hold the point sort (keys, indices, sorted points,
// radix-sort temporaries), the rest the partial rows of the pair chunks.  Each
// device holds its own copy for the life of the process; host code reaches it
// only through the address that fe_noalloc_init (called by fe_init, with the
// context's device current) stores in the context.
#include <hip/hip_runtime.h>

#include "fe_internal.h"

#define FE_NOALLOC_SCRATCH_BYTES ((unsigned long long)FE_NOALLOC_SCRATCH_MIB << 20)
#define FE_NOALLOC_SORT_BYTES (8ull << 20)

static __device__ __align__(256) char scratch[FE_NOALLOC_SCRATCH_BYTES];

int fe_noalloc_init(fe_ctx* c) {
  void* sp = nullptr;
  const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !sp) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_HIP, "fe_init: address of the scratch array: %s",
                        hipGetErrorString(e));
  }
  c->noalloc_scratch = (char*)sp;
  return FE_OK;
}

// fe_esp_forward on device buffers with no allocation: the point sort and the
// partial rows live in the file-scope scratch array (through the address taken
// at fe_init), and the pair-chunk count is lowered until the rows fit.
extern "C" int fe_esp_forward_noalloc(fe_table* t, const double* pts, int64_t n, int precision,
                                      int parts, void* stream, double* V) {
  if (!t || (!pts && n > 0) || (!V && n > 0))
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_noalloc: null argument");
  if (n < 0 || n > FE_NOALLOC_MAX_POINTS)
    return fe_set_error(FE_ERR_ARG,
                        "fe_esp_forward_noalloc: %lld points; this entry point takes at most %d",
                        (long long)n, FE_NOALLOC_MAX_POINTS);
  if (precision != FE_PREC_FP64 && precision != FE_PREC_MIXED)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_noalloc: bad precision %d", precision);
  if (parts < 1 || parts > 3)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_noalloc: bad parts %d", parts);
  if (n == 0) return FE_OK;
  fe_ctx* c = t->ctx;
  FE_CTX_LOCK(c);
  if (!c->noalloc_scratch)
    return fe_set_error(FE_ERR_STATE, "fe_esp_forward_noalloc: context has no scratch");
  if (precision == FE_PREC_MIXED)
    for (int L = 0; L <= FE_LMAX; ++L)
      if (t->cap_L[L] && !t->d_rec32[L])
        return fe_set_error(FE_ERR_STATE,
                            "fe_esp_forward_noalloc: the FP32 coefficients of this table are not "
                            "built yet (run fe_esp_forward once with FE_PREC_MIXED)");
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // phase events would be created on first use: no profiling in this call
  struct ProfOff {
    fe_ctx* c;
    bool was;
    ~ProfOff() { c->prof_on = was; }
  } prof_off{c, c->prof_on};
  c->prof_on = false;
  char* sort_buf = c->noalloc_scratch;
  double* partial = (double*)(c->noalloc_scratch + FE_NOALLOC_SORT_BYTES);
  const size_t part_cap = (FE_NOALLOC_SCRATCH_BYTES - FE_NOALLOC_SORT_BYTES) / sizeof(double);
  const double* d_pts = pts;
  const int32_t* perm = nullptr;
  int rc = FE_OK;
  if (c->sort_points && n >= 2 * FE_WG && (parts & FE_PART_ELECTRON)) {
    rc = fe_sort_points(c, t, pts, (int)n, st, &d_pts, &perm, sort_buf, FE_NOALLOC_SORT_BYTES);
    if (rc) return rc;
  }
  int target = c->target_wg;
  FwdPlan pl = fe_plan(t, n, c->ppl_L, target);
  while ((size_t)pl.rows * pl.ld > part_cap && target > 1) {
    target /= 2;
    pl = fe_plan(t, n, c->ppl_L, target);
  }
  if ((size_t)pl.rows * pl.ld > part_cap)
    return fe_set_error(FE_ERR_ARG, "fe_esp_forward_noalloc: partial rows do not fit the scratch");
  const int nrows = (parts & FE_PART_ELECTRON) ? pl.rows : 0;
  if (nrows > 0) {
    rc = fe_launch_eval(t, pl, d_pts, (int)n, partial, 0, false, precision == FE_PREC_MIXED, st);
    if (rc) return rc;
  }
  fe_launch_esp_reduce(nrows ? partial : nullptr, nrows, pl.ld, d_pts, (int)n, t, parts, perm, V,
                       st);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

