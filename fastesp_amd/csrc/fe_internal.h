// Internal types shared by the translation units of libfastesp_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/fastesp_hip.h"

#ifndef FE_LMAX
#define FE_LMAX 6
#endif

#define FE_WG 256  // threads per workgroup of the evaluation kernels (4 waves)

int fe_set_error(int code, const char* fmt, ...);

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
  FE_WS_SLOTS
};

struct fe_ctx {
  int device = 0;
  int n_cu = 256;
  double* d_boys = nullptr;  // (FE_BOYS_MMAX+1) x NR x COLS
  void* ws[FE_WS_SLOTS] = {};
  size_t ws_bytes[FE_WS_SLOTS] = {};
  int ppl = 1;          // points per lane
  int target_wg = 2048; // workgroups per class launch
  bool tile_screen = false;
};

struct fe_table {
  fe_ctx* ctx = nullptr;
  int n_pairs = 0;
  int n_terms = 0;
  int n_atoms = 0;
  int64_t n_candidates = 0;
  double* d_rec[FE_LMAX + 1] = {};    // per-L records (FP64)
  float* d_rec32[FE_LMAX + 1] = {};   // per-L g in FP32 (mixed precision), lazily built
  int64_t n_L[FE_LMAX + 1] = {};
  std::vector<int32_t> orig[FE_LMAX + 1];  // flat-table index of every record
  double* d_atoms = nullptr;          // (x, y, z, Z) per atom
  size_t dev_bytes = 0;
};

struct FwdPlan {
  int n_pb = 0;   // point blocks
  int ld = 0;     // leading dimension of a partial row
  int rows = 0;   // total partial rows
  int chunks[FE_LMAX + 1] = {};
  int ppc[FE_LMAX + 1] = {};
  int row0[FE_LMAX + 1] = {};
};

int fe_ctx_reserve(fe_ctx* c, int slot, size_t bytes);
FwdPlan fe_plan(const fe_table* t, int64_t n, int ppl, int target_wg);
int fe_stage_points(fe_ctx* c, const double* pts, int64_t n, int ptr_kind, hipStream_t st,
                    const double** d_pts);
int fe_launch_fwd_mixed(fe_table* t, const FwdPlan& pl, const double* pts, int n,
                        double* partial, hipStream_t st);
