// Electron density at points and vdW-surface point candidates.
//
// rho(C) = sum_{mu nu} phi_mu(C) P_{mu nu} phi_nu(C)   (benchmarks/
// check_density_mwfn.py:60-75), phi_mu the contracted Cartesian functions of
// the basis.bin fields (fastesp/table.py:268-293).
//
// Layout and kernels (one chunk of up to FE_DENS_CHUNK points at a time):
//  * k_basis_values: phiT[mu][i] = (x-A)^lx (y-A)^ly (z-A)^lz sum_k c_k
//    exp(-a_k r^2), functions padded to a multiple of 64 (zero rows), points
//    padded to the chunk width (zero columns).  HBM: 8 B per (point, function).
//  * k_density_mm: per tile of 64 points x block J of 64 functions,
//    Y = Phi[:, K <= J] Pu[K <= J, J] on the FP64 matrix cores
//    (v_mfma_f64_16x16x4f64), then the row-dot sum_j Phi[i,j] Y[i,j] in the
//    epilogue; Y never leaves registers.  Pu is the symmetrised density with the
//    blocks above the diagonal doubled and those below dropped, so the
//    quadratic form costs N n (n + 64) FLOPs instead of 2 N n^2.
//  * k_density_reduce: rho[i] = sum_J partial[J][i] in block order
//    (bitwise reproducible).
// The roofline is the FP64 matrix peak: a 64 x 64 output tile reads
// 2 x 64 x 32 doubles (mostly L2 hits) per 2 x 64 x 64 x 32 FLOPs.
#include <algorithm>
#include <cmath>

#include "fe_internal.h"

#define FE_DENS_TILE 64
#define FE_DENS_BK 32
#define FE_DENS_CHUNK 32768

typedef double fe_v4d __attribute__((ext_vector_type(4)));

struct fe_density {
  fe_ctx* ctx = nullptr;
  int n = 0, npad = 0;
  double* d_Pu = nullptr;      // npad x npad
  double* d_fn = nullptr;      // per function: Ax, Ay, Az
  int* d_fi = nullptr;         // per function: lx, ly, lz, prim offset, prim count
  double* d_prim = nullptr;    // per primitive: exponent, coefficient
  double* d_phiT = nullptr;    // npad x chunk
  double* d_part = nullptr;    // (npad / 64) x chunk
  int chunk_cap = 0;
  double* d_pts = nullptr;     // staged host points
  double* d_rho = nullptr;     // staged host output
  int64_t io_cap = 0;
  int64_t last_flops = 0;      // FLOPs of the matrix kernel in the last evaluation
  bool tile128 = false;        // 128 x 64 tiles (k_density_mm128) instead of 64 x 64
};

__global__ __launch_bounds__(256) void k_density_pack(const double* __restrict__ P, const int n,
                                                      const int npad, double* __restrict__ Pu) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)npad * npad) return;
  const int k = (int)(e / npad), j = (int)(e - (int64_t)k * npad);
  double v = 0.0;
  if (k < n && j < n) {
    const int kb = k / FE_DENS_TILE, jb = j / FE_DENS_TILE;
    const double s = P[(size_t)k * n + j] + P[(size_t)j * n + k];
    v = kb < jb ? s : kb == jb ? 0.5 * s : 0.0;
  }
  Pu[e] = v;
}

__global__ __launch_bounds__(256) void k_basis_values(const double* __restrict__ pts,
                                                      const int64_t base, const int64_t n_pts,
                                                      const int ld, const double* __restrict__ fn,
                                                      const int* __restrict__ fi,
                                                      const double* __restrict__ prim,
                                                      const int n, double* __restrict__ phiT) {
  const int i = blockIdx.x * 256 + threadIdx.x;  // point within the chunk
  const int mu = blockIdx.y;
  if (i >= ld) return;
  double v = 0.0;
  if (mu < n && base + i < n_pts) {
    const double* c = pts + 3 * (base + i);
    const double dx = c[0] - fn[3 * mu], dy = c[1] - fn[3 * mu + 1], dz = c[2] - fn[3 * mu + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const int* f = fi + 5 * mu;
    double acc = 0.0;
    for (int k = 0; k < f[4]; ++k) {
      const double* pr = prim + 2 * (f[3] + k);
      acc += pr[1] * exp(-pr[0] * r2);
    }
    double m = 1.0;
    for (int q = 0; q < f[0]; ++q) m *= dx;
    for (int q = 0; q < f[1]; ++q) m *= dy;
    for (int q = 0; q < f[2]; ++q) m *= dz;
    v = m * acc;
  }
  phiT[(size_t)mu * ld + i] = v;
}

// grid (chunk / 64, npad / 64), 256 threads: wave w owns rows 32 (w / 2) ..
// and columns 32 (w % 2) .. of the 64 x 64 tile, as 2 x 2 MFMA 16 x 16 blocks.
// MFMA 16x16x4 f64 lane layout: A[i][k] and B[k][j] with i, j = lane % 16,
// k = lane / 16; D[i][j] in register i / 4 of lane j + 16 (i % 4).
// The K loop is software-pipelined: the next K slab is loaded into registers
// while the matrix cores work on the slab in LDS.
__global__ __launch_bounds__(256) void k_density_mm(const double* __restrict__ phiT, const int ld,
                                                    const double* __restrict__ Pu, const int npad,
                                                    double* __restrict__ partial) {
  constexpr int NQ = FE_DENS_BK * FE_DENS_TILE / 256;  // doubles per thread per slab
  __shared__ double As[FE_DENS_BK][FE_DENS_TILE];
  __shared__ double Bs[FE_DENS_BK][FE_DENS_TILE];
  __shared__ double red[2][FE_DENS_TILE];
  const int it = blockIdx.x, J = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const size_t col0 = (size_t)it * FE_DENS_TILE;
  fe_v4d acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
  const int kend = (J + 1) * FE_DENS_TILE;
  // slab element q of this thread: row k = (tid >> 6) + 4 q, column c = lane
  const double* pa = phiT + (size_t)(tid >> 6) * ld + col0 + lane;
  const double* pb = Pu + (size_t)(tid >> 6) * npad + (size_t)J * FE_DENS_TILE + lane;
  double ra[NQ], rb[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    ra[q] = pa[(size_t)4 * q * ld];
    rb[q] = pb[(size_t)4 * q * npad];
  }
  for (int k0 = 0; k0 < kend; k0 += FE_DENS_BK) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      As[(tid >> 6) + 4 * q][lane] = ra[q];
      Bs[(tid >> 6) + 4 * q][lane] = rb[q];
    }
    __syncthreads();
    if (k0 + FE_DENS_BK < kend) {
      const size_t ka = (size_t)(k0 + FE_DENS_BK);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        ra[q] = pa[(ka + 4 * q) * ld];
        rb[q] = pb[(ka + 4 * q) * npad];
      }
    }
#pragma unroll
    for (int kk = 0; kk < FE_DENS_BK; kk += 4) {
      const double a0 = As[kk + lk][wr + li], a1 = As[kk + lk][wr + 16 + li];
      const double b0 = Bs[kk + lk][wc + li], b1 = Bs[kk + lk][wc + 16 + li];
      acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: s[rb][r] = sum over this lane's two columns of Phi * Y for row
  // wr + 16 rb + 4 r + lk; then a sum over the 16 lanes that share lk
  const size_t jc0 = (size_t)J * FE_DENS_TILE + wc + li, jc1 = jc0 + 16;
  double s[2][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const size_t row0 = col0 + wr + 4 * r + lk, row1 = row0 + 16;
    s[0][r] = acc00[r] * phiT[jc0 * ld + row0] + acc01[r] * phiT[jc1 * ld + row0];
    s[1][r] = acc10[r] * phiT[jc0 * ld + row1] + acc11[r] * phiT[jc1 * ld + row1];
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = s[rb][r];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      s[rb][r] = v;
    }
  if (li == 0) {
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w & 1][wr + 16 * rb + 4 * r + lk] = s[rb][r];
  }
  __syncthreads();
  if (tid < FE_DENS_TILE) partial[(size_t)J * ld + col0 + tid] = red[0][tid] + red[1][tid];
}

// 128 points x 64 functions per workgroup: wave w owns rows 32 w .. 32 w + 31
// and all 64 columns (2 x 4 MFMA blocks), so each LDS fragment feeds more
// matrix instructions.  Same slab pipeline and epilogue as k_density_mm.
__global__ __launch_bounds__(256) void k_density_mm128(const double* __restrict__ phiT,
                                                       const int ld,
                                                       const double* __restrict__ Pu,
                                                       const int npad,
                                                       double* __restrict__ partial) {
  constexpr int BM = 2 * FE_DENS_TILE, BK = 16;  // BK 16 keeps the slab registers small
  constexpr int NQA = BK * BM / 256, NQB = BK * FE_DENS_TILE / 256;
  __shared__ double As[BK][BM];
  __shared__ double Bs[BK][FE_DENS_TILE];
  const int it = blockIdx.x, J = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = 32 * w;
  const int li = lane & 15, lk = lane >> 4;
  const size_t col0 = (size_t)it * BM;
  fe_v4d acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = fe_v4d{0, 0, 0, 0};
  const int kend = (J + 1) * FE_DENS_TILE;
  // A slab element q of this thread: row (tid >> 7) + 2 q, column tid & 127;
  // B slab element q: row (tid >> 6) + 4 q, column lane
  const double* pa = phiT + (size_t)(tid >> 7) * ld + col0 + (tid & 127);
  const double* pb = Pu + (size_t)(tid >> 6) * npad + (size_t)J * FE_DENS_TILE + lane;
  double ra[NQA], rb[NQB];
#pragma unroll
  for (int q = 0; q < NQA; ++q) ra[q] = pa[(size_t)2 * q * ld];
#pragma unroll
  for (int q = 0; q < NQB; ++q) rb[q] = pb[(size_t)4 * q * npad];
  for (int k0 = 0; k0 < kend; k0 += BK) {
#pragma unroll
    for (int q = 0; q < NQA; ++q) As[(tid >> 7) + 2 * q][tid & 127] = ra[q];
#pragma unroll
    for (int q = 0; q < NQB; ++q) Bs[(tid >> 6) + 4 * q][lane] = rb[q];
    __syncthreads();
    if (k0 + BK < kend) {
      const size_t ka = (size_t)(k0 + BK);
#pragma unroll
      for (int q = 0; q < NQA; ++q) ra[q] = pa[(ka + 2 * q) * ld];
#pragma unroll
      for (int q = 0; q < NQB; ++q) rb[q] = pb[(ka + 4 * q) * npad];
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[2], bv[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[kk + lk][wr + 16 * a + li];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk + lk][16 * b + li];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // row wr + 16 a + 4 r + lk, columns 16 b + li of block J
  double s[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const size_t row = col0 + wr + 16 * a + 4 * r + lk;
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        v += acc[a][b][r] * phiT[((size_t)J * FE_DENS_TILE + 16 * b + li) * ld + row];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      s[a][r] = v;
    }
  if (li == 0) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        partial[(size_t)J * ld + col0 + wr + 16 * a + 4 * r + lk] = s[a][r];
  }
}

__global__ __launch_bounds__(256) void k_density_reduce(const double* __restrict__ partial,
                                                        const int nJ, const int ld, const int m,
                                                        double* __restrict__ rho) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int J = 0; J < nJ; ++J) s += partial[(size_t)J * ld + i];
  rho[i] = s;
}

// vdW-surface candidates: point k of atom a's Fibonacci sphere of cnt[a]
// points (examples/vdw_surface_esp.py:28-52), and whether it lies outside
// every sphere (|p - R_b| >= rad_b - 1e-9 for all b).
__global__ __launch_bounds__(256) void k_vdw_candidates(const double* __restrict__ xyz,
                                                        const double* __restrict__ rad,
                                                        const int n_atoms,
                                                        const int64_t* __restrict__ off,
                                                        const int64_t n_total,
                                                        double* __restrict__ pts,
                                                        int* __restrict__ ok) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n_total) return;
  int lo = 0, hi = n_atoms;  // last atom with off[a] <= e
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid; else hi = mid;
  }
  const int a = lo;
  const double n = (double)(off[a + 1] - off[a]);
  const double i = (double)(e - off[a]) + 0.5;
  const double phi = acos(1.0 - 2.0 * i / n);
  const double theta = M_PI * (1.0 + sqrt(5.0)) * i;
  const double sp = sin(phi);
  const double p0 = xyz[3 * a] + rad[a] * (sp * cos(theta));
  const double p1 = xyz[3 * a + 1] + rad[a] * (sp * sin(theta));
  const double p2 = xyz[3 * a + 2] + rad[a] * cos(phi);
  int keep = 1;
  for (int b = 0; b < n_atoms; ++b) {
    const double dx = p0 - xyz[3 * b], dy = p1 - xyz[3 * b + 1], dz = p2 - xyz[3 * b + 2];
    if (!(sqrt(dx * dx + dy * dy + dz * dz) >= rad[b] - 1e-9)) keep = 0;
  }
  pts[3 * e] = p0;
  pts[3 * e + 1] = p1;
  pts[3 * e + 2] = p2;
  ok[e] = keep;
}

// ---------------------------------------------------------------------------
static void dens_free_ws(fe_density* d) {
  if (d->d_phiT) (void)hipFree(d->d_phiT);
  if (d->d_part) (void)hipFree(d->d_part);
  d->d_phiT = d->d_part = nullptr;
  d->chunk_cap = 0;
}

extern "C" int fe_density_free(fe_density* d) {
  if (!d) return FE_OK;
  (void)hipSetDevice(d->ctx->device);
  dens_free_ws(d);
  for (void* p : {(void*)d->d_Pu, (void*)d->d_fn, (void*)d->d_fi, (void*)d->d_prim,
                  (void*)d->d_pts, (void*)d->d_rho})
    if (p) (void)hipFree(p);
  delete d;
  return FE_OK;
}

extern "C" int fe_density_create(fe_ctx* c, int n_func, int n_prim, const int32_t* fn_nprim,
                                 const int32_t* fn_off, const int32_t* fn_lx,
                                 const int32_t* fn_ly, const int32_t* fn_lz, const double* fc,
                                 const double* pexp, const double* pcoef, const double* dens,
                                 int ptr_kind, void* stream, fe_density** out) {
  if (!c || !out) return fe_set_error(FE_ERR_ARG, "fe_density_create: null argument");
  *out = nullptr;
  if (n_func <= 0 || n_prim <= 0 || n_func > 65535 - FE_DENS_TILE)
    return fe_set_error(FE_ERR_ARG, "fe_density_create: bad sizes (n_func %d)", n_func);
  if (!fn_nprim || !fn_off || !fn_lx || !fn_ly || !fn_lz || !fc || !pexp || !pcoef || !dens)
    return fe_set_error(FE_ERR_ARG, "fe_density_create: null array");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_density_create: bad ptr_kind");
  for (int f = 0; f < n_func; ++f) {
    if (fn_nprim[f] < 0 || fn_off[f] < 0 || (int64_t)fn_off[f] + fn_nprim[f] > n_prim)
      return fe_set_error(FE_ERR_ARG, "fe_density_create: function %d primitives out of range", f);
    if (fn_lx[f] < 0 || fn_ly[f] < 0 || fn_lz[f] < 0 || fn_lx[f] + fn_ly[f] + fn_lz[f] > 8)
      return fe_set_error(FE_ERR_ARG, "fe_density_create: function %d angular part", f);
  }
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  fe_density* d = new fe_density();
  d->ctx = c;
  d->n = n_func;
  d->npad = (n_func + FE_DENS_TILE - 1) / FE_DENS_TILE * FE_DENS_TILE;
  std::vector<double> h_fn(3 * (size_t)n_func), h_prim(2 * (size_t)n_prim);
  std::vector<int> h_fi(5 * (size_t)n_func);
  for (int f = 0; f < n_func; ++f) {
    for (int k = 0; k < 3; ++k) h_fn[3 * f + k] = fc[3 * f + k];
    h_fi[5 * f] = fn_lx[f];
    h_fi[5 * f + 1] = fn_ly[f];
    h_fi[5 * f + 2] = fn_lz[f];
    h_fi[5 * f + 3] = fn_off[f];
    h_fi[5 * f + 4] = fn_nprim[f];
  }
  for (int p = 0; p < n_prim; ++p) {
    h_prim[2 * p] = pexp[p];
    h_prim[2 * p + 1] = pcoef[p];
  }
  const size_t nn = (size_t)d->npad * d->npad;
  double* d_P = nullptr;
  auto fail = [&](int rc) {
    if (d_P && ptr_kind == FE_PTR_HOST) (void)hipFree(d_P);
    fe_density_free(d);
    return rc;
  };
  if (hipMalloc(&d->d_Pu, nn * sizeof(double)) != hipSuccess ||
      hipMalloc(&d->d_fn, h_fn.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&d->d_fi, h_fi.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->d_prim, h_prim.size() * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(fe_set_error(FE_ERR_ALLOC, "fe_density_create: allocation failed"));
  }
  if (hipMemcpyAsync(d->d_fn, h_fn.data(), h_fn.size() * sizeof(double), hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      hipMemcpyAsync(d->d_fi, h_fi.data(), h_fi.size() * sizeof(int), hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      hipMemcpyAsync(d->d_prim, h_prim.data(), h_prim.size() * sizeof(double),
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_density_create: upload failed"));
  if (ptr_kind == FE_PTR_HOST) {
    if (hipMalloc(&d_P, (size_t)n_func * n_func * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(fe_set_error(FE_ERR_ALLOC, "fe_density_create: allocation failed"));
    }
    if (hipMemcpyAsync(d_P, dens, (size_t)n_func * n_func * sizeof(double),
                       hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(fe_set_error(FE_ERR_HIP, "fe_density_create: upload failed"));
  } else {
    d_P = (double*)dens;
  }
  hipLaunchKernelGGL(k_density_pack, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, d_P,
                     n_func, d->npad, d->d_Pu);
  if (hipGetLastError() != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_density_create: launch failed"));
  if (hipStreamSynchronize(st) != hipSuccess)
    return fail(fe_set_error(FE_ERR_HIP, "fe_density_create: %s",
                             hipGetErrorString(hipGetLastError())));
  if (ptr_kind == FE_PTR_HOST) (void)hipFree(d_P);
  *out = d;
  return FE_OK;
}

extern "C" int fe_density_eval(fe_density* d, const double* pts, int64_t n_points, int ptr_kind,
                               void* stream, double* rho) {
  if (!d || (n_points > 0 && (!pts || !rho)))
    return fe_set_error(FE_ERR_ARG, "fe_density_eval: null argument");
  if (n_points < 0 || n_points > (int64_t)INT32_MAX)
    return fe_set_error(FE_ERR_ARG, "fe_density_eval: point count out of range");
  if (ptr_kind != FE_PTR_HOST && ptr_kind != FE_PTR_DEVICE)
    return fe_set_error(FE_ERR_ARG, "fe_density_eval: bad ptr_kind");
  if (n_points == 0) return FE_OK;
  fe_ctx* c = d->ctx;
  FE_CTX_LOCK(c);
  FE_HIPCHK(hipSetDevice(c->device));
  hipStream_t st = (hipStream_t)stream;
  // points padded to 128 so both tile widths divide every chunk
  const int64_t chunk = std::min<int64_t>(FE_DENS_CHUNK, (n_points + 127) / 128 * 128);
  const int nJ = d->npad / FE_DENS_TILE;
  if (chunk > d->chunk_cap) {
    dens_free_ws(d);
    if (hipMalloc(&d->d_phiT, (size_t)d->npad * chunk * sizeof(double)) != hipSuccess ||
        hipMalloc(&d->d_part, (size_t)nJ * chunk * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      dens_free_ws(d);
      return fe_set_error(FE_ERR_ALLOC, "fe_density_eval: workspace allocation failed");
    }
    d->chunk_cap = (int)chunk;
  }
  const double* d_pts = pts;
  double* d_rho = rho;
  if (ptr_kind == FE_PTR_HOST) {
    if (n_points > d->io_cap) {
      if (d->d_pts) (void)hipFree(d->d_pts);
      if (d->d_rho) (void)hipFree(d->d_rho);
      d->d_pts = d->d_rho = nullptr;
      d->io_cap = 0;
      if (hipMalloc(&d->d_pts, (size_t)n_points * 3 * sizeof(double)) != hipSuccess ||
          hipMalloc(&d->d_rho, (size_t)n_points * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fe_set_error(FE_ERR_ALLOC, "fe_density_eval: staging allocation failed");
      }
      d->io_cap = n_points;
    }
    FE_HIPCHK(hipMemcpyAsync(d->d_pts, pts, (size_t)n_points * 3 * sizeof(double),
                             hipMemcpyHostToDevice, st));
    d_pts = d->d_pts;
    d_rho = d->d_rho;
  }
  d->last_flops = 0;
  for (int64_t base = 0; base < n_points; base += chunk) {
    const int m = (int)std::min<int64_t>(chunk, n_points - base);
    const int ld = (m + 127) / 128 * 128;
    int ps = fe_prof_begin(c, FE_PH_DENS_PHI, st);
    hipLaunchKernelGGL(k_basis_values, dim3((unsigned)((ld + 255) / 256), (unsigned)d->npad),
                       dim3(256), 0, st, d_pts, base, n_points, ld, d->d_fn, d->d_fi, d->d_prim,
                       d->n, d->d_phiT);
    fe_prof_end(c, ps, st);
    ps = fe_prof_begin(c, FE_PH_DENS_MM, st);
    if (d->tile128)
      hipLaunchKernelGGL(k_density_mm128, dim3((unsigned)(ld / (2 * FE_DENS_TILE)), (unsigned)nJ),
                         dim3(256), 0, st, d->d_phiT, ld, d->d_Pu, d->npad, d->d_part);
    else
      hipLaunchKernelGGL(k_density_mm, dim3((unsigned)(ld / FE_DENS_TILE), (unsigned)nJ),
                         dim3(256), 0, st, d->d_phiT, ld, d->d_Pu, d->npad, d->d_part);
    fe_prof_end(c, ps, st);
    ps = fe_prof_begin(c, FE_PH_DENS_RED, st);
    hipLaunchKernelGGL(k_density_reduce, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                       d->d_part, nJ, ld, m, d_rho + base);
    fe_prof_end(c, ps, st);
    d->last_flops += (int64_t)ld * 2 * FE_DENS_TILE * FE_DENS_TILE * ((int64_t)nJ * (nJ + 1) / 2);
  }
  FE_HIPCHK(hipGetLastError());
  if (ptr_kind == FE_PTR_HOST) {
    FE_HIPCHK(hipMemcpyAsync(rho, d_rho, (size_t)n_points * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    FE_HIPCHK(hipStreamSynchronize(st));
  }
  return FE_OK;
}

extern "C" int fe_density_set_tuning(fe_density* d, int key, int64_t value) {
  if (!d) return fe_set_error(FE_ERR_ARG, "fe_density_set_tuning: null argument");
  if (key != 0) return fe_set_error(FE_ERR_ARG, "fe_density_set_tuning: unknown key %d", key);
  d->tile128 = value != 0;
  return FE_OK;
}

extern "C" int fe_density_info(const fe_density* d, int64_t* info, int n_info) {
  if (!d || !info) return fe_set_error(FE_ERR_ARG, "fe_density_info: null argument");
  const int64_t v[4] = {d->n, d->npad, d->last_flops, (int64_t)d->chunk_cap};
  for (int i = 0; i < n_info && i < 4; ++i) info[i] = v[i];
  return FE_OK;
}

extern "C" int fe_vdw_candidates(fe_ctx* c, int n_atoms, const double* xyz, const double* rad,
                                 const int64_t* counts, double* pts, int32_t* ok) {
  if (!c || n_atoms <= 0 || !xyz || !rad || !counts || !pts || !ok)
    return fe_set_error(FE_ERR_ARG, "fe_vdw_candidates: bad argument");
  std::vector<int64_t> off(n_atoms + 1, 0);
  for (int a = 0; a < n_atoms; ++a) {
    if (counts[a] <= 0 || !(rad[a] > 0.0))
      return fe_set_error(FE_ERR_ARG, "fe_vdw_candidates: atom %d count/radius", a);
    off[a + 1] = off[a] + counts[a];
  }
  const int64_t n = off[n_atoms];
  if (n > (int64_t)INT32_MAX) return fe_set_error(FE_ERR_ARG, "fe_vdw_candidates: too many points");
  FE_HIPCHK(hipSetDevice(c->device));
  double *d_xyz = nullptr, *d_rad = nullptr, *d_pts = nullptr;
  int64_t* d_off = nullptr;
  int* d_ok = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)d_xyz, (void*)d_rad, (void*)d_pts, (void*)d_off, (void*)d_ok})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc(&d_xyz, 3 * (size_t)n_atoms * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_rad, (size_t)n_atoms * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_off, (size_t)(n_atoms + 1) * sizeof(int64_t)) != hipSuccess ||
      hipMalloc(&d_pts, 3 * (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_ok, (size_t)n * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return fe_set_error(FE_ERR_ALLOC, "fe_vdw_candidates: allocation failed");
  }
  hipError_t e = hipMemcpy(d_xyz, xyz, 3 * (size_t)n_atoms * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_rad, rad, (size_t)n_atoms * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_off, off.data(), (size_t)(n_atoms + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_vdw_candidates, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       d_xyz, d_rad, n_atoms, d_off, n, d_pts, d_ok);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(pts, d_pts, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ok, d_ok, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fe_set_error(FE_ERR_HIP, "fe_vdw_candidates: %s", hipGetErrorString(e));
  }
  return FE_OK;
}
