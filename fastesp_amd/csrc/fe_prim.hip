// Device primitives of the table build and the point sort: a stable LSD radix
// sort of (uint32 key, int32 value) pairs and exclusive prefix sums.
//
// Both are deterministic, issue no memsets (every buffer they clear is
// written by a kernel) and never synchronise with the host, so a stream that
// is being captured into a hipGraph can record them (the library sorts from
// rocPRIM clear their counters with hipMemsetAsync, which a captured graph
// must not contain).
//
// Radix sort: 8-bit digits, one pass per digit.  A pass is a digit histogram
// per tile of 1024 keys (LDS counters), an exclusive scan of the histograms in
// digit-major order (digit d of tile b lands after every tile's smaller digits
// and after digit d of tiles < b), and a scatter that ranks equal digits
// within a tile in input order: per wavefront by ballots over the digit's
// eight bits, across the four wavefronts and the four rounds of a tile by LDS
// counters.  So the sort is stable, and the output does not depend on timing.
#include <hip/hip_runtime.h>

#include "fe_internal.h"

namespace {

constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = 256 * RS_ITEMS;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = 256 * SC_ITEMS;

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// hist[d * gridDim.x + tile] = number of keys of the tile with digit d
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t* __restrict__ keys, const int n,
                                                 const int shift, int* __restrict__ hist) {
  __shared__ int s_cnt[256];
  s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_TILE;
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int i = base + r * 256 + (int)threadIdx.x;
    if (i < n) atomicAdd(&s_cnt[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = s_cnt[threadIdx.x];
}

// stable scatter of one tile to the offsets of the scanned histogram
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                    const int32_t* __restrict__ vin, const int n,
                                                    const int shift,
                                                    const int* __restrict__ hist_scan,
                                                    uint32_t* __restrict__ kout,
                                                    int32_t* __restrict__ vout) {
  __shared__ int s_base[256];   // next output position of digit d for this tile
  __shared__ int s_wc[4][256];  // keys of digit d in wavefront w of the current round
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  s_base[tid] = hist_scan[tid * gridDim.x + blockIdx.x];
#pragma unroll
  for (int w = 0; w < 4; ++w) s_wc[w][tid] = 0;
  __syncthreads();
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int base = blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int i = base + r * 256 + tid;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const int32_t v = valid ? vin[i] : 0;
    const int d = (int)((k >> shift) & 255u);
    // lanes of this wavefront holding a valid key with the same digit
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const int rank = __popcll(same & below);
    if (valid && rank == 0) s_wc[wave][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      int o = s_base[d] + rank;
      for (int w = 0; w < wave; ++w) o += s_wc[w][d];
      kout[o] = k;
      vout[o] = v;
    }
    __syncthreads();
    s_base[tid] += (s_wc[0][tid] + s_wc[1][tid]) + (s_wc[2][tid] + s_wc[3][tid]);
#pragma unroll
    for (int w = 0; w < 4; ++w) s_wc[w][tid] = 0;
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_pairs(const T* __restrict__ a, const int32_t* __restrict__ b,
                                                    const int n, T* __restrict__ ao,
                                                    int32_t* __restrict__ bo) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    ao[i] = a[i];
    bo[i] = b[i];
  }
}

// ---------------------------------------------------------------------------
// exclusive prefix sums (integer: any grouping gives the same result)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T fe_block_excl_scan(const T v, T* s_w, T& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  T pre = 0;
  for (int w = 0; w < wave; ++w) pre += s_w[w];
  total = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  __syncthreads();
  return pre + x - v;
}

// sums[b] = sum of tile b
template <typename T>
__global__ __launch_bounds__(256) void k_scan_reduce(const T* __restrict__ in, const int n,
                                                     T* __restrict__ sums) {
  __shared__ T s_w[4];
  const int b0 = blockIdx.x * SC_TILE;
  T s = 0;
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) {
    const int i = b0 + r * 256 + (int)threadIdx.x;
    if (i < n) s += in[i];
  }
  T total;
  (void)fe_block_excl_scan<T>(s, s_w, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of sums[0..nb) in place by one block
template <typename T>
__global__ __launch_bounds__(256) void k_scan_top(T* __restrict__ sums, const int nb) {
  __shared__ T s_w[4];
  T carry = 0;
  for (int c0 = 0; c0 < nb; c0 += SC_TILE) {
    // thread t holds the SC_ITEMS consecutive entries c0 + t*SC_ITEMS + j
    T v[SC_ITEMS];
    T s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
      const int i = c0 + (int)threadIdx.x * SC_ITEMS + j;
      v[j] = i < nb ? sums[i] : (T)0;
      s += v[j];
    }
    T total;
    T pre = fe_block_excl_scan<T>(s, s_w, total) + carry;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
      const int i = c0 + (int)threadIdx.x * SC_ITEMS + j;
      if (i < nb) sums[i] = pre;
      pre += v[j];
    }
    carry += total;
  }
}

// out[i] = sums[tile] + exclusive prefix of in within the tile
template <typename T>
__global__ __launch_bounds__(256) void k_scan_down(const T* __restrict__ in, const int n,
                                                   const T* __restrict__ sums,
                                                   T* __restrict__ out) {
  __shared__ T s_w[4];
  const int b0 = blockIdx.x * SC_TILE + (int)threadIdx.x * SC_ITEMS;
  T v[SC_ITEMS];
  T s = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    v[j] = b0 + j < n ? in[b0 + j] : (T)0;
    s += v[j];
  }
  T total;
  T pre = fe_block_excl_scan<T>(s, s_w, total) + sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    if (b0 + j < n) out[b0 + j] = pre;
    pre += v[j];
  }
}

// ---------------------------------------------------------------------------
template <typename T>
size_t fe_scan_scratch_t(int64_t n) {
  const int64_t nb = (n + SC_TILE - 1) / SC_TILE;
  return al256((size_t)std::max<int64_t>(nb, 1) * sizeof(T));
}

size_t fe_scan_scratch(int64_t n) { return fe_scan_scratch_t<unsigned long long>(n); }

template <typename T>
static int scan_excl(const T* in, T* out, int n, void* scratch, hipStream_t st) {
  if (n <= 0) return FE_OK;
  const int nb = (n + SC_TILE - 1) / SC_TILE;
  T* sums = (T*)scratch;
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(256), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(256), 0, st, sums, nb);
  hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(256), 0, st, in, n, sums, out);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

int fe_scan_excl_i32(const int* in, int* out, int n, void* scratch, hipStream_t st) {
  return scan_excl<int>(in, out, n, scratch, st);
}

int fe_scan_excl_u64(const unsigned long long* in, unsigned long long* out, int n, void* scratch,
                     hipStream_t st) {
  return scan_excl<unsigned long long>(in, out, n, scratch, st);
}

size_t fe_radix_sort_scratch(int64_t n) {
  const int64_t nt = std::max<int64_t>((n + RS_TILE - 1) / RS_TILE, 1);
  return 2 * al256((size_t)std::max<int64_t>(n, 1) * 4) + al256((size_t)256 * nt * sizeof(int)) +
         fe_scan_scratch_t<int>(256 * nt);
}

int fe_radix_sort(uint32_t* keys, int32_t* vals, int n, int key_bits, void* scratch,
                  hipStream_t st) {
  if (n <= 1) return FE_OK;
  if (key_bits < 1 || key_bits > 32)
    return fe_set_error(FE_ERR_ARG, "radix sort: key_bits %d out of range", key_bits);
  const int nt = (n + RS_TILE - 1) / RS_TILE;
  char* s = (char*)scratch;
  uint32_t* kalt = (uint32_t*)s;
  s += al256((size_t)n * 4);
  int32_t* valt = (int32_t*)s;
  s += al256((size_t)n * 4);
  int* hist = (int*)s;
  s += al256((size_t)256 * nt * sizeof(int));
  void* scan_tmp = s;
  const int passes = (key_bits + 7) / 8;
  uint32_t* ks = keys;
  int32_t* vs = vals;
  uint32_t* kd = kalt;
  int32_t* vd = valt;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(nt), dim3(256), 0, st, ks, n, shift, hist);
    if (int rc = scan_excl<int>(hist, hist, 256 * nt, scan_tmp, st)) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nt), dim3(256), 0, st, ks, vs, n, shift, hist, kd, vd);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != keys)
    hipLaunchKernelGGL(k_copy_pairs<uint32_t>, dim3((n + 255) / 256), dim3(256), 0, st, ks, vs, n,
                       keys, vals);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}
