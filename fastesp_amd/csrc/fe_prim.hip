// Device primitives of the table build and the point sort: a stable LSD radix
// sort of (uint32 key, int32 value) pairs and exclusive prefix sums.
//
// Both are deterministic, issue no memsets (every buffer they clear is
// written by a kernel) and never synchronise with the host, so a stream that
// is being captured into a hipGraph can record them (the library sorts from
// rocPRIM clear their counters with hipMemsetAsync, which a captured graph
// must not contain).
//
// Radix sort: 8-bit digits, one pass per digit, three launches a pass.  A
// pass is a digit histogram per tile of 1024 keys (LDS counters); an exclusive
// scan of each digit's row of tile counts, one block per digit, which also
// leaves the digit's total; and a scatter that starts digit d of tile b after
// every key with a smaller digit (a block scan of the 256 totals) and after
// digit d of the tiles < b, and ranks equal digits within a tile in input
// order: per wavefront by ballots over the digit's eight bits, across the four
// wavefronts and the four rounds of a tile by LDS counters.  So the sort is
// stable, and the output does not depend on timing.
#include <hip/hip_runtime.h>

#include "fe_internal.h"

namespace {

constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = 256 * RS_ITEMS;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = 256 * SC_ITEMS;

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

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

// hist[d * nt + b] -> the count of digit d in tiles < b (exclusive scan of the
// row, one block per digit d); tot[d] = the row's total
__global__ __launch_bounds__(256) void k_rs_rowscan(int* __restrict__ hist, const int nt,
                                                    int* __restrict__ tot) {
  __shared__ int s_w[4];
  int* row = hist + (size_t)blockIdx.x * nt;
  int carry = 0;
  for (int c0 = 0; c0 < nt; c0 += SC_TILE) {
    // thread t holds the SC_ITEMS consecutive entries c0 + t*SC_ITEMS + j
    int v[SC_ITEMS];
    int s = 0;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
      const int i = c0 + (int)threadIdx.x * SC_ITEMS + j;
      v[j] = i < nt ? row[i] : 0;
      s += v[j];
    }
    int total;
    int pre = fe_block_excl_scan<int>(s, s_w, total) + carry;
#pragma unroll
    for (int j = 0; j < SC_ITEMS; ++j) {
      const int i = c0 + (int)threadIdx.x * SC_ITEMS + j;
      if (i < nt) row[i] = pre;
      pre += v[j];
    }
    carry += total;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// stable scatter of one tile: digit d starts after every key of a smaller
// digit (exclusive scan of tot) and after digit d of the tiles before it
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                    const int32_t* __restrict__ vin, const int n,
                                                    const int shift,
                                                    const int* __restrict__ row_scan,
                                                    const int* __restrict__ tot,
                                                    uint32_t* __restrict__ kout,
                                                    int32_t* __restrict__ vout) {
  __shared__ int s_base[256];   // next output position of digit d for this tile
  __shared__ int s_wc[4][256];  // keys of digit d in wavefront w of the current round
  __shared__ int s_w[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int total;
  s_base[tid] = fe_block_excl_scan<int>(tot[tid], s_w, total) +
                row_scan[tid * gridDim.x + blockIdx.x];
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

// out[i] = (sum of the tiles before this one) + exclusive prefix of in within
// the tile; each block adds the tile sums before it itself (a few hundred)
template <typename T>
__global__ __launch_bounds__(256) void k_scan_down(const T* __restrict__ in, const int n,
                                                   const T* __restrict__ sums,
                                                   T* __restrict__ out) {
  __shared__ T s_w[4];
  T before = 0;
  for (int b = (int)threadIdx.x; b < (int)blockIdx.x; b += 256) before += sums[b];
  {
    T all;
    (void)fe_block_excl_scan<T>(before, s_w, all);
    before = all;
  }
  const int b0 = blockIdx.x * SC_TILE + (int)threadIdx.x * SC_ITEMS;
  T v[SC_ITEMS];
  T s = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    v[j] = b0 + j < n ? in[b0 + j] : (T)0;
    s += v[j];
  }
  T total;
  T pre = fe_block_excl_scan<T>(s, s_w, total) + before;
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
  hipLaunchKernelGGL(k_scan_down<T>, dim3(nb), dim3(256), 0, st, in, n, sums, out);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

int fe_scan_excl_u64(const unsigned long long* in, unsigned long long* out, int n, void* scratch,
                     hipStream_t st) {
  return scan_excl<unsigned long long>(in, out, n, scratch, st);
}

size_t fe_radix_sort_scratch(int64_t n) {
  const int64_t nt = std::max<int64_t>((n + RS_TILE - 1) / RS_TILE, 1);
  return 2 * al256((size_t)std::max<int64_t>(n, 1) * 4) + al256((size_t)256 * nt * sizeof(int)) +
         al256(256 * sizeof(int));
}

int fe_radix_sort_bits(uint32_t* keys, int32_t* vals, int n, int lo_bit, int hi_bit,
                       void* scratch, hipStream_t st) {
  if (n <= 1) return FE_OK;
  if (lo_bit < 0 || hi_bit > 32 || hi_bit <= lo_bit)
    return fe_set_error(FE_ERR_ARG, "radix sort: key bits [%d, %d) out of range", lo_bit, hi_bit);
  const int nt = (n + RS_TILE - 1) / RS_TILE;
  char* s = (char*)scratch;
  uint32_t* kalt = (uint32_t*)s;
  s += al256((size_t)n * 4);
  int32_t* valt = (int32_t*)s;
  s += al256((size_t)n * 4);
  int* hist = (int*)s;
  s += al256((size_t)256 * nt * sizeof(int));
  int* tot = (int*)s;
  const int passes = (hi_bit - lo_bit + 7) / 8;
  uint32_t* ks = keys;
  int32_t* vs = vals;
  uint32_t* kd = kalt;
  int32_t* vd = valt;
  for (int p = 0; p < passes; ++p) {
    const int shift = lo_bit + 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(nt), dim3(256), 0, st, ks, n, shift, hist);
    hipLaunchKernelGGL(k_rs_rowscan, dim3(256), dim3(256), 0, st, hist, nt, tot);
    hipLaunchKernelGGL(k_rs_scatter, dim3(nt), dim3(256), 0, st, ks, vs, n, shift, hist, tot, kd,
                       vd);
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != keys)
    hipLaunchKernelGGL(k_copy_pairs<uint32_t>, dim3((n + 255) / 256), dim3(256), 0, st, ks, vs, n,
                       keys, vals);
  FE_HIPCHK(hipGetLastError());
  return FE_OK;
}

int fe_radix_sort(uint32_t* keys, int32_t* vals, int n, int key_bits, void* scratch,
                  hipStream_t st) {
  if (key_bits < 1 || key_bits > 32)
    return fe_set_error(FE_ERR_ARG, "radix sort: key_bits %d out of range", key_bits);
  return fe_radix_sort_bits(keys, vals, n, 0, key_bits, scratch, st);
}
