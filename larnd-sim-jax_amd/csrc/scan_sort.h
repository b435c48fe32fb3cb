// Deterministic device-wide exclusive scan and stable LSD radix sort (key, value) pairs.
// Hand-written instead of a library call so that every launch is a plain kernel (no memset or
// host synchronisation inside), which keeps the forward capturable into a hipGraph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d) {
  return __shfl_up(v, d, kWave);
}

// Block-wide exclusive scan of one value per thread (256 threads).  Returns the exclusive
// prefix; `total` receives the block sum.  Uses `sh` (>= 4 entries of T) as scratch.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* sh, T& total) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  T inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T u = shfl_up_t(inc, d);
    if (lane >= d) inc += u;
  }
  // the exclusive prefix is the previous lane's inclusive one (not inc - v: that would round, and
  // a NaN or inf of this thread's own values would reach its prefix, which a cumulative sum's
  // earlier entries never see)
  T exc = shfl_up_t(inc, 1);
  if (lane == 0) exc = 0;
  if (lane == kWave - 1) sh[wid] = inc;
  __syncthreads();
  const int nw = blockDim.x / kWave;
  T base = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wid) base += sh[w];
    tot += sh[w];
  }
  total = tot;
  __syncthreads();
  return base + exc;
}

// 1) per-tile sums
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in, int64_t n, T* __restrict__ tile_sums) {
  __shared__ T sh[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  T tot;
  (void)block_exclusive_scan<T>(s, sh, tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// 2) exclusive scan of the tile sums in place (one block, loops over chunks); total -> *total
template <typename T, typename TT>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(T* __restrict__ sums, int64_t ntiles, TT* __restrict__ total) {
  __shared__ T sh[kScanThreads / kWave];
  T carry = 0;
  for (int64_t c0 = 0; c0 < ntiles; c0 += kScanThreads) {
    const int64_t i = c0 + threadIdx.x;
    T v = i < ntiles ? sums[i] : T(0);
    T tot;
    T ex = block_exclusive_scan<T>(v, sh, tot);
    if (i < ntiles) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = (TT)carry;
}

// 3) scan within each tile plus the tile offset
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                                            const T* __restrict__ tile_off) {
  __shared__ T sh[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : T(0);
    s += v[k];
  }
  T tot;
  T ex = block_exclusive_scan<T>(s, sh, tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

inline int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

// Small inputs (up to 16 values per thread of one 1024-thread workgroup): the whole scan in one
// launch instead of three (the radix sort's digit histograms and the per-row counts are this
// size, and each launch costs a few microseconds on its own).
constexpr int kScanSmallThreads = 1024;
constexpr int kScanSmallItems = 16;
constexpr int64_t kScanSmallMax = (int64_t)kScanSmallThreads * kScanSmallItems;

template <typename T, typename TT>
__global__ __launch_bounds__(kScanSmallThreads) void k_scan_small(const T* in, T* out, int64_t n,
                                                                 TT* __restrict__ total) {  // in may equal out
  __shared__ T sh[kScanSmallThreads / kWave];
  const int per = (int)((n + kScanSmallThreads - 1) / kScanSmallThreads);  // <= kScanSmallItems
  const int64_t base = (int64_t)threadIdx.x * per;
  T v[kScanSmallItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanSmallItems; ++k) {
    v[k] = (k < per && base + k < n) ? in[base + k] : T(0);
    s += v[k];
  }
  T tot;
  T ex = block_exclusive_scan<T>(s, sh, tot);
#pragma unroll
  for (int k = 0; k < kScanSmallItems; ++k) {
    if (k < per && base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0 && total) *total = (TT)tot;
}

// Exclusive scan of n values (in may equal out).  tmp holds scan_tiles(n) values of T.
template <typename T, typename TT>
inline hipError_t exclusive_scan(const T* in, T* out, int64_t n, T* tmp, TT* total, hipStream_t st) {
  const int64_t nt = scan_tiles(n);
  if (nt == 0) {
    return hipSuccess;
  }
  if (n <= kScanSmallMax) {
    hipLaunchKernelGGL((k_scan_small<T, TT>), dim3(1), dim3(kScanSmallThreads), 0, st, in, out, n, total);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_scan_reduce<T>), dim3((unsigned)nt), dim3(kScanThreads), 0, st, in, n, tmp);
  hipLaunchKernelGGL((k_scan_tiles<T, TT>), dim3(1), dim3(kScanThreads), 0, st, tmp, nt, total);
  hipLaunchKernelGGL((k_scan_apply<T>), dim3((unsigned)nt), dim3(kScanThreads), 0, st, in, out, n, tmp);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits, tiles of 2048 items (256 threads x 8 rounds)
// ------------------------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kRadix = 256;

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                             int ntiles, int* __restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t i = base + r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter with stable ranks: items are ranked in index order (round by round, lane order within
// a round), so equal digits keep their input order.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, int64_t n, int shift,
                                                                int ntiles, const int* __restrict__ offs) {
  constexpr int NW = kSortThreads / kWave;
  __shared__ int base[kRadix];
  __shared__ int wcnt[NW][kRadix];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  base[threadIdx.x] = offs[(int64_t)threadIdx.x * ntiles + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
    for (int w = 0; w < NW; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = tile0 + r * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0, val = 0;
    int dg = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      dg = (key >> shift) & (kRadix - 1);
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dg >> b) & 1;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const int rank_w = __popcll(peers & lt);
    if (valid && rank_w == 0) wcnt[wid][dg] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = base[dg] + rank_w;
      for (int w = 0; w < wid; ++w) pos += wcnt[w][dg];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) add += wcnt[w][threadIdx.x];
    base[threadIdx.x] += add;
    __syncthreads();
  }
}

inline int sort_tiles(int64_t n) { return (int)((n + kSortTile - 1) / kSortTile); }

// Bytes of scratch the sort needs for n items (histogram + scan tiles).
inline size_t radix_sort_scratch(int64_t n) {
  const int64_t nt = sort_tiles(n);
  const int64_t hist = nt * kRadix;
  return (size_t)(hist + scan_tiles(hist) + 4) * sizeof(int);
}

// Sort (keys, vals) by the low `bits` bits of the key.  Ping-pongs between (k0, v0) and
// (k1, v1); returns the buffer index (0 or 1) holding the result.
inline int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int bits, void* scratch,
                            hipStream_t st, hipError_t* err) {
  const int nt = sort_tiles(n);
  int* hist = (int*)scratch;
  int* tmp = hist + (int64_t)nt * kRadix;
  int cur = 0;
  *err = hipSuccess;
  if (n == 0) return 0;
  for (int shift = 0; shift < bits; shift += 8) {
    uint32_t* ki = cur ? k1 : k0;
    uint32_t* vi = cur ? v1 : v0;
    uint32_t* ko = cur ? k0 : k1;
    uint32_t* vo = cur ? v0 : v1;
    hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(kSortThreads), 0, st, ki, n, shift, nt, hist);
    *err = exclusive_scan<int, int>(hist, hist, (int64_t)nt * kRadix, tmp, (int*)nullptr, st);
    if (*err != hipSuccess) return cur;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(kSortThreads), 0, st, ki, vi, ko, vo, n, shift, nt, hist);
    cur ^= 1;
  }
  *err = hipGetLastError();
  return cur;
}

}  // namespace lsim
