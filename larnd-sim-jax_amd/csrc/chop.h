// Track chopping and batch padding (optimize/dataio.py:63-106 chop_tracks, :47-61
// remap_event_ids_to_local, :347-373 _invalidate_rows / pad_batch).
//
// Every input track row of the float32 (N, F) array becomes n = max(ceil(length / precision), 1)
// rows of the output.  Two launches after the step counts' exclusive scan:
//   k_chop_count: one thread per track -> n (int64), the scan gives each track its first row;
//   k_chop_fill:  one thread per output row builds the row in LDS (the block finds its first track
//                 by a 256-ary search of the offsets, stages the next 256 offsets in LDS and each
//                 thread searches those), then the block stores its rows as one contiguous run of
//                 floats (float4 when aligned), so HBM writes are fully coalesced.
// The rounding follows NumPy 2's rules on the reference's float32 array exactly:
//   length, direction, the step count and the per-step dE in float32; step positions
//   s + (k precision) d and s + (precision (k + 1)) d in float64; the last step's dE and dx in
//   float64; the midpoint 0.5 (start + end) in float32.  Contraction into FMAs is switched off in
//   these functions so every product and sum rounds on its own, as NumPy's do.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

namespace lsim {

constexpr int kChopMaxFields = 64;
constexpr int kChopRows = 256;  // output rows per block of k_chop_fill (= threads)
constexpr int kChopMaxCols = 8;

struct ChopLayout {
  int F;
  int s[3], e[3], m[3];
  int dE, dx, ev;
  int nzero, zero[kChopMaxCols];
  int nneg, neg[kChopMaxCols];
};

struct ChopTrack {
  float sx, sy, sz;     // start
  float dx, dy, dz;     // direction (float32)
  float length, den;    // |end - start| and length + 1e-10f
};

__device__ __forceinline__ ChopTrack chop_track(const ChopLayout& L, const float* __restrict__ row) {
#pragma clang fp contract(off)
  ChopTrack t;
  t.sx = row[L.s[0]];
  t.sy = row[L.s[1]];
  t.sz = row[L.s[2]];
  const float gx = row[L.e[0]] - t.sx, gy = row[L.e[1]] - t.sy, gz = row[L.e[2]] - t.sz;
  const float qx = gx * gx, qy = gy * gy, qz = gz * gz;
  // __builtin_sqrtf is the correctly rounded square root (__fsqrt_rn is the 1-ulp instruction)
  t.length = __builtin_sqrtf((qx + qy) + qz);
  t.den = t.length + 1e-10f;
  t.dx = __fdiv_rn(gx, t.den);
  t.dy = __fdiv_rn(gy, t.den);
  t.dz = __fdiv_rn(gz, t.den);
  return t;
}

// Steps of one track; 0 flags a length that is not finite or a count beyond 2^31 - 1.
__device__ __forceinline__ int64_t chop_steps(float length, float pf) {
  const float q = ceilf(__fdiv_rn(length, pf));
  if (!(q <= 2147483647.f)) return 0;  // NaN, inf or too many steps
  return q < 1.f ? 1 : (int64_t)q;
}

__global__ __launch_bounds__(256) void k_chop_count(ChopLayout L, const float* __restrict__ tracks, int64_t n, float pf,
                                                   int64_t* __restrict__ cnt, int64_t* __restrict__ info) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ChopTrack t = chop_track(L, tracks + i * L.F);
  int64_t k = chop_steps(t.length, pf);
  if (k == 0) {
    atomicAdd((unsigned long long*)(info + 1), 1ull);
    k = 1;
  }
  cnt[i] = k;
}

// Rows of an invalidated (padding or foreign-event) row: eventID -1, the zero columns 0 and the
// -1 columns -1 (dataio.py:347-355).
__device__ __forceinline__ void chop_invalidate(const ChopLayout& L, float* row) {
  if (L.ev >= 0) row[L.ev] = -1.f;
  for (int c = 0; c < L.nzero; ++c) row[L.zero[c]] = 0.f;
  for (int c = 0; c < L.nneg; ++c) row[L.neg[c]] = -1.f;
}

// Block-cooperative search for the last track whose first row is <= r0 (off[0] = 0 <= r0): each
// round the 256 threads test 256 evenly spaced offsets and the barrier count of true tests picks
// the bucket, so n tracks take ceil(log_256 n) dependent loads instead of log_2 n.
__device__ __forceinline__ int64_t chop_block_track(const int64_t* __restrict__ off, int64_t n, int64_t r0) {
  int64_t lo = 0, hi = n;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const int64_t stride = (hi - lo + kChopRows - 1) / kChopRows;
    const int64_t idx = lo + (int64_t)threadIdx.x * stride;
    const int pred = idx < hi && off[idx] <= r0;
    const int cnt = __syncthreads_count(pred);  // >= 1: off[lo] <= r0
    lo += (int64_t)(cnt - 1) * stride;
    hi = lo + stride < hi ? lo + stride : hi;
  }
  return lo;
}

template <bool V4>
__global__ __launch_bounds__(kChopRows) void k_chop_fill(ChopLayout L, const float* __restrict__ tracks, int64_t n,
                                                        double prec, const int64_t* __restrict__ off, int64_t n_steps,
                                                        int64_t n_rows, const int64_t* __restrict__ src_ev,
                                                        const int64_t* __restrict__ batch_ev, int64_t n_batch,
                                                        int64_t n_local, float* __restrict__ out,
                                                        int64_t* __restrict__ info) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) float tile[];
  __shared__ int64_t s_off[kChopRows + 1];
  const int F = L.F;
  const int64_t r0 = (int64_t)blockIdx.x * kChopRows;
  const int64_t r = r0 + threadIdx.x;
  float* row = tile + threadIdx.x * F;
  // The block's rows belong to tracks t0 .. t0 + m - 1 (m <= 256: every track has a row); their
  // first rows go to LDS and each thread finds its track there.
  int64_t t0 = 0;
  int m = 0;
  if (r0 < n_steps) {  // block-uniform
    t0 = chop_block_track(off, n, r0);
    const int64_t ti = t0 + threadIdx.x;
    s_off[threadIdx.x] = ti < n ? off[ti] : INT64_MAX;
    m = (int)(n - t0 < kChopRows ? n - t0 : kChopRows);
    __syncthreads();
  }
  if (r < n_steps) {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_off[mid] <= r) lo = mid;
      else hi = mid - 1;
    }
    const int64_t tr = t0 + lo;
    const float* in = tracks + tr * F;
    for (int c = 0; c < F; ++c) row[c] = in[c];
    const ChopTrack t = chop_track(L, in);
    const float pf = (float)prec;
    int64_t nst = chop_steps(t.length, pf);
    if (nst == 0) nst = 1;
    const int64_t k = r - s_off[lo];
    const bool last = k == nst - 1;
    row[L.dE] = __fdiv_rn(in[L.dE] * pf, t.den);
    const double kp = (double)k * prec, k1p = prec * (double)(k + 1);
    const float d3[3] = {t.dx, t.dy, t.dz}, s3[3] = {t.sx, t.sy, t.sz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      row[L.s[c]] = (float)((double)s3[c] + kp * (double)d3[c]);
      row[L.e[c]] = last ? in[L.e[c]] : (float)((double)s3[c] + k1p * (double)d3[c]);
    }
    row[L.dx] = pf;
    if (last) {
      const double nm1 = (double)(nst - 1);
      row[L.dE] = (float)((double)in[L.dE] * (1.0 - (prec * nm1) / (double)t.den));
      row[L.dx] = (float)((double)t.length - prec * nm1);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) row[L.m[c]] = 0.5f * (row[L.s[c]] + row[L.e[c]]);
    if (src_ev) {
      // searchsorted(batch_event_ids, source_event_id) with a membership check (dataio.py:47-61)
      const int64_t g = src_ev[tr];
      int64_t a = 0, b = n_batch;
      while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (batch_ev[mid] < g) a = mid + 1;
        else b = mid;
      }
      if (a == n_batch || batch_ev[a] != g) {
        if (k == 0) atomicAdd((unsigned long long*)(info + 2), 1ull);
        row[L.ev] = -1.f;
      } else {
        row[L.ev] = (float)a;
      }
    }
    // rows whose event is not one of the batch's local ids 0 .. n_local - 1 (dataio.py:368-371)
    if (n_local >= 0 && L.ev >= 0 && row[L.ev] >= 0.f && (double)row[L.ev] >= (double)n_local) chop_invalidate(L, row);
  } else if (r < n_rows) {
    for (int c = 0; c < F; ++c) row[c] = 0.f;
    chop_invalidate(L, row);
  }
  __syncthreads();
  const int64_t rows = n_rows - r0 < kChopRows ? n_rows - r0 : kChopRows;
  const int cnt = (int)rows * F;
  float* dst = out + r0 * F;
  if (V4) {  // out is 16-byte aligned and 256 * F floats are a whole number of float4
    const int c4 = cnt >> 2;
    for (int j = threadIdx.x; j < c4; j += kChopRows) reinterpret_cast<float4*>(dst)[j] = reinterpret_cast<const float4*>(tile)[j];
    for (int j = (c4 << 2) + threadIdx.x; j < cnt; j += kChopRows) dst[j] = tile[j];
  } else {
    for (int j = threadIdx.x; j < cnt; j += kChopRows) dst[j] = tile[j];
  }
}

}  // namespace lsim
