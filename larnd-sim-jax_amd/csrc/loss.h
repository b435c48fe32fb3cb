// Event-blocked Gaussian-kernel sums for the hit-cloud MMD of the fit loss (losses_jax.py:13-39,
// 58-80) and the intensity of nll_loss (:329-361): out[i][k] = sum_j exp(-|x_i - y_j|^2 * inv)
// * c[j][k], k < 4, over the points y_j of x_i's event, in D = 3 or 4 dimensions.  Points of
// different events never form a kernel value (the reference's event offsets make those exactly
// zero).  Distances and the exponential are float32; terms are summed in float32 over runs of
// kRbfRun consecutive terms (at most 32: a relative rounding of ~32 * 2^-24 = 2e-6 of the run's sum
// in the worst case, ~3e-7 typically) and the runs in float64 (a float64 add per term would double
// the work per pair).  Fixed summation order: bitwise reproducible.
//
// The lanes of a wave walk one candidate span together (wave-uniform indices, so y_j and c_j are
// scalar loads shared by the wave): lanes are grouped by their span (range kernel) or by their event
// (grid kernel), and a wave whose lanes straddle two events walks each group in turn.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kRbfRun = 32;
constexpr float kLog2e = 1.4426950408889634f;

struct RbfAcc {
  double s0, s1, s2, s3;  // closed runs
  float f0, f1, f2, f3;   // the open run
  int run;                // terms in the open run
};

// Terms j = ja .. jb - 1 (wave-uniform bounds) into the open run; full runs close into float64.
template <int D>
__device__ __forceinline__ void rbf_span(const float4& p, const float4* __restrict__ y, const float4* __restrict__ c,
                                         int ja, int jb, float inv_l2, RbfAcc& a) {
  // the lanes walking a span share its run position: keep it (and so the loop) wave-uniform
  int run = __builtin_amdgcn_readfirstlane(a.run);
  while (ja < jb) {
    const int n = jb - ja < kRbfRun - run ? jb - ja : kRbfRun - run;
#pragma unroll 4
    for (int u = 0; u < n; ++u) {
      const float4 q = y[ja + u];
      const float4 w = c[ja + u];
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      float d2 = dx * dx + dy * dy + dz * dz;
      if (D == 4) {
        const float dw = p.w - q.w;
        d2 += dw * dw;
      }
      const float k = __builtin_amdgcn_exp2f(-d2 * inv_l2);  // exp(-d2 inv), inv_l2 = inv log2(e)
      a.f0 += k * w.x;
      a.f1 += k * w.y;
      a.f2 += k * w.z;
      a.f3 += k * w.w;
    }
    ja += n;
    run += n;
    if (run == kRbfRun) {
      a.s0 += (double)a.f0;
      a.s1 += (double)a.f1;
      a.s2 += (double)a.f2;
      a.s3 += (double)a.f3;
      a.f0 = a.f1 = a.f2 = a.f3 = 0.f;
      run = 0;
    }
  }
  a.run = run;
}

__device__ __forceinline__ void rbf_store(const RbfAcc& a, double* o) {
  o[0] = a.s0 + (double)a.f0;
  o[1] = a.s1 + (double)a.f1;
  o[2] = a.s2 + (double)a.f2;
  o[3] = a.s3 + (double)a.f3;
}

// Span form: point i sums over y[lo_i .. hi_i) (the caller's event ranges).
template <int D>
__global__ __launch_bounds__(256) void k_rbf_sums(const float4* __restrict__ x, int64_t nx, const int32_t* __restrict__ lo,
                                                  const int32_t* __restrict__ hi, const float4* __restrict__ y,
                                                  const float4* __restrict__ c, float inv, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < nx;
  const float4 p = valid ? x[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int mlo = valid ? lo[i] : 0, mhi = valid ? hi[i] : 0;
  RbfAcc a = {0.0, 0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0};
  bool pending = valid;
  for (uint64_t pm = __ballot(pending); pm != 0; pm = __ballot(pending)) {
    const int lead = __builtin_ctzll(pm);
    const int ja = __builtin_amdgcn_readlane(mlo, lead), jb = __builtin_amdgcn_readlane(mhi, lead);
    const bool mine = pending && mlo == ja && mhi == jb;
    if (mine) rbf_span<D>(p, y, c, ja, jb, inv * kLog2e, a);
    pending = pending && !mine;
  }
  if (valid) rbf_store(a, out + i * 4);
}

// ---- grid form: only the grid columns that can hold a non-zero kernel value are walked ----
//
// The candidates are sorted by key = ((event * Gx + cx) * Gy + cy) * Gz + cz, the cell of a point
// being clamp(floor((v - lo) / h), 0, G - 1) per axis (grid[] = lo xyz, h xyz, G xyz); the edge
// cells reach to infinity.  For the queries of one event in a wave (bounding box B), a cell column
// (cx, cy) is walked over the z cells within the cull radius rc of B, and skipped when its
// x-y distance from B exceeds rc.  Every skipped pair is at least rc apart in the first three
// coordinates, where exp(-d^2 inv) is exactly 0 in float32 (rc^2 inv > 106: e^-106 < 2^-152, below
// the smallest denormal): the sums are those of all pairs of the event, in another order.

__device__ __forceinline__ int rbf_cell(float v, float lo, float h, int G) {
  const float c = floorf((v - lo) / h);
  return (int)fminf(fmaxf(c, 0.f), (float)(G - 1));
}

// distance along one axis between [b0, b1] and cell c
__device__ __forceinline__ float rbf_gap(int c, float b0, float b1, float lo, float h, int G) {
  const float c0 = c == 0 ? -__builtin_inff() : lo + (float)c * h;
  const float c1 = c == G - 1 ? __builtin_inff() : lo + (float)(c + 1) * h;
  return fmaxf(fmaxf(c0 - b1, b0 - c1), 0.f);
}

__device__ __forceinline__ int rbf_lower_bound(const int64_t* __restrict__ k, int n, int64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
    if (k[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, kWave));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, kWave));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <int D>
__global__ __launch_bounds__(256) void k_rbf_grid(const float4* __restrict__ x, const int32_t* __restrict__ ev, int64_t nx,
                                                  const float4* __restrict__ y, const float4* __restrict__ c,
                                                  const int64_t* __restrict__ ykey, int32_t ny,
                                                  const float* __restrict__ grid, float inv, float rc,
                                                  double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool valid = i < nx;
  const float4 p = valid ? x[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const int my_ev = valid ? ev[i] : 0;
  const float lox = grid[0], loy = grid[1], loz = grid[2], hx = grid[3], hy = grid[4], hz = grid[5];
  const int Gx = (int)grid[6], Gy = (int)grid[7], Gz = (int)grid[8];
  const float inf = __builtin_inff();
  RbfAcc a = {0.0, 0.0, 0.0, 0.0, 0.f, 0.f, 0.f, 0.f, 0};
  bool pending = valid;
  for (uint64_t pm = __ballot(pending); pm != 0; pm = __ballot(pending)) {
    const int e = __builtin_amdgcn_readlane(my_ev, __builtin_ctzll(pm));
    const bool mine = pending && my_ev == e;
    const float bx0 = wave_min_f(mine ? p.x : inf), bx1 = wave_max_f(mine ? p.x : -inf);
    const float by0 = wave_min_f(mine ? p.y : inf), by1 = wave_max_f(mine ? p.y : -inf);
    const float bz0 = wave_min_f(mine ? p.z : inf), bz1 = wave_max_f(mine ? p.z : -inf);
    const int cxa = rbf_cell(bx0 - rc, lox, hx, Gx), cxb = rbf_cell(bx1 + rc, lox, hx, Gx);
    const int cya = rbf_cell(by0 - rc, loy, hy, Gy), cyb = rbf_cell(by1 + rc, loy, hy, Gy);
    const int ncy = cyb - cya + 1, ncol = cxb < cxa || cyb < cya ? 0 : (cxb - cxa + 1) * ncy;
    for (int b0 = 0; b0 < ncol; b0 += kWave) {
      // lane l finds the candidate span of column b0 + l
      int j0 = 0, j1 = 0;
      const int col = b0 + lane;
      if (col < ncol) {
        const int cx = cxa + col / ncy, cy = cya + col % ncy;
        const float gx = rbf_gap(cx, bx0, bx1, lox, hx, Gx), gy = rbf_gap(cy, by0, by1, loy, hy, Gy);
        const float r2 = rc * rc - gx * gx - gy * gy;
        if (r2 >= 0.f) {
          const float rz = sqrtf(r2);
          const int cza = rbf_cell(bz0 - rz, loz, hz, Gz), czb = rbf_cell(bz1 + rz, loz, hz, Gz);
          const int64_t kb = (((int64_t)e * Gx + cx) * Gy + cy) * Gz;
          j0 = rbf_lower_bound(ykey, ny, kb + cza);
          j1 = rbf_lower_bound(ykey, ny, kb + czb + 1);
        }
      }
      const int nc = ncol - b0 < kWave ? ncol - b0 : kWave;
      for (int cc = 0; cc < nc; ++cc) {
        const int ja = __builtin_amdgcn_readlane(j0, cc), jb = __builtin_amdgcn_readlane(j1, cc);
        if (mine) rbf_span<D>(p, y, c, ja, jb, inv * kLog2e, a);
      }
    }
    pending = pending && !mine;
  }
  if (valid) rbf_store(a, out + i * 4);
}

// Event-blocked nearest neighbour for the chamfer loss (losses_jax.py:195-229
// chamfer_distance_3d_og): for query i, the index of the nearest point among y[lo_i .. hi_i) by
// squared distance (dx^2 + dy^2) + dz^2 in float32, the first on ties as argmin takes it, and that
// distance.  The first n_origin queries also see the origin, a zero-weight padding point that
// comes after every real point (it wins only when strictly nearer; index -1).  No candidate:
// index -1, distance +inf.
__global__ __launch_bounds__(256) void k_nearest(const float4* __restrict__ x, int64_t nx, const int32_t* __restrict__ lo,
                                                 const int32_t* __restrict__ hi, const float4* __restrict__ y,
                                                 int64_t n_origin, int32_t* __restrict__ idx, float* __restrict__ dmin) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const float4 p = x[i];
  float best = __builtin_inff();
  int32_t bi = -1;
  for (int j = lo[i], j1 = hi[i]; j < j1; ++j) {
    const float4 q = y[j];
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float d = (dx * dx + dy * dy) + dz * dz;
    if (d < best) {
      best = d;
      bi = j;
    }
  }
  if (i < n_origin) {
    const float d = (p.x * p.x + p.y * p.y) + p.z * p.z;
    if (d < best) {
      best = d;
      bi = -1;
    }
  }
  idx[i] = bi;
  dmin[i] = best;
}

}  // namespace lsim
