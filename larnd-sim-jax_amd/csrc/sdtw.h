// Soft dynamic time warping (src/larndsim/softdtw_jax.py:205-275, the ott-jax SoftDTW with the
// squared-Euclidean ground cost) and its gradient, for the sdtw fit losses (losses_jax.py:309-327).
//
// R[i,j] = D[i,j] + softmin_g(R[i-1,j-1], R[i-1,j], R[i,j-1]),  softmin_g(a) = -g log sum exp(-a/g),
// with R[-1,-1] = 0 and every other outside value 1e10 (the reference's padding), D[i,j] =
// (-2 x_i y_j + x_i^2) + y_j^2; the value is R[n-1,m-1].  The gradient of the value with respect to
// D is E, E[n-1,m-1] = 1 and E[i,j] = sum over the successors s of (i,j) of
// E[s] exp((R[s] - D[s] - R[i,j]) / g); d value / d x_i = sum_j E[i,j] (2 x_i - 2 y_j).
//
// Wavefront over 64 x 64 tiles: one launch per tile anti-diagonal, one wave per tile, lane l owns
// row l of the tile and walks the tile's 127 inner anti-diagonals; the row above comes from lane
// l-1 by a lane shuffle, the tiles above / to the left (finished in earlier launches) from global
// memory.  R and E are kept in global memory (n x m float32 each).  The backward walks the same
// tiles in reverse; each tile writes its rows' partial gradient sums, which one pass adds in a
// fixed order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kDtwTile = 64;  // = kWave
constexpr float kDtwBig = 1e10f;

__device__ __forceinline__ float dtw_cost(float x, float y) { return (-2.f * x * y + x * x) + y * y; }

__device__ __forceinline__ float dtw_softmin(float a, float b, float c, float g) {
  const float m = fminf(a, fminf(b, c));
  return m - g * logf(expf(-(a - m) / g) + expf(-(b - m) / g) + expf(-(c - m) / g));
}

// R outside the table: 0 at (-1,-1), 1e10 elsewhere.
__device__ __forceinline__ float dtw_r(const float* R, int64_t m, int64_t i, int64_t j) {
  if (i < 0 || j < 0) return (i < 0 && j < 0) ? 0.f : kDtwBig;
  return R[i * m + j];
}

__global__ __launch_bounds__(kDtwTile) void k_sdtw_fwd(const float* __restrict__ x, int64_t n, const float* __restrict__ y,
                                                      int64_t m, float g, int k, float* __restrict__ R) {
#pragma clang fp contract(off)
  const int nTJ = (int)((m + kDtwTile - 1) / kDtwTile);
  const int I = max(0, k - (nTJ - 1)) + (int)blockIdx.x;
  const int J = k - I;
  const int nTI = (int)((n + kDtwTile - 1) / kDtwTile);
  if (I >= nTI || J < 0) return;  // grid covers the longest diagonal; the rest have no tile
  const int l = threadIdx.x;
  const int64_t i = (int64_t)I * kDtwTile + l;
  const float xi = i < n ? x[i] : 0.f;
  float v1 = kDtwBig, v2 = kDtwBig;  // this lane's values one and two steps ago
  for (int s = 0; s < 2 * kDtwTile - 1; ++s) {
    const int jl = s - l;
    const int64_t j = (int64_t)J * kDtwTile + jl;
    const float up_l = __shfl_up(v1, 1, kWave);
    const float dg_l = __shfl_up(v2, 1, kWave);
    float r = kDtwBig;
    const bool act = jl >= 0 && jl < kDtwTile && i < n && j < m;
    if (act) {
      const float left = jl == 0 ? dtw_r(R, m, i, j - 1) : v1;
      const float up = l == 0 ? dtw_r(R, m, i - 1, j) : up_l;
      const float dg = (l == 0 || jl == 0) ? dtw_r(R, m, i - 1, j - 1) : dg_l;
      r = dtw_cost(xi, y[j]) + dtw_softmin(dg, up, left, g);
      R[i * m + j] = r;
    }
    v2 = v1;
    v1 = r;
  }
}

// E outside the table (and beyond the last cell) is 0.
__device__ __forceinline__ float dtw_e(const float* E, int64_t n, int64_t m, int64_t i, int64_t j) {
  return (i < n && j < m) ? E[i * m + j] : 0.f;
}

__global__ __launch_bounds__(kDtwTile) void k_sdtw_bwd(const float* __restrict__ x, int64_t n, const float* __restrict__ y,
                                                      int64_t m, float g, int k, const float* __restrict__ R,
                                                      float* __restrict__ E, float* __restrict__ part) {
#pragma clang fp contract(off)
  const int nTJ = (int)((m + kDtwTile - 1) / kDtwTile);
  const int I = max(0, k - (nTJ - 1)) + (int)blockIdx.x;
  const int J = k - I;
  const int nTI = (int)((n + kDtwTile - 1) / kDtwTile);
  if (I >= nTI || J < 0) return;
  const int l = threadIdx.x;
  const int64_t i = (int64_t)I * kDtwTile + l;
  const float xi = i < n ? x[i] : 0.f;
  const float xn = i + 1 < n ? x[i + 1] : 0.f;
  float e1 = 0.f, e2 = 0.f;  // this lane's E one and two steps ago (columns j+1, j+2)
  float gsum = 0.f;
  for (int s = 2 * kDtwTile - 2; s >= 0; --s) {
    const int jl = s - l;
    const int64_t j = (int64_t)J * kDtwTile + jl;
    const float dn1 = __shfl_down(e1, 1, kWave);  // lane l+1: E[i+1, j]
    const float dn2 = __shfl_down(e2, 1, kWave);  // lane l+1: E[i+1, j+1]
    float e = 0.f;
    const bool act = jl >= 0 && jl < kDtwTile && i < n && j < m;
    if (act) {
      const float rij = R[i * m + j];
      if (i == n - 1 && j == m - 1) {
        e = 1.f;
      } else {
        // successor (i+1, j)
        const float e_d = l == kDtwTile - 1 ? dtw_e(E, n, m, i + 1, j) : (i + 1 < n ? dn1 : 0.f);
        // successor (i, j+1)
        const float e_r = jl == kDtwTile - 1 ? dtw_e(E, n, m, i, j + 1) : (j + 1 < m ? e1 : 0.f);
        // successor (i+1, j+1)
        const float e_dr = (l == kDtwTile - 1 || jl == kDtwTile - 1) ? dtw_e(E, n, m, i + 1, j + 1)
                                                                    : ((i + 1 < n && j + 1 < m) ? dn2 : 0.f);
        if (e_d != 0.f) e += e_d * expf((R[(i + 1) * m + j] - dtw_cost(xn, y[j]) - rij) / g);
        if (e_r != 0.f) e += e_r * expf((R[i * m + j + 1] - dtw_cost(xi, y[j + 1]) - rij) / g);
        if (e_dr != 0.f) e += e_dr * expf((R[(i + 1) * m + j + 1] - dtw_cost(xn, y[j + 1]) - rij) / g);
      }
      E[i * m + j] = e;
      gsum += e * (2.f * xi - 2.f * y[j]);
    }
    e2 = e1;
    e1 = e;
  }
  if (i < n) part[i * nTJ + J] = gsum;
}

// grad[i] = sum over column tiles of part[i][J], in order.
__global__ __launch_bounds__(256) void k_sdtw_grad(const float* __restrict__ part, int64_t n, int nTJ,
                                                  double* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int J = 0; J < nTJ; ++J) s += (double)part[i * nTJ + J];
  grad[i] = s;
}

}  // namespace lsim
