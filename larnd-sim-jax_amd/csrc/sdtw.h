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
// l-1 by a lane shuffle.  The values the tile needs from the tiles of earlier launches (the row
// above, the column to the left, the corner; in the backward the row below, the column to the
// right and their E) and the y values are loaded once per tile, one per lane, and handed out by
// shuffles, so no step waits on memory; the backward stages its tile of R in LDS.  R and E are
// kept in global memory (n x m float32 each).  The backward walks the tiles in reverse; each tile
// writes its rows' partial gradient sums, which one pass adds in a fixed order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kDtwTile = 64;  // = kWave
constexpr float kDtwBig = 1e10f;

__device__ __forceinline__ float dtw_cost(float x, float y) { return (-2.f * x * y + x * x) + y * y; }

// -g log(exp(-a/g) + exp(-b/g) + exp(-c/g)) = m - g log(1 + exp((m - o1)/g) + exp((m - o2)/g)), m the
// smallest and o1, o2 the other two (the smallest's own term is exactly 1).
__device__ __forceinline__ float dtw_softmin(float a, float b, float c, float g, float ig) {
  const float m = fminf(a, fminf(b, c));
  const float o1 = a == m ? b : a;
  const float o2 = a == m ? c : (b == m ? c : b);
  // logf(1 + s) rather than log1pf(s): s <= 2 and the absolute error g ulp(1) is what matters;
  // the library log1pf is several times longer on this wave's serial chain
  return m - g * logf(1.f + (expf((m - o1) * ig) + expf((m - o2) * ig)));
}

// Lane l receives lane l+1's value (0 on lane 63), resp. lane l-1's (0 on lane 0): DPP wave shifts,
// no LDS round trip on the wavefront's serial chain.  A DPP move reads 0 from lanes that are switched
// off, so it must run with the whole wave active: the empty asm operand barrier keeps the compiler
// from sinking the move into a divergent branch where only its consumers live (it emits no
// instruction).
__device__ __forceinline__ float dtw_from_next(float v) {
  int r = __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
  asm volatile("" : "+v"(r));
  return __int_as_float(r);
}
__device__ __forceinline__ float dtw_from_prev(float v) {
  int r = __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
  asm volatile("" : "+v"(r));
  return __int_as_float(r);
}

// v of lane src (the same lane for the whole wave).
__device__ __forceinline__ float dtw_lane(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src & (kWave - 1)));
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
  const int64_t j0 = (int64_t)J * kDtwTile;
  // tile inputs held one per lane: x_i, y_{j0+l}, the row above R[i0-1, j0+l], the column to the
  // left R[i, j0-1] and the corner R[i0-1, j0-1] (tiles of earlier launches, or the boundary)
  const float xi = i < n ? x[i] : 0.f;
  const float yv = j0 + l < m ? y[j0 + l] : 0.f;
  const float above = j0 + l < m ? dtw_r(R, m, (int64_t)I * kDtwTile - 1, j0 + l) : kDtwBig;
  const float leftc = i < n ? dtw_r(R, m, i, j0 - 1) : kDtwBig;
  const float corner = dtw_r(R, m, (int64_t)I * kDtwTile - 1, j0 - 1);
  const float ig = 1.f / g;
  const float lc_l = __shfl_up(leftc, 1, kWave);
  float v1 = kDtwBig, v2 = kDtwBig;  // this lane's values one and two steps ago
  for (int s = 0; s < 2 * kDtwTile - 1; ++s) {
    const int jl = s - l;
    const int64_t j = j0 + jl;
    const float up_l = dtw_from_prev(v1);
    const float dg_l = dtw_from_prev(v2);
    const float yj = __shfl(yv, jl & (kWave - 1), kWave);
    const float ab = dtw_lane(above, s);       // lane 0's column is s
    const float abm = dtw_lane(above, s - 1);
    // every lane evaluates (no branch before the lane exchanges of the next step: a lane shift
    // executed with lanes switched off would read 0 from them); only the store is predicated
    const bool act = jl >= 0 && jl < kDtwTile && i < n && j < m;
    const float left = jl == 0 ? leftc : v1;
    const float up = l == 0 ? ab : up_l;
    const float dg = jl == 0 ? (l == 0 ? corner : lc_l) : (l == 0 ? abm : dg_l);
    const float rv = dtw_cost(xi, yj) + dtw_softmin(dg, up, left, g, ig);
    const float r = act ? rv : kDtwBig;
    if (act) R[i * m + j] = r;
    v2 = v1;
    v1 = r;
  }
}

// E outside the table (and beyond the last cell) is 0.
__device__ __forceinline__ float dtw_e(const float* E, int64_t n, int64_t m, int64_t i, int64_t j) {
  return (i >= 0 && j >= 0 && i < n && j < m) ? E[i * m + j] : 0.f;
}

__device__ __forceinline__ float dtw_rr(const float* R, int64_t n, int64_t m, int64_t i, int64_t j) {
  return (i < n && j < m) ? R[i * m + j] : 0.f;
}

__global__ __launch_bounds__(kDtwTile) void k_sdtw_bwd(const float* __restrict__ x, int64_t n, const float* __restrict__ y,
                                                      int64_t m, float g, int k, const float* __restrict__ R,
                                                      float* __restrict__ E, float* __restrict__ part) {
#pragma clang fp contract(off)
  // this tile's R and the three successor factors exp((R[s] - D[s] - R[i,j]) / g) of every cell
  // (down, right, diagonal; 0 where the successor is outside the table), row-major 64 x 64: the
  // sweep's lane l reads column s - l of row l, bank (s - l) mod 64, conflict-free
  __shared__ float tile[kDtwTile * kDtwTile];
  __shared__ float w_d[kDtwTile * kDtwTile], w_r[kDtwTile * kDtwTile], w_dr[kDtwTile * kDtwTile];
  const int nTJ = (int)((m + kDtwTile - 1) / kDtwTile);
  const int I = max(0, k - (nTJ - 1)) + (int)blockIdx.x;
  const int J = k - I;
  const int nTI = (int)((n + kDtwTile - 1) / kDtwTile);
  if (I >= nTI || J < 0) return;
  const int l = threadIdx.x;
  const int64_t i0 = (int64_t)I * kDtwTile, j0 = (int64_t)J * kDtwTile;
  const int64_t i = i0 + l;
  for (int r = 0; r < kDtwTile; ++r) tile[r * kDtwTile + l] = dtw_rr(R, n, m, i0 + r, j0 + l);
  // successors outside the tile (tiles of earlier launches): the row below (lane = column) and the
  // column to the right (lane = row), and the corner
  const int64_t ib = i0 + kDtwTile, jr = j0 + kDtwTile;
  const float eb = dtw_e(E, n, m, ib, j0 + l), rb = dtw_rr(R, n, m, ib, j0 + l);
  const float er = dtw_e(E, n, m, i, jr), rr = dtw_rr(R, n, m, i, jr);
  const float ec = dtw_e(E, n, m, ib, jr), rc = dtw_rr(R, n, m, ib, jr);
  const float xv = i < n ? x[i] : 0.f;        // lane = row
  const float xb = ib < n ? x[ib] : 0.f;      // the row below
  const float yv = j0 + l < m ? y[j0 + l] : 0.f;  // lane = column
  const float yr = jr < m ? y[jr] : 0.f;
  const float ig = 1.f / g;
  __syncthreads();
  // factors, all lanes in parallel: lane c handles column c of every row r
  {
    const int c = l;
    const int64_t j = j0 + c;
    const float yn = dtw_from_next(yv);  // outside the select: a ternary evaluates one side only
    const float rb1 = dtw_from_next(rb);
    const float y1 = c == kDtwTile - 1 ? yr : yn;
    for (int r = 0; r < kDtwTile; ++r) {
      const int64_t ii = i0 + r;
      const float xr = dtw_lane(xv, r);
      const float x1 = r == kDtwTile - 1 ? xb : dtw_lane(xv, r + 1);
      const float rr_r = dtw_lane(rr, r), rr_r1 = dtw_lane(rr, r + 1);
      // branch-free (see the forward): invalid cells and successors are selected to 0
      const int r1 = r < kDtwTile - 1 ? r + 1 : r, c1 = c < kDtwTile - 1 ? c + 1 : c;
      const float rij = tile[r * kDtwTile + c];
      const float rs_d = r == kDtwTile - 1 ? rb : tile[r1 * kDtwTile + c];
      const float rs_r = c == kDtwTile - 1 ? rr_r : tile[r * kDtwTile + c1];
      const float rs_dr = r == kDtwTile - 1 ? (c == kDtwTile - 1 ? rc : rb1)
                                            : (c == kDtwTile - 1 ? rr_r1 : tile[r1 * kDtwTile + c1]);
      const bool cell = ii < n && j < m, down = ii + 1 < n, right = j + 1 < m;
      const float fd = cell && down ? expf((rs_d - dtw_cost(x1, yv) - rij) * ig) : 0.f;
      const float fr = cell && right ? expf((rs_r - dtw_cost(xr, y1) - rij) * ig) : 0.f;
      const float fdr = cell && down && right ? expf((rs_dr - dtw_cost(x1, y1) - rij) * ig) : 0.f;
      w_d[r * kDtwTile + c] = fd;
      w_r[r * kDtwTile + c] = fr;
      w_dr[r * kDtwTile + c] = fdr;
    }
  }
  __syncthreads();
  float e1 = 0.f, e2 = 0.f;  // this lane's E one and two steps ago (columns j+1, j+2)
  float gsum = 0.f;
  for (int s = 2 * kDtwTile - 2; s >= 0; --s) {
    const int jl = s - l;
    const int64_t j = j0 + jl;
    const float dn1 = dtw_from_next(e1);  // lane l+1: E[i+1, j]
    const float dn2 = dtw_from_next(e2);  // lane l+1: E[i+1, j+1]
    // the lane on the tile's last column (l = s - 63) takes lane l+1's right-column E; the last lane
    // (column s - 63) takes the row below at columns s - 63 and s - 62
    const float er_n = dtw_lane(er, s - 62);
    const float eb0 = dtw_lane(eb, s - 63), eb1 = dtw_lane(eb, s - 62);
    const float yj = __shfl(yv, jl & (kWave - 1), kWave);
    // branch-free up to the store (lane shifts must see every lane), the LDS index clamped
    const bool act = jl >= 0 && jl < kDtwTile && i < n && j < m;
    const bool last_row = l == kDtwTile - 1, last_col = jl == kDtwTile - 1;
    const int o = l * kDtwTile + (jl < 0 ? 0 : (jl > kDtwTile - 1 ? kDtwTile - 1 : jl));
    const float es_d = last_row ? eb0 : dn1;
    const float es_r = last_col ? er : e1;
    const float es_dr = last_row ? (last_col ? ec : eb1) : (last_col ? er_n : dn2);
    const float ev = (i == n - 1 && j == m - 1) ? 1.f : es_d * w_d[o] + es_r * w_r[o] + es_dr * w_dr[o];
    const float e = act ? ev : 0.f;
    if (act) E[i * m + j] = e;
    gsum += e * (2.f * xv - 2.f * yj);
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
