// Event-blocked Gaussian-kernel sums for the hit-cloud MMD of the fit loss (losses_jax.py:13-39,
// 58-80) and the intensity of nll_loss (:329-361): out[i][k] = sum_{j = lo_i .. hi_i - 1}
// exp(-|x_i - y_j|^2 * inv) * c[j][k], k < 4, over 4-D points (the 3-D callers set w = 0).
// The caller sorts both clouds by event and gives each point of x the index range of its event
// in y, so no kernel value between events is ever formed (the reference's 1e5 cm event offset
// makes those exactly zero).  Threads of a wave almost always share an event, so the y point and
// its coefficients are wave-uniform loads; distances and the exponential are float32.  Terms are
// summed in float32 over runs of kRbfRun consecutive y points (at most 32 terms: a relative
// rounding of ~32 * 2^-24 = 2e-6 of the run's sum in the worst case, ~3e-7 typically) and the
// runs in float64; a float64 add per term would double the work per pair.  Fixed summation
// order: bitwise reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kRbfRun = 32;

__global__ __launch_bounds__(256) void k_rbf_sums(const float4* __restrict__ x, int64_t nx, const int32_t* __restrict__ lo,
                                                  const int32_t* __restrict__ hi, const float4* __restrict__ y,
                                                  const float4* __restrict__ c, float inv, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const float4 p = x[i];
  const int j0 = lo[i], j1 = hi[i];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int jb = j0; jb < j1; jb += kRbfRun) {
    const int je = jb + kRbfRun < j1 ? jb + kRbfRun : j1;
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, f3 = 0.f;
    for (int j = jb; j < je; ++j) {
      const float4 q = y[j];
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z, dw = p.w - q.w;
      const float k = __expf(-(dx * dx + dy * dy + dz * dz + dw * dw) * inv);
      const float4 w = c[j];
      f0 += k * w.x;
      f1 += k * w.y;
      f2 += k * w.z;
      f3 += k * w.w;
    }
    s0 += (double)f0;
    s1 += (double)f1;
    s2 += (double)f2;
    s3 += (double)f3;
  }
  double* o = out + i * 4;
  o[0] = s0;
  o[1] = s1;
  o[2] = s2;
  o[3] = s3;
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
