// Probabilistic self-trigger readout (fee_jax.py:334-461 _find_one_hit_step and
// get_adc_values_average_noise_vmap): a beam search of V charge-reset paths per pixel row over
// the tick grid, with Gaussian reset noise sigma, giving per hit index s and tick t the log
// probability that hit s starts at t and its expected charge.
//
// Per row, with q_sum the cumulative charge (T ticks), c_v and p_v the reset charge and log
// probability of path v, sh(t) = min(t + interval + 1, T - 1) and z(x) = (x - threshold) / sigma:
//   log_guess[v,t] = logdiffPhi(z(CM[t+1] - c_v), z(CM[t] - c_v)),  CM = cumulative max of q_sum
//                    (cummax(q_sum - c) = cummax(q_sum) - c, so one CM serves every path);
//   log_event[v,t] = min(logdiffPhi(z(q_sum[sh] - c_v), z(q_sum[t] - c_v)), log_guess), then
//                    soft-clamped at -1000 and soft-masked by the expected charge
//                    esp[t] = q_sum[sh] + threshold - (q_sum[t+1] + q_sum[t]) / 2;
//   out_logp[s,t]  = logsumexp_v(log_event + p_v),  out_charge[s,t] = esp[t];
//   total[t]       = logsumexp_v(log_guess + p_v);
//   select[t]      = LF[t] + total[t], LF[t] = log Phi(z(RCM[e(t)] - q_sum[sh(t)+1])), RCM the
//                    reverse cumulative max, e(t) = min(sh(t) + interval + 1, T - 1) (the path
//                    charge cancels in this term, so it is one value per tick);
//   the V ticks of largest select (lower tick first on ties) become the next paths:
//   p'_k = total[t_k], c'_k = q_sum[min(sh(t_k) + 1, T - 1)].
// The reference stops all rows at once when no row has logsumexp(p') > log(stop) any more; the
// step launches read and set that flag in device memory, and stopped steps write -1000 / 0.
// Rows are independent: one workgroup per row and step, ticks in contiguous chunks per thread,
// the row's tables in LDS, the top-V selection by V rounds of a block arg-max.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsim {

constexpr int kPfThreads = 256;
constexpr int kPfMaxTicks = 2048;
constexpr int kPfMaxPaths = 32;
constexpr int kPfChunk = (kPfMaxTicks + kPfThreads - 1) / kPfThreads;  // ticks per thread

struct PfDev {
  int T;          // ticks per row
  int V;          // paths kept (fee_paths_scaling)
  int S;          // hit indices (MAX_ADC_VALUES)
  int interval;   // ADC hold interval in ticks
  float theta;    // discrimination threshold (e-)
  float zs;       // 1 / reset-noise sigma
  float ts;       // t_sampling
  float log_stop; // log of the stop threshold
};

// log Phi(x) in float32 from one scaled complementary error function, u = |x| / sqrt 2:
// x < 0: -u^2 + log(erfcx(u) / 2) (accurate far into the tail); x >= 0: log1p(-exp(-u^2) erfcx(u) / 2).
// No branch on the sign around the expensive part, so lanes of either sign share one erfcx.
__device__ __forceinline__ float pf_log_ndtr(float x) {
  const float u = fabsf(x) * 0.70710678118654752f;
  const float ex = erfcxf(u);
  const float u2 = u * u;
  return x < 0.f ? -u2 + logf(0.5f * ex) : log1pf(-0.5f * expf(-u2) * ex);
}

// softplus(x) = max(x, 0) + log1p(exp(-|x|)); beyond |x| = 30 the correction is below float32
// resolution (x itself, or exp(x) = log1p(exp(x)) to 1e-13), so the transcendental calls are skipped.
__device__ __forceinline__ float pf_softplus(float x) {
  if (x > 30.f) return x;
  if (x < -30.f) return expf(x);
  return fmaxf(x, 0.f) + log1pf(expf(-fabsf(x)));
}

// lo + softplus((x - lo) s) / s (fee_jax.py:13-20)
__device__ __forceinline__ float pf_soft_max(float x, float lo, float s) { return lo + pf_softplus((x - lo) * s) / s; }

// sigmoid(c s) t + (1 - sigmoid(c s)) f (fee_jax.py:22-30)
__device__ __forceinline__ float pf_soft_where(float c, float t, float f, float s) {
  const float x = c * s;
  // sigmoid(x): 1 in float32 above 30 (1 + exp(-30) rounds to 1); exp(x) below -30 (to 1e-13)
  const float w = x > 30.f ? 1.f : (x < -30.f ? expf(x) : 1.f / (1.f + expf(-x)));
  return w * t + (1.f - w) * f;
}

// log(Phi(a) - Phi(b)) with the reference's smooth clamps (fee_jax.py:32-54), from la = log Phi(a)
// and lb = log Phi(b).
__device__ __forceinline__ float pf_log_diff(float a, float b, float la, float lb) {
  const float sd = a > b ? lb - la : -1.f;
  float ne = -expm1f(sd);
  ne = pf_soft_max(ne, 1e-30f, 1e10f);
  float lt = logf(ne);
  lt = pf_soft_max(lt, -100.f, 10.f);
  return pf_soft_where(a - b, la + lt, -1000.f, 1000.f);
}

// Running logsumexp: (m, sum) with sum = sum_i exp(x_i - m).
__device__ __forceinline__ void pf_lse_add(float& m, float& sum, float x) {
  if (x > m) {
    sum = sum * expf(m - x) + 1.f;
    m = x;
  } else {
    sum += expf(x - m);
  }
}

// Block-wide inclusive scan (sum, double) of one value per thread; returns the inclusive prefix.
__device__ __forceinline__ double pf_block_scan(double v, double* sh) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  double inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double u = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += u;
  }
  if (lane == kWave - 1) sh[wid] = inc;
  __syncthreads();
  double base = 0.0;
  for (int w = 0; w < wid; ++w) base += sh[w];
  __syncthreads();
  return base + inc;
}

// Per row: q_sum, its cumulative max, LF and esp (all independent of the step).
// tab layout per row: [q_sum T][CM T][LF T-1 (+1 pad)][esp T-1 (+1 pad)].
__global__ __launch_bounds__(kPfThreads) void k_pf_prep(PfDev d, const float* __restrict__ wfs, int64_t P,
                                                       float* __restrict__ tab) {
  __shared__ float s_q[kPfMaxTicks];
  __shared__ float s_r[kPfMaxTicks];
  __shared__ double s_w[kPfThreads / kWave];
  __shared__ float s_wm[kPfThreads / kWave];
  const int64_t row = blockIdx.x;
  const int T = d.T;
  const float* w = wfs + row * T;
  const int t0 = threadIdx.x * kPfChunk;
  // cumulative sum of q = wfs * t_sampling (double prefix, rounded once)
  double loc[kPfChunk];
  double run = 0.0;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    const double q = t < T ? (double)(w[t] * d.ts) : 0.0;
    run += q;
    loc[k] = run;
  }
  const double base = pf_block_scan(run, s_w) - run;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k)
    if (t0 + k < T) s_q[t0 + k] = (float)(base + loc[k]);
  __syncthreads();
  float* qs = tab + row * 4 * (int64_t)T;
  float* cm = qs + T;
  float* lf = cm + T;
  float* esp = lf + T;
  // forward and reverse cumulative max: chunk maxima, exclusive prefix / suffix maxima over
  // threads (wave shuffles, then the wave totals in LDS), then a walk over the chunk
  const float ninf = -__builtin_inff();
  float m = ninf;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k)
    if (t0 + k < T) m = fmaxf(m, s_q[t0 + k]);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  float inc = m;
#pragma unroll
  for (int dd = 1; dd < kWave; dd <<= 1) {
    const float u = __shfl_up(inc, dd, kWave);
    if (lane >= dd) inc = fmaxf(inc, u);
  }
  const float up1 = __shfl_up(inc, 1, kWave);
  if (lane == kWave - 1) s_wm[wid] = inc;
  __syncthreads();
  float excl_f = lane > 0 ? up1 : ninf;
  for (int ww = 0; ww < wid; ++ww) excl_f = fmaxf(excl_f, s_wm[ww]);
  __syncthreads();
  float incr = m;
#pragma unroll
  for (int dd = 1; dd < kWave; dd <<= 1) {
    const float u = __shfl_down(incr, dd, kWave);
    if (lane + dd < kWave) incr = fmaxf(incr, u);
  }
  const float dn1 = __shfl_down(incr, 1, kWave);
  if (lane == 0) s_wm[wid] = incr;
  __syncthreads();
  float excl_r = lane + 1 < kWave ? dn1 : ninf;
  for (int ww = wid + 1; ww < kPfThreads / kWave; ++ww) excl_r = fmaxf(excl_r, s_wm[ww]);
  float cf = excl_f;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < T) {
      cf = fmaxf(cf, s_q[t]);
      cm[t] = cf;
    }
  }
  float cr = excl_r;
#pragma unroll
  for (int k = kPfChunk - 1; k >= 0; --k) {
    const int t = t0 + k;
    if (t < T) {
      cr = fmaxf(cr, s_q[t]);
      s_r[t] = cr;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < T) qs[t] = s_q[t];
    if (t < T - 1) {
      const int sh = min(t + d.interval + 1, T - 1);
      const int nx = min(sh + 1, T - 1);
      const int e = min(sh + d.interval + 1, T - 1);
      lf[t] = pf_log_ndtr((s_r[e] - s_q[nx] - d.theta) * d.zs);
      esp[t] = s_q[sh] + d.theta - 0.5f * (s_q[t + 1] + s_q[t]);
    }
  }
}

// One beam-search step for every row.  The paths are kept per step (the tape of the adjoint):
// tc[s][row][v], tp[s][row][v] = reset charge and log probability of path v entering step s
// (s = 0 .. S; slot 0 holds the start paths), tk[s][row][k] = the tick of the k-th selected path of
// step s.  flags[s] != 0: step s runs; a row whose new paths keep probability sets flags[s + 1].
__global__ __launch_bounds__(kPfThreads) void k_pf_step(PfDev d, int s, const float* __restrict__ tab, int64_t P,
                                                       float* __restrict__ tc, float* __restrict__ tp,
                                                       int* __restrict__ tk, int* __restrict__ flags,
                                                       float* __restrict__ out_logp, float* __restrict__ out_q) {
  __shared__ float s_q[kPfMaxTicks];
  __shared__ float s_cm[kPfMaxTicks];
  // per-path tables log Phi(z(q_sum - c)) and log Phi(z(CM - c)); after the path loop the same
  // memory holds the per-tick totals and selection scores
  __shared__ float s_lq[kPfMaxTicks];
  __shared__ float s_lc[kPfMaxTicks];
  float* const s_tot = s_lq;
  float* const s_sel = s_lc;
  __shared__ float s_c[kPfMaxPaths], s_p[kPfMaxPaths];
  __shared__ float s_bv[kPfThreads / kWave];
  __shared__ int s_bi[kPfThreads / kWave];
  __shared__ int s_top[kPfMaxPaths];
  const int64_t row = blockIdx.x;
  const int T = d.T, V = d.V, n = T - 1;
  float* ol = out_logp + (row * d.S + s) * (int64_t)n;
  float* oq = out_q + (row * d.S + s) * (int64_t)n;
  if (!flags[s]) {  // every row has stopped: the reference's inactive branch
    for (int t = threadIdx.x; t < n; t += kPfThreads) {
      ol[t] = -1000.f;
      oq[t] = 0.f;
    }
    return;
  }
  const float* qs = tab + row * 4 * (int64_t)T;
  const float* cm = qs + T;
  const float* lf = cm + T;
  const float* esp = lf + T;
  for (int t = threadIdx.x; t < T; t += kPfThreads) {
    s_q[t] = qs[t];
    s_cm[t] = cm[t];
  }
  float* in_c = tc + ((int64_t)s * P + row) * V;
  float* in_p = tp + ((int64_t)s * P + row) * V;
  float* st_c = tc + ((int64_t)(s + 1) * P + row) * V;
  float* st_p = tp + ((int64_t)(s + 1) * P + row) * V;
  __shared__ int s_nv;
  if (threadIdx.x < V) {
    s_c[threadIdx.x] = s == 0 ? 0.f : in_c[threadIdx.x];
    s_p[threadIdx.x] = s == 0 ? (threadIdx.x == 0 ? 0.f : -1000.f) : in_p[threadIdx.x];
    if (s == 0) {  // the start paths, for the adjoint
      in_c[threadIdx.x] = s_c[threadIdx.x];
      in_p[threadIdx.x] = s_p[threadIdx.x];
    }
  }
  __syncthreads();
  // Paths with the same reset charge have the same log_event and log_guess at every tick, so their
  // terms differ only by p_v: merge them into one path with p = logsumexp of theirs (in path
  // order).  At step 0 all V start paths share charge 0 and collapse into one.
  if (threadIdx.x == 0) {
    int nv = 0;
    for (int v = 0; v < V; ++v) {
      int u = 0;
      while (u < nv && s_c[u] != s_c[v]) ++u;
      if (u == nv) {
        s_c[nv] = s_c[v];
        s_p[nv] = s_p[v];
        ++nv;
      } else {
        const float a = s_p[u], b = s_p[v], m = fmaxf(a, b);
        s_p[u] = m + log1pf(expf(fminf(a, b) - m));
      }
    }
    s_nv = nv;
  }
  __syncthreads();
  const int NV = s_nv;
  const int t0 = threadIdx.x * kPfChunk;
  float hmax[kPfChunk], hsum[kPfChunk], gmax[kPfChunk], gsum[kPfChunk];
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    hmax[k] = -__builtin_inff();
    gmax[k] = -__builtin_inff();
    hsum[k] = 0.f;
    gsum[k] = 0.f;
  }
  float espk[kPfChunk];
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) espk[k] = t0 + k < n ? esp[t0 + k] : 0.f;
  // one pass over the (merged) paths with running logsumexp accumulators (max and shifted sum);
  // each path's log Phi values are tabulated once per tick (each is used by two ticks)
  for (int v = 0; v < NV; ++v) {
    const float c = s_c[v], pv = s_p[v];
    for (int t = threadIdx.x; t < T; t += kPfThreads) {
      s_lq[t] = pf_log_ndtr((s_q[t] - c - d.theta) * d.zs);
      s_lc[t] = pf_log_ndtr((s_cm[t] - c - d.theta) * d.zs);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPfChunk; ++k) {
      const int t = t0 + k;
      if (t < n) {
        const float a1 = (s_cm[t + 1] - c - d.theta) * d.zs, b1 = (s_cm[t] - c - d.theta) * d.zs;
        const float lg = pf_log_diff(a1, b1, s_lc[t + 1], s_lc[t]);
        const int sh = min(t + d.interval + 1, T - 1);
        const float a2 = (s_q[sh] - c - d.theta) * d.zs, b2 = (s_q[t] - c - d.theta) * d.zs;
        float le = pf_log_diff(a2, b2, s_lq[sh], s_lq[t]);
        le = fminf(le, lg);
        le = pf_soft_max(le, -1000.f, 1.f);
        le = pf_soft_where(espk[k] - d.theta, le, -1000.f, 10.f);
        pf_lse_add(hmax[k], hsum[k], le + pv);
        pf_lse_add(gmax[k], gsum[k], lg + pv);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < n) {
      ol[t] = hmax[k] + logf(hsum[k]);
      oq[t] = espk[k];
      const float tot = gmax[k] + logf(gsum[k]);
      s_tot[t] = tot;
      s_sel[t] = lf[t] + tot;
    }
  }
  __syncthreads();
  // top V ticks of select, lower tick first on ties
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (int r = 0; r < V; ++r) {
    float bv = -__builtin_inff();
    int bi = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < kPfChunk; ++k) {
      const int t = t0 + k;
      if (t < n && (s_sel[t] > bv || (s_sel[t] == bv && t < bi))) {
        bv = s_sel[t];
        bi = t;
      }
    }
#pragma unroll
    for (int dd = kWave / 2; dd > 0; dd >>= 1) {
      const float ov = __shfl_xor(bv, dd, kWave);
      const int oi = __shfl_xor(bi, dd, kWave);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      s_bv[wid] = bv;
      s_bi[wid] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float v0 = s_bv[0];
      int i0 = s_bi[0];
      for (int w = 1; w < kPfThreads / kWave; ++w)
        if (s_bv[w] > v0 || (s_bv[w] == v0 && s_bi[w] < i0)) {
          v0 = s_bv[w];
          i0 = s_bi[w];
        }
      if (i0 < 0 || i0 >= n) i0 = 0;  // all -inf (cannot happen for finite inputs)
      s_top[r] = i0;
      s_sel[i0] = -__builtin_inff();
    }
    __syncthreads();
  }
  if (threadIdx.x < V) {
    const int t = s_top[threadIdx.x];
    const int nx = min(min(t + d.interval + 1, T - 1) + 1, T - 1);
    st_c[threadIdx.x] = s_q[nx];
    st_p[threadIdx.x] = s_tot[t];
    tk[((int64_t)s * P + row) * V + threadIdx.x] = t;
  }
  if (threadIdx.x == 0) {
    float m = -__builtin_inff();
    for (int v = 0; v < V; ++v) m = fmaxf(m, s_tot[s_top[v]]);
    float sum = 0.f;
    for (int v = 0; v < V; ++v) sum += expf(s_tot[s_top[v]] - m);
    if (m + logf(sum) > d.log_stop) atomicOr(flags + s + 1, 1);
  }
}

// ------------------------------------------------------------------------------------------
// Adjoint of the beam search
// ------------------------------------------------------------------------------------------
//
// Reverse-mode derivative of (out_logp, out_q) with respect to the waveforms, with the path
// selections held fixed (top_k is under stop_gradient in the reference, fee_jax.py:377), as
// jax.grad takes it.  Per executed step s, in reverse, with G and Gq the cotangents of out_logp[s]
// and out_q[s] and (dc', dp') those of the paths the step selected:
//   p'_k = total[t_k]  -> d total[t_k] += dp'_k;  c'_k = q_sum[nx(t_k)] -> d q_sum[nx(t_k)] += dc'_k
//   total[t] = lse_v(lg[v,t] + p_v), out_logp[t] = lse_v(le[v,t] + p_v): softmax weights to lg, le, p
//   le = soft_where(esp - theta, soft_max(min(ld, lg), -1000, 1), -1000, 10)   (ties of min: half each)
//   ld = logdiffPhi(z(q_sum[sh] - c), z(q_sum[t] - c)), lg = logdiffPhi(z(CM[t+1] - c), z(CM[t] - c))
//   esp[t] = q_sum[sh] + theta - (q_sum[t+1] + q_sum[t]) / 2
// which gives d q_sum, d CM and, per path, d c and d p; paths merged in the forward (equal charge)
// share their group's derivative in proportion exp(p_v - p_group).  Then CM = cummax(q_sum): each
// d CM[t] goes to the last tick <= t where q_sum attains the running maximum (torch.cummax's
// choice; jax splits ties, a corner no fixture reaches).  Finally q_sum = cumsum(wfs t_sampling):
// d wfs[t] = t_sampling * sum_{u >= t} d q_sum[u].  Sums over paths and steps are float64, in a
// fixed order.

// d log(Phi(a) - Phi(b)) with the smooth clamps of pf_log_diff (fee_jax.py:32-54): returns the value
// and, for a cotangent g on it, the cotangents of a and b.
__device__ __forceinline__ float pf_sigmoid(float x) {
  return x > 30.f ? 1.f : (x < -30.f ? expf(x) : 1.f / (1.f + expf(-x)));
}

// d log Phi(x) / dx = phi(x) / Phi(x), from lx = log Phi(x)
__device__ __forceinline__ float pf_dlog_ndtr(float x, float lx) {
  return expf(-0.5f * x * x - lx) * 0.3989422804014327f;
}

__device__ __forceinline__ float pf_log_diff_adj(float a, float b, float la, float lb, float g, float& ga, float& gb) {
  const bool gt = a > b;
  const float sd = gt ? lb - la : -1.f;
  const float ne = -expm1f(sd);
  const float ne2 = pf_soft_max(ne, 1e-30f, 1e10f);
  const float lt = logf(ne2);
  const float lt2 = pf_soft_max(lt, -100.f, 10.f);
  const float lp = la + lt2;
  const float w = pf_sigmoid((a - b) * 1000.f);
  const float out = w * lp + (1.f - w) * (-1000.f);
  const float g_lp = w * g;
  const float g_amb = 1000.f * w * (1.f - w) * (lp + 1000.f) * g;
  const float g_lt = g_lp * pf_sigmoid((lt + 100.f) * 10.f);
  const float g_ne = g_lt / ne2 * pf_sigmoid((ne - 1e-30f) * 1e10f);
  const float g_sd = gt ? -expf(sd) * g_ne : 0.f;
  const float g_la = g_lp - g_sd;
  const float g_lb = g_sd;
  ga = g_la * pf_dlog_ndtr(a, la) + g_amb;
  gb = g_lb * pf_dlog_ndtr(b, lb) - g_amb;
  return out;
}

template <typename T>
__device__ __forceinline__ T pf_block_sum(T v, T* sh) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  T tot = 0;
  for (int w = 0; w < kPfThreads / kWave; ++w) tot += sh[w];
  __syncthreads();
  return tot;
}

// One workgroup per row: every executed step in reverse, then the cumulative max and sum.
__global__ __launch_bounds__(kPfThreads) void k_pf_adj(PfDev d, const float* __restrict__ tab, int64_t P,
                                                      const float* __restrict__ tc, const float* __restrict__ tp,
                                                      const int* __restrict__ tk, const int* __restrict__ flags,
                                                      const float* __restrict__ out_logp,
                                                      const float* __restrict__ g_logp, const float* __restrict__ g_q,
                                                      float* __restrict__ d_wfs) {
  __shared__ float s_q[kPfMaxTicks];
  __shared__ float s_cm[kPfMaxTicks];
  __shared__ double s_buf[kPfMaxTicks];  // per-path log Phi tables; at the end d q_sum
  float* const s_lq = (float*)s_buf;
  float* const s_lc = s_lq + kPfMaxTicks;
  __shared__ float s_dtot[kPfMaxTicks];  // d total at the selected ticks, 0 elsewhere
  __shared__ short s_kof[kPfMaxTicks];   // selected tick -> its path slot k, -1 elsewhere
  __shared__ double s_dqx[kPfMaxTicks];  // d q_sum at the selected paths' reset ticks; at the end d CM
  __shared__ float s_c[kPfMaxPaths], s_p[kPfMaxPaths], s_mc[kPfMaxPaths], s_mp[kPfMaxPaths], s_tot[kPfMaxPaths];
  __shared__ int s_grp[kPfMaxPaths];
  __shared__ double s_dcN[kPfMaxPaths], s_dpN[kPfMaxPaths], s_dcm[kPfMaxPaths], s_dpm[kPfMaxPaths];
  __shared__ double s_red[kPfThreads / kWave];
  __shared__ int s_nv;
  const int64_t row = blockIdx.x;
  const int T = d.T, V = d.V, n = T - 1;
  const int t0 = threadIdx.x * kPfChunk;
  const float* qs = tab + row * 4 * (int64_t)T;
  const float* cm = qs + T;
  const float* espg = cm + 2 * T;
  for (int t = threadIdx.x; t < T; t += kPfThreads) {
    s_q[t] = qs[t];
    s_cm[t] = cm[t];
    s_dqx[t] = 0.0;
    s_dtot[t] = 0.f;
    s_kof[t] = -1;
  }
  if (threadIdx.x < kPfMaxPaths) {
    s_dcN[threadIdx.x] = 0.0;
    s_dpN[threadIdx.x] = 0.0;
  }
  // per tick (own chunk): d q_sum[t], d q_sum[t+1], d q_sum[sh(t)], d CM[t], d CM[t+1]
  double aq[kPfChunk], bq[kPfChunk], cq[kPfChunk], dc0[kPfChunk], dc1[kPfChunk];
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) aq[k] = bq[k] = cq[k] = dc0[k] = dc1[k] = 0.0;
  __syncthreads();
  int n_run = 0;
  while (n_run < d.S && flags[n_run]) ++n_run;
  for (int s = n_run - 1; s >= 0; --s) {
    const float* ol = out_logp + (row * d.S + s) * (int64_t)n;
    const float* gl = g_logp + (row * d.S + s) * (int64_t)n;
    const float* gq = g_q + (row * d.S + s) * (int64_t)n;
    if (threadIdx.x < V) {
      const int v = threadIdx.x;
      s_c[v] = tc[((int64_t)s * P + row) * V + v];
      s_p[v] = tp[((int64_t)s * P + row) * V + v];
      s_tot[v] = tp[((int64_t)(s + 1) * P + row) * V + v];  // total[t_k] of this step's selections
      const int t = tk[((int64_t)s * P + row) * V + v];
      s_dtot[t] = (float)s_dpN[v];
      s_kof[t] = (short)v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // c'_k = q_sum[nx(t_k)] (in slot order; several k can share a reset tick)
      for (int v = 0; v < V; ++v) {
        const int t = tk[((int64_t)s * P + row) * V + v];
        const int nx = min(min(t + d.interval + 1, T - 1) + 1, T - 1);
        s_dqx[nx] += s_dcN[v];
      }
      // the forward's merge of equal-charge paths (same comparisons, same order)
      int nv = 0;
      for (int v = 0; v < V; ++v) {
        int u = 0;
        while (u < nv && s_mc[u] != s_c[v]) ++u;
        if (u == nv) {
          s_mc[nv] = s_c[v];
          s_mp[nv] = s_p[v];
          ++nv;
        } else {
          const float a = s_mp[u], b = s_p[v], m = fmaxf(a, b);
          s_mp[u] = m + log1pf(expf(fminf(a, b) - m));
        }
        s_grp[v] = u;
      }
      s_nv = nv;
    }
    __syncthreads();
    const int NV = s_nv;
    float gk[kPfChunk], olk[kPfChunk], espk[kPfChunk], desp[kPfChunk];
#pragma unroll
    for (int k = 0; k < kPfChunk; ++k) {
      const int t = t0 + k;
      gk[k] = t < n ? gl[t] : 0.f;
      olk[k] = t < n ? ol[t] : 0.f;
      espk[k] = t < n ? espg[t] : 0.f;
      desp[k] = 0.f;
    }
    for (int u = 0; u < NV; ++u) {
      const float c = s_mc[u], pv = s_mp[u];
      for (int t = threadIdx.x; t < T; t += kPfThreads) {
        s_lq[t] = pf_log_ndtr((s_q[t] - c - d.theta) * d.zs);
        s_lc[t] = pf_log_ndtr((s_cm[t] - c - d.theta) * d.zs);
      }
      __syncthreads();
      double dcu = 0.0, dpu = 0.0;
#pragma unroll
      for (int k = 0; k < kPfChunk; ++k) {
        const int t = t0 + k;
        if (t < n) {
          const int sh = min(t + d.interval + 1, T - 1);
          const float a1 = (s_cm[t + 1] - c - d.theta) * d.zs, b1 = (s_cm[t] - c - d.theta) * d.zs;
          const float a2 = (s_q[sh] - c - d.theta) * d.zs, b2 = (s_q[t] - c - d.theta) * d.zs;
          // values first (the cotangents of ld and lg depend on them)
          float ga1, gb1, ga2, gb2;
          const float lg = pf_log_diff(a1, b1, s_lc[t + 1], s_lc[t]);
          const float ld = pf_log_diff(a2, b2, s_lq[sh], s_lq[t]);
          const float mn = fminf(ld, lg);
          const float sm = pf_soft_max(mn, -1000.f, 1.f);
          const float we = pf_sigmoid((espk[k] - d.theta) * 10.f);
          const float le = we * sm + (1.f - we) * (-1000.f);
          // out_logp = lse(le + p): softmax weight of this path
          const float wh = expf(le + pv - olk[k]);
          const float g_le = gk[k] * wh;
          dpu += (double)g_le;
          float g_lg = 0.f;
          const int kk = s_kof[t];
          if (kk >= 0) {  // total[t] = lse(lg + p) feeds a selected path
            const float wt = expf(lg + pv - s_tot[kk]);
            const float g = s_dtot[t] * wt;
            g_lg += g;
            dpu += (double)g;
          }
          desp[k] += g_le * (sm + 1000.f) * 10.f * we * (1.f - we);
          const float g_mn = g_le * we * pf_sigmoid(mn + 1000.f);
          float g_ld = 0.f;
          if (ld < lg) g_ld = g_mn;
          else if (lg < ld) g_lg += g_mn;
          else {
            g_ld = 0.5f * g_mn;
            g_lg += 0.5f * g_mn;
          }
          (void)pf_log_diff_adj(a2, b2, s_lq[sh], s_lq[t], g_ld, ga2, gb2);
          (void)pf_log_diff_adj(a1, b1, s_lc[t + 1], s_lc[t], g_lg, ga1, gb1);
          cq[k] += (double)(d.zs * ga2);
          aq[k] += (double)(d.zs * gb2);
          dc1[k] += (double)(d.zs * ga1);
          dc0[k] += (double)(d.zs * gb1);
          dcu -= (double)(d.zs * (ga1 + gb1 + ga2 + gb2));
        }
      }
      dcu = pf_block_sum<double>(dcu, s_red);
      dpu = pf_block_sum<double>(dpu, s_red);
      if (threadIdx.x == 0) {
        s_dcm[u] = dcu;
        s_dpm[u] = dpu;
      }
      __syncthreads();
    }
    // esp[t] = q_sum[sh] + theta - (q_sum[t+1] + q_sum[t]) / 2
#pragma unroll
    for (int k = 0; k < kPfChunk; ++k) {
      const int t = t0 + k;
      if (t < n) {
        const double ge = (double)desp[k] + (double)gq[t];
        cq[k] += ge;
        bq[k] -= 0.5 * ge;
        aq[k] -= 0.5 * ge;
      }
    }
    // cotangents of the paths entering this step (the previous step's selections)
    if (threadIdx.x < V) {
      const int v = threadIdx.x, u = s_grp[v];
      const double share = (double)expf(s_p[v] - s_mp[u]);
      s_dcN[v] = s_dcm[u] * share;
      s_dpN[v] = s_dpm[u] * share;
      const int t = tk[((int64_t)s * P + row) * V + v];
      s_dtot[t] = 0.f;
      s_kof[t] = -1;
    }
    __syncthreads();
  }
  // d q_sum: own ticks, then t + 1 (crossing into the next chunk at the last tick), then sh(t)
  // (injective below the clamp), the clamped ticks sh(t) = T - 1 in tick order by one thread
  double* dqs = s_buf;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < T) dqs[t] = s_dqx[t] + aq[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < n && k < kPfChunk - 1) dqs[t + 1] += bq[k];
  }
  __syncthreads();
  if (t0 + kPfChunk - 1 < n) dqs[t0 + kPfChunk] += bq[kPfChunk - 1];
  __syncthreads();
  const int t_clamp = T - 1 - d.interval - 1;  // ticks t >= t_clamp have sh(t) = T - 1
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < n && t < t_clamp) dqs[t + d.interval + 1] += cq[k];
  }
  __syncthreads();
  // the clamped ones and d CM into s_dqx (free now): own ticks, then t + 1
  double* dcmv = s_dqx;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < T) dcmv[t] = t < n ? dc0[k] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < n && k < kPfChunk - 1) dcmv[t + 1] += dc1[k];
  }
  __syncthreads();
  if (t0 + kPfChunk - 1 < n) dcmv[t0 + kPfChunk] += dc1[kPfChunk - 1];
  // clamped sh contributions: written to s_lq-free scratch via a shuffle-free path: each thread's
  // clamped ticks go to LDS (reusing s_dtot as float? no: keep float64 in registers) -> one thread
  __shared__ double s_cl[kPfMaxTicks / kPfChunk + 1];
  {
    double cl = 0.0;
#pragma unroll
    for (int k = 0; k < kPfChunk; ++k) {
      const int t = t0 + k;
      if (t < n && t >= t_clamp) cl += cq[k];
    }
    s_cl[threadIdx.x] = cl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double cl = 0.0;
    for (int w = 0; w < kPfThreads; ++w) cl += s_cl[w];
    dqs[T - 1] += cl;
    // CM = cummax(q_sum): d CM[t] to the last tick <= t where q_sum reaches the running maximum
    int rec = 0;
    for (int t = 0; t < T; ++t) {
      if (s_q[t] == s_cm[t]) rec = t;
      dqs[rec] += dcmv[t];
    }
  }
  __syncthreads();
  // d wfs[t] = t_sampling * sum_{u >= t} d q_sum[u]: chunk sums, an exclusive suffix scan over
  // threads, then the walk down the chunk
  double loc = 0.0;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k)
    if (t0 + k < T) loc += dqs[t0 + k];
  const double incl = pf_block_scan(loc, s_red);  // inclusive prefix over threads
  const double total = pf_block_sum<double>(loc, s_red);
  double suffix = total - incl + loc;  // sum over u >= t0
  float* dw = d_wfs + row * (int64_t)T;
#pragma unroll
  for (int k = 0; k < kPfChunk; ++k) {
    const int t = t0 + k;
    if (t < T) {
      dw[t] = (float)(suffix * d.ts);
      suffix -= dqs[t];
    }
  }
}

}  // namespace lsim
