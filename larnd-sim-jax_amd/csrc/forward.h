// Forward kernels: segment keys and records, cell / pixel-row enumeration, the row-owned
// waveform kernel with the fused self-trigger, hit compaction, and the response tables.
#pragma once
#include "device.h"
#include "scan_sort.h"

namespace lsim {

// ------------------------------------------------------------------------------------------
// small utility kernels
// ------------------------------------------------------------------------------------------

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

template <typename T>
inline void fill(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

// ------------------------------------------------------------------------------------------
// 1. cell key per segment (input order)
// ------------------------------------------------------------------------------------------

// Key = main-pixel cell of the segment in the halo-extended grid, or n_cells (sorts last) for
// segments that neither create a pixel row nor carry charge: padding rows (eventID < 0),
// events out of range, and zero-charge segments whose main pixel is off the pixel plane.
__global__ void k_keys(Dev d, SegIn in, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, int* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in.n) return;
  Phys ph;
  seg_physics(d, in, i, ph);
  int64_t c = cell_of(d, ph.ev, ph.plane, ph.mpx, ph.mpy);
  const bool on_plane = ph.mpx >= 0 && ph.mpx < d.nx && ph.mpy >= 0 && ph.mpy < d.ny;
  bool keep = c >= 0 && (ph.mask ? true : on_plane);
  if (status && ph.ev >= d.n_events) atomicOr(status, LSIM_STATUS_BAD_EVENT);
  keys[i] = keep ? (uint32_t)c : (uint32_t)d.n_cells;
  vals[i] = (uint32_t)i;
}

// ------------------------------------------------------------------------------------------
// 2. records in cell order, cell ranges, pixel rows
// ------------------------------------------------------------------------------------------

__global__ void k_records(Dev d, SegIn in, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n,
                          SegRec* __restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || keys[i] >= (uint32_t)d.n_cells) return;
  Phys ph;
  seg_physics(d, in, vals[i], ph);
  double wx[kMaxNb], wy[kMaxNb];
  tran_weights(d, ph.x0, ph.st, wx, nullptr);
  tran_weights(d, ph.y0, ph.st, wy, nullptr);
  SegRec r;
  r.bx = ph.bx;
  r.by = ph.by;
  r.k0 = ph.k0;
  r.idx = ph.idx;
  r.ne = (float)ph.ne;
  r.f = (float)ph.f;
  r.la = (float)ph.la;
  r.lb = (float)ph.lb;
  r.lc = (float)ph.lc;
#pragma unroll
  for (int k = 0; k < kMaxNb; ++k) {
    r.wx[k] = k < d.nb ? (float)wx[k] : 0.f;
    r.wy[k] = k < d.nb ? (float)wy[k] : 0.f;
  }
  r.pad = 0.f;
  rec[i] = r;
}

// Cell ranges [begin, end) in the sorted order and the "first segment of an on-plane cell" flag
// whose prefix sum numbers the pixel rows (ascending pixel id = ascending cell key).
__global__ void k_cells(Dev d, const uint32_t* __restrict__ keys, int64_t n, int2* __restrict__ rng, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  int fl = 0;
  if (k < (uint32_t)d.n_cells) {
    const bool first = (i == 0) || keys[i - 1] != k;
    const bool last = (i == n - 1) || keys[i + 1] != k;
    if (first) rng[k].x = (int)i;
    if (last) rng[k].y = (int)(i + 1);
    if (first) {
      const int cx = (int)(k % d.gx) - d.halo;
      const int cy = (int)((k / d.gx) % d.gy) - d.halo;
      fl = (cx >= 0 && cx < d.nx && cy >= 0 && cy < d.ny) ? 1 : 0;
    }
  }
  flag[i] = fl;
}

__global__ void k_rows_map(Dev d, const uint32_t* __restrict__ keys, const int* __restrict__ flag, const int* __restrict__ ridx,
                           int64_t n, int* __restrict__ row_cell, int* __restrict__ cell_row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int r = ridx[i];
  const uint32_t k = keys[i];
  row_cell[r] = (int)k;
  cell_row[k] = r;
}

__global__ void k_copy_count(const int* __restrict__ src, int64_t* __restrict__ dst, int64_t max_rows, int* __restrict__ status) {
  const int64_t v = (int64_t)*src;
  if (dst) *dst = v;
  if (status && max_rows >= 0 && v > max_rows) atomicOr(status, LSIM_STATUS_ROWS_OVERFLOW);
}

// ------------------------------------------------------------------------------------------
// 3. the waveform kernel: one workgroup owns one pixel row
// ------------------------------------------------------------------------------------------

struct Tables {
  const float* TM;    // [nm][cm][cm][L]   main-bin template windows
  const float* TN;    // [lx][ly][L]       raw response window (neighbours)
  const float2* DM;   // [nm][cm][cm][nt]  (C[nt-L]-C[k], C[min(k+1,nt-1)]-C[k]) of template m
  const float2* DN;   // [lx][ly][nt]      same for the raw response
  const double* CTM;  // [nm][cm][cm][L+1] window prefix sums (adjoint)
  const double* CTN;  // [lx][ly][L+1]
};

struct RowOut {
  float* wfs;          // [max_rows][nout] or null
  float* adc;          // [max_rows][H]
  int* ticks;          // [max_rows][H]
  FeeTape* tape;       // [max_rows] or null
  int* hit_cnt;        // [max_rows]
  int64_t* pixel_ids;  // [max_rows]
};

__device__ __forceinline__ double digitize(const Dev& d, double q) {
  double x = q * d.gain + d.vped - d.vcm;
  x = x > 0 ? x : 0.0;
  double y = x * d.adcc / (d.vref - d.vcm);
  return y < d.adcc ? y : d.adcc;
}

// JAX's derivative of digitize (ties of max/min split the gradient evenly).
__device__ __forceinline__ double digitize_slope(const Dev& d, double q) {
  const double x = q * d.gain + d.vped - d.vcm;
  const double fx = x > 0 ? 1.0 : (x == 0 ? 0.5 : 0.0);
  const double y = (x > 0 ? x : 0.0) * d.adcc / (d.vref - d.vcm);
  const double fy = y < d.adcc ? 1.0 : (y == d.adcc ? 0.5 : 0.0);
  return fx * fy * d.gain * d.adcc / (d.vref - d.vcm);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    int u = __shfl_xor(v, m, kWave);
    v = u < v ? u : v;
  }
  return v;
}

// Self-trigger of one row (fee_jax.py:168-279, noise off) on Q held in LDS, then digitisation
// (fee_jax.py:55-71).  Q and Q0 hold the float64 cumulative charge; Q is consumed.
template <int NT>
__device__ void fee_row(const Dev& d, double* Q, const double* Q0, int r, int plane, const RowOut& out, int* sh_int) {
  constexpr int TPER = kMaxOut / NT;
  const int t = threadIdx.x;
  const int nout = d.nout;
  const int lane = t & (kWave - 1), wid = t / kWave;
  int pe[kMaxH], pg[kMaxH], pi[kMaxH];
  double pa[kMaxH];
  for (int k = 0; k < d.H; ++k) {
    int cand = 0x7fffffff;
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int o = t * TPER + u;
      if (o <= nout - 2) {
        if (Q[o + 1] >= d.thr && Q[o] <= d.thr && o < cand) cand = o;
      }
    }
    cand = wave_min_i(cand);
    if (lane == 0) sh_int[wid] = cand;
    __syncthreads();
    int istar = 0x7fffffff;
    for (int w = 0; w < NT / kWave; ++w) istar = sh_int[w] < istar ? sh_int[w] : istar;
    if (istar == 0x7fffffff) istar = nout - 2;
    int e = istar + 1 + d.interval;
    if (e >= nout) e = nout - 1;
    int g = istar + 1 + d.interval + 1;
    if (g >= nout) g = nout - 1;
    const double a = Q[e];
    const double Qg = Q[g];
    pe[k] = e;
    pg[k] = g;
    pi[k] = istar;
    pa[k] = a;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int o = t * TPER + u;
      if (o < nout) {
        const double x = Q[o] - Qg;
        Q[o] = x < 0 ? 0.0 : x;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    int hits = 0;
    const int H = d.H;
    for (int k = 0; k < H; ++k) {
      const bool invalid = (pa[k] < d.thr) || (pi[k] == nout - 2);
      const double q = invalid ? 0.0 : pa[k];
      const int tick = invalid ? nout - 2 : pi[k];
      out.adc[(int64_t)r * H + k] = (float)digitize(d, q);
      out.ticks[(int64_t)r * H + k] = tick;
      hits += (tick < nout - 3) ? 1 : 0;
    }
    out.hit_cnt[r] = hits;
    if (out.tape) {
      FeeTape tp;
      double val[2 * kMaxH];
      for (int k = 0; k < H; ++k) {
        tp.pt[k] = (int16_t)pe[k];
        tp.pt[H + k] = (int16_t)pg[k];
        val[k] = Q0[pe[k]];
        val[H + k] = Q0[pg[k]];
      }
      for (int k = 0; k < H; ++k) {
        const bool invalid = (pa[k] < d.thr) || (pi[k] == nout - 2);
        tp.slope[k] = invalid ? 0.f : (float)digitize_slope(d, pa[k]);
        tp.tick[k] = invalid ? nout - 2 : pi[k];
        const double qg = val[H + k];
        uint32_t m = 0;
        for (int j = 0; j < 2 * H; ++j) {
          const double x = val[j] - qg;
          if (x >= 0) m |= (1u << j);
          val[j] = x < 0 ? 0.0 : x;
        }
        tp.mask[k] = m;
      }
      tp.plane = plane;
      out.tape[r] = tp;
    }
  }
}

// Sum of the per-wave partial rows, waveform store, float64 cumulative charge into Q and Q0.
template <int NT>
__device__ void finish_row(const Dev& d, float* acc, int nwaves, double* Q, double* Q0, int r, const RowOut& out, double* shd) {
  constexpr int TPER = kMaxOut / NT;
  const int t = threadIdx.x;
  float row[TPER];
  double loc[TPER];
  double run = 0.0;
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int o = t * TPER + u;
    float v = 0.f;
    for (int w = 0; w < nwaves; ++w) v += acc[w * kMaxOut + o];
    row[u] = v;
    if (o < d.nout) run += (double)v * d.ts;
    loc[u] = run;
  }
  if (out.wfs) {
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int o = t * TPER + u;
      if (o < d.nout) out.wfs[(int64_t)r * d.nout + o] = row[u];
    }
  }
  double tot;
  const double ex = block_exclusive_scan<double>(run, shd, tot);
  __syncthreads();  // all partial rows read before Q overwrites the LDS
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int o = t * TPER + u;
    Q[o] = ex + loc[u];
    Q0[o] = ex + loc[u];
  }
  __syncthreads();
}

// Accumulate the contribution of segment record `R` (main pixel at offset (-a, -b) from this
// row's pixel) into the wave's partial row.  The segment's stripes landing on this row share
// one time window: V[j] = sum over stripes n_st * X_st[j], j < L, placed as (1-f) V at output
// tick S+j and f V at S-1+j (S = nt - L - k0), plus the cumsum boundary correction Delta at
// output ticks S-1 ((1-f) Delta) and S-2 (f Delta) (sim_jax.py:170-261).
__device__ __forceinline__ void add_pair(const Dev& d, const Tables& tb, const SegRec& R, int a, int b, int px, int py,
                                         float* accw, float* stage) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nsb = d.nsb, L = d.L;
  const bool has_nb = (a != 0) || (b != 0);
  const int ux = pymod_i(R.bx, nsb), uy = pymod_i(R.by, nsb);
  int cxN = 0, cyN = 0;
  if (has_nb) {
    cxN = (int)fabsf((float)(ux - nsb / 2) + 0.5f - (float)(a * nsb));
    cyN = (int)fabsf((float)(uy - nsb / 2) + 0.5f - (float)(b * nsb));
  }
  // diffusion bins whose pixel is this row's pixel
  unsigned imask = 0, jmask = 0;
  if (a >= -1 && a <= 1 && b >= -1 && b <= 1) {
    for (int i = 0; i < d.nb; ++i) {
      if (floordiv_i(R.bx + i - d.nbsym, nsb) == px) imask |= 1u << i;
      if (floordiv_i(R.by + i - d.nbsym, nsb) == py) jmask |= 1u << i;
    }
  }
  const bool has_bins = imask && jmask;
  if (!has_nb && !has_bins) return;

  const int cm = d.cm;
  const int idx = R.idx;
  const int64_t tstride = (int64_t)cm * cm * L;
  const float* TM0 = tb.TM + (int64_t)(idx - 1) * tstride;
  const float* TNp = tb.TN + ((int64_t)cxN * d.ly + cyN) * L;
  const float ne = R.ne;

  // window values
  for (int j0 = 0; j0 < L; j0 += kWave) {
    const int j = j0 + lane;
    if (j < L) {
      float v = 0.f;
      if (has_nb) v = ne * TNp[j];
      if (has_bins) {
        for (int i = 0; i < d.nb; ++i) {
          if (!((imask >> i) & 1)) continue;
          const int cx = (int)fabsf((float)(pymod_i(R.bx + i - d.nbsym, nsb) - nsb / 2) + 0.5f);
          const float nwx = ne * R.wx[i];
          for (int jj = 0; jj < d.nb; ++jj) {
            if (!((jmask >> jj) & 1)) continue;
            const int cy = (int)fabsf((float)(pymod_i(R.by + jj - d.nbsym, nsb) - nsb / 2) + 0.5f);
            const float* tp = TM0 + ((int64_t)cx * cm + cy) * L + j;
            const float t1 = tp[0], t2 = tp[tstride], t3 = tp[2 * tstride];
            v += (nwx * R.wy[jj]) * (t2 * R.lb + t1 * R.la + t3 * R.lc);
          }
        }
      }
      stage[2 + j] = v;
    }
  }
  if (lane < 2) stage[lane] = 0.f;
  if (lane == 2) stage[L + 2] = 0.f;

  // boundary correction: one lane per stripe, then a fixed-order wave sum
  float dl = 0.f;
  const int k0 = R.k0;
  const float f = R.f;
  if (lane == 0 && has_nb) {
    const float2 D = tb.DN[((int64_t)cxN * d.ly + cyN) * d.nt + k0];
    dl = ne * (D.x - f * D.y);
  } else if (lane >= 1 && lane <= kMaxNb * kMaxNb && has_bins) {
    const int i = (lane - 1) / kMaxNb, jj = (lane - 1) % kMaxNb;
    if (i < d.nb && jj < d.nb && ((imask >> i) & 1) && ((jmask >> jj) & 1)) {
      const int cx = (int)fabsf((float)(pymod_i(R.bx + i - d.nbsym, nsb) - nsb / 2) + 0.5f);
      const int cy = (int)fabsf((float)(pymod_i(R.by + jj - d.nbsym, nsb) - nsb / 2) + 0.5f);
      const float2 D = tb.DM[(((int64_t)idx * cm + cx) * cm + cy) * d.nt + k0];
      dl = (ne * R.wx[i] * R.wy[jj]) * (D.x - f * D.y);
    }
  }
  const float delta = wave_sum(dl);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const int S = d.nt - L - k0;
  const float omf = 1.f - f;
  for (int k0l = 0; k0l < L + 2; k0l += kWave) {
    const int k = k0l + lane;
    if (k < L + 2) {
      const int m = k - 2;
      const int o = S + m;
      // window ticks are kept for 1 <= o <= nout-1, corrections for 0 <= o <= nout-2
      // (everything else goes to the reference's garbage column)
      if (o >= 0 && o < d.nout) {
        float val = 0.f;
        if (o >= 1) val = omf * stage[m + 2] + f * stage[m + 3];
        if (o <= d.nout - 2) {
          if (m == -1) val += omf * delta;
          if (m == -2) val += f * delta;
        }
        accw[o] += val;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kStage = 264;  // >= L + 3 for L <= 260

template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_rows_fwd(Dev d, Tables tb, const SegRec* __restrict__ rec,
                                                            const int2* __restrict__ rng, const int* __restrict__ row_cell,
                                                            const int* __restrict__ n_rows_dev, int64_t max_rows,
                                                            RowOut out) {
  constexpr int NT = WAVES * kWave;
  constexpr int ACC_BYTES = WAVES * kMaxOut * 4;
  constexpr int Q_BYTES = 2 * kMaxOut * 8;
  constexpr int SM = ACC_BYTES > Q_BYTES ? ACC_BYTES : Q_BYTES;
  __shared__ __align__(16) char smem[SM];
  __shared__ float stage_all[WAVES][kStage];
  __shared__ double shd[NT / kWave];
  __shared__ int shi[NT / kWave];

  const int r = blockIdx.x;
  const int P = *n_rows_dev;
  if (r >= P || r >= max_rows) {
    if (r < max_rows && threadIdx.x == 0) out.hit_cnt[r] = 0;
    return;
  }
  float* acc = (float*)smem;
  const int t = threadIdx.x, wid = t / kWave;
  for (int o = t; o < WAVES * kMaxOut; o += NT) acc[o] = 0.f;
  const int cell = row_cell[r];
  const int px = cell % d.gx - d.halo;
  const int py = (cell / d.gx) % d.gy - d.halo;
  const int evp = cell / (d.gx * d.gy);  // ev * ntpc + plane
  const int plane = evp % d.ntpc;
  if (t == 0) out.pixel_ids[r] = ((int64_t)evp * d.ny + py) * d.nx + px;
  __syncthreads();

  float* accw = acc + wid * kMaxOut;
  float* stage = stage_all[wid];
  for (int a = -d.nn; a <= d.nn; ++a) {
    const int qx = px - a;
    if (qx < -d.halo || qx >= d.nx + d.halo) continue;
    for (int b = -d.nn; b <= d.nn; ++b) {
      const int qy = py - b;
      if (qy < -d.halo || qy >= d.ny + d.halo) continue;
      const int64_t qc = ((int64_t)evp * d.gy + (qy + d.halo)) * d.gx + (qx + d.halo);
      const int2 rg = rng[qc];
      for (int s = rg.x + wid; s < rg.y; s += WAVES) {
        const SegRec R = rec[s];
        if (R.ne == 0.f) continue;
        add_pair(d, tb, R, a, b, px, py, accw, stage);
      }
    }
  }
  __syncthreads();
  double* Q = (double*)smem;
  double* Q0 = Q + kMaxOut;
  finish_row<NT>(d, acc, WAVES, Q, Q0, r, out, shd);
  fee_row<NT>(d, Q, Q0, r, plane, out, shi);
}

// Self-trigger on caller waveforms (simulate_stochastic on an arbitrary wfs): one workgroup per
// row, the same finish/fee code as the fused path.
template <int NT>
__global__ __launch_bounds__(NT) void k_fee_rows(Dev d, const float* __restrict__ wfs, const int64_t* __restrict__ pids,
                                                 int64_t n_rows, RowOut out) {
  __shared__ __align__(16) char smem[2 * kMaxOut * 8 > kMaxOut * 4 ? 2 * kMaxOut * 8 : kMaxOut * 4];
  __shared__ float accs[kMaxOut];
  __shared__ double shd[NT / kWave];
  __shared__ int shi[NT / kWave];
  const int64_t r = blockIdx.x;
  if (r >= n_rows) return;
  for (int o = threadIdx.x; o < kMaxOut; o += NT) accs[o] = o < d.nout ? wfs[r * d.nout + o] : 0.f;
  __syncthreads();
  const int64_t pid = pids[r];
  const int npl = d.nx * d.ny;
  const int plane = pid >= 0 ? (int)((pid / npl) % d.ntpc) : 0;
  double* Q = (double*)smem;
  double* Q0 = Q + kMaxOut;
  RowOut o2 = out;
  o2.wfs = nullptr;
  finish_row<NT>(d, accs, 1, Q, Q0, (int)r, o2, shd);
  fee_row<NT>(d, Q, Q0, (int)r, plane, o2, shi);
  if (threadIdx.x == 0 && pid < 0) out.hit_cnt[r] = 0;
}

// ------------------------------------------------------------------------------------------
// 4. hit compaction (sim_jax.py:619-647) with coordinates (detsim_jax.py:296-319)
// ------------------------------------------------------------------------------------------

struct HitOut {
  float *adc, *x, *y, *z, *tick, *prob;
  int64_t *event, *pixel;
  int32_t* index;
};

__global__ void k_hits(Dev d, const int64_t* __restrict__ pids, const float* __restrict__ adc, const int* __restrict__ ticks,
                       const int* __restrict__ hit_off, const int64_t* __restrict__ n_rows_dev, int64_t max_rows,
                       int64_t n_rows_static, HitOut h) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t P = n_rows_dev ? *n_rows_dev : n_rows_static;
  if (P > max_rows) P = max_rows;
  const int H = d.H;
  if (i >= P * H) return;
  const int64_t r = i / H;
  const int k = (int)(i % H);
  const int tick = ticks[i];
  if (!(tick < d.nout - 3)) return;
  const int64_t pid = pids[r];
  if (pid < 0) return;
  int before = 0;
  for (int kk = 0; kk < k; ++kk) before += (ticks[r * H + kk] < d.nout - 3) ? 1 : 0;
  const int64_t pos = (int64_t)hit_off[r] + before;
  const int npl = d.nx * d.ny;
  const int px = (int)(pid % d.nx), py = (int)((pid / d.nx) % d.ny);
  const int plane = (int)((pid / npl) % d.ntpc);
  const int64_t ev = pid / ((int64_t)npl * d.ntpc);
  const double za = d.tpc[plane][2][0], zc = d.tpc[plane][2][1];
  const double sgn = (zc - za) > 0 ? 1.0 : ((zc - za) < 0 ? -1.0 : 0.0);
  h.adc[pos] = adc[i];
  h.x[pos] = (float)((double)px * d.pitch + d.tpc[plane][0][0] + d.pitch / 2);
  h.y[pos] = (float)((double)py * d.pitch + d.tpc[plane][1][0] + d.pitch / 2);
  h.z[pos] = (float)(za + (double)tick * d.ts * d.v * sgn);
  h.tick[pos] = (float)tick;
  h.prob[pos] = 1.0f;
  h.event[pos] = ev;
  h.pixel[pos] = pid;
  h.index[pos] = (int32_t)i;
}

// ------------------------------------------------------------------------------------------
// 5. response tables (consts_jax.py:427-447), float64 arithmetic
// ------------------------------------------------------------------------------------------

// One workgroup per (template m, cx, cy) series: template values over the full time range,
// prefix sums, then the window, boundary-difference and window-prefix tables.
// main != 0: m in [0, nm), (cx, cy) in [0, cm)^2 -> TM, DM, CTM; else the raw response over all
// (cx, cy) -> TN, DN, CTN.
__global__ __launch_bounds__(256) void k_lut_tables(Dev d, const float* __restrict__ raw, int main_tables, float* __restrict__ Tw,
                                                    float2* __restrict__ Dt, double* __restrict__ CT) {
  __shared__ double T[kMaxOut];
  __shared__ double g[128];
  __shared__ double shd[4];
  const int series = blockIdx.x;
  int m, cx, cy;
  if (main_tables) {
    m = series / (d.cm * d.cm);
    cx = (series / d.cm) % d.cm;
    cy = series % d.cm;
  } else {
    m = 0;
    cx = series / d.ly;
    cy = series % d.ly;
  }
  const int nt = d.nt, ext = d.ext;
  const float* R = raw + ((int64_t)cx * d.ly + cy) * nt;
  // normalised Gaussian kernel of template m (norm.pdf(k, scale=sigma), consts_jax.py:427-428)
  if (threadIdx.x == 0) {
    const double s = d.sig[m];
    double sum = 0.0;
    for (int k = -ext; k <= ext; ++k) {
      const double z = (double)k / s;
      const double v = exp(-0.5 * z * z) / (sqrt(2.0 * M_PI) * s);
      g[k + ext] = v;
      sum += v;
    }
    for (int k = 0; k <= 2 * ext; ++k) g[k] /= sum;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    double v;
    if (m == 0) {
      v = (double)R[t];
    } else {
      v = 0.0;
      for (int k = -ext; k <= ext; ++k) {
        const int tt = t + k;
        if (tt >= 0 && tt < nt) v += (double)R[tt] * g[k + ext];
      }
    }
    T[t] = v;
  }
  __syncthreads();
  // inclusive prefix sum C[t] (blocked: each thread 8 consecutive entries)
  constexpr int PER = kMaxOut / 256;
  double loc[PER];
  double run = 0.0;
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x * PER + u;
    if (t < nt) run += T[t];
    loc[u] = run;
  }
  double tot;
  const double ex = block_exclusive_scan<double>(run, shd, tot);
  __syncthreads();
  const int L = d.L;
  // keep the window values (float32, as used by the forward) before T is overwritten
  __shared__ float win[kMaxOut];
  for (int j = threadIdx.x; j < L; j += blockDim.x) win[j] = (float)T[nt - L + j];
  __syncthreads();
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x * PER + u;
    if (t < nt) T[t] = ex + loc[u];  // T now holds C
  }
  __syncthreads();
  const int64_t sbase = (int64_t)series;
  for (int j = threadIdx.x; j < L; j += blockDim.x) Tw[sbase * L + j] = win[j];
  const double cref = T[nt - L];
  for (int k = threadIdx.x; k < nt; k += blockDim.x) {
    const double c0 = T[k];
    const double c1 = T[k + 1 < nt ? k + 1 : nt - 1];
    Dt[sbase * nt + k] = make_float2((float)(cref - c0), (float)(c1 - c0));
  }
  if (threadIdx.x == 0) {
    double c = 0.0;
    CT[sbase * (L + 1)] = 0.0;
    for (int j = 0; j < L; ++j) {
      c += (double)win[j];
      CT[sbase * (L + 1) + j + 1] = c;
    }
  }
}

}  // namespace lsim
