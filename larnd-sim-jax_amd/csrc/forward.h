// Forward kernels: segment keys and records, cell / pixel-row enumeration, the row-owned
// waveform kernel with the fused self-trigger, hit compaction, and the response tables.
#pragma once
#include "device.h"
#include "scan_sort.h"

namespace lsim {

// Window tables used by the waveform kernels are zero padded: P[p] = W[p - 2] for 2 <= p < L + 2
// and 0 elsewhere, kWinPad entries per series, so chunk c of lane l reads P[63 c + l] with no
// bounds test (windows up to 3 * 63 - 2 = 187 samples).
constexpr int kWinPad = 3 * kWave;
constexpr int kXcd = 8;  // XCDs of an MI355X (workgroup i runs on XCD i % 8)

// Per-segment record of the window-store paths (sorted order, 32 bytes = one scalar load): window
// start, charge, fractional tick, the interior-window flag, and byte offsets from which the
// padded raw-window row of any neighbour offset follows with a few scalar operations.  For a
// sub-bin position u and a pixel offset a != 0, offset_bin(u, a) = nsb a + nsb/2 - 1 - u (a > 0)
// or u - nsb/2 + nsb |a| (a < 0): affine in u, so the row is a per-cell constant plus +-U.
struct __align__(32) PairRec {
  int S;          // first output tick of the (1-f) window: nt - L - k0
  float ne, f;
  int flags;      // bit 0: 2 <= S and S + L <= nout (interior window); bits 8-15: u; bits 16-23: v
  int U, V;       // u * ly * kWinPad * 4 and v * kWinPad * 4
  int U0, V0;     // offset_bin(u, 0) * ly * kWinPad * 4 and offset_bin(v, 0) * kWinPad * 4
};
static_assert(sizeof(PairRec) == 32, "PairRec layout");

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

// Sort keys of a segment.  Cell key = main-pixel cell of the segment in the halo-extended grid,
// or n_cells (sorts last) for segments that neither create a pixel row nor carry charge: padding
// rows (eventID < 0), events out of range, and zero-charge segments whose main pixel is off the
// pixel plane.  Window key = the first output tick S = nt - L - k0 of the segment's windows,
// clamped to [0, nout + 2]: segments are ordered by cell, then by S (an adjoint row's live pairs,
// those whose windows start before its gradient vanishes, then lead each cell).
__device__ __forceinline__ void seg_keys(const Dev& d, const SegIn& in, int64_t i, uint32_t& cell_key, uint32_t& s_key,
                                         int* status) {
  Phys ph;
  seg_physics(d, in, i, ph);
  int64_t c = cell_of(d, ph.ev, ph.plane, ph.mpx, ph.mpy);
  const bool on_plane = ph.mpx >= 0 && ph.mpx < d.nx && ph.mpy >= 0 && ph.mpy < d.ny;
  bool keep = c >= 0 && (ph.mask ? true : on_plane);
  if (status && ph.ev >= d.n_events) atomicOr(status, LSIM_STATUS_BAD_EVENT);
  cell_key = keep ? (uint32_t)c : (uint32_t)d.n_cells;
  const int S = d.nt - d.L - ph.k0;
  s_key = (uint32_t)(S < 0 ? 0 : (S > d.nout + 2 ? d.nout + 2 : S));
}

// First sort stage: window keys (input order).
__global__ void k_keys(Dev d, SegIn in, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, int* __restrict__ status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in.n) return;
  uint32_t ck, sk;
  seg_keys(d, in, i, ck, sk, status);
  keys[i] = sk;
  vals[i] = (uint32_t)i;
}

// Second stage: the cell keys of the window-sorted order (a stable sort by cell follows).
__global__ void k_cell_keys(Dev d, SegIn in, const uint32_t* __restrict__ vals, int64_t n, uint32_t* __restrict__ keys) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t ck, sk;
  seg_keys(d, in, vals[j], ck, sk, nullptr);
  keys[j] = ck;
}

// ------------------------------------------------------------------------------------------
// 2. records in cell order, cell ranges, pixel rows
// ------------------------------------------------------------------------------------------

// Boundary corrections of every (row offset, segment) pair of record R, written where the record
// is built: dtot[ab][i] = n_e (D.x - f D.y) of the neighbour stripe (raw response, unless the offset
// is (0, 0)) plus, for |a|, |b| <= 1, the sum over the diffusion bins landing on that row of
// n_bin (D.x - f D.y) of the idx template (sim_jax.py:227-261), D = (C[nt-L] - C[k0],
// C[k0+1] - C[k0]), ab = (a + n)(2n + 1) + (b + n).  The D tables are tick-major, so all of one
// segment's entries lie in the tick-k0 plane; the waveform kernels then read dtot[ab][cell range]
// contiguously.
//
// The stripe entries come from DNS, the raw response's D of every (tick, sub-bin x, sub-bin y)
// regrouped so that one segment's WW entries are contiguous (k_lut_dns).  A wave stages each batch
// of kCorrBatch entries of its 64 segments through LDS: lane e of load m fetches element
// e + 64 m of the batch (segment (e + 64 m) / kCorrBatch, entry (e + 64 m) % kCorrBatch), so one
// load instruction reads about seven segments' runs of kCorrBatch contiguous entries instead of 64
// scattered ones; each lane then takes its own segment's entries from LDS.  The diffusion-bin loads
// are issued one x-bin row at a time so that each thread has several in flight; each entry still
// receives its terms in the same order (its stripe term, then its diffusion bins by x-bin, y-bin).
constexpr int kCorrBatch = 9;
constexpr int kRecThreads = 256;

template <int NB>
__device__ __forceinline__ void seg_corrections(const Dev& d, const float2* __restrict__ DNS, const float2* __restrict__ DM,
                                                const SegRec& R, bool live, int ux, int uy, int64_t i, int64_t n,
                                                float2* __restrict__ stage, int* __restrict__ base,
                                                float* __restrict__ dtot) {
  constexpr int SYM = (NB - 1) / 2;
  const int W = 2 * d.ne + 1, WW = W * W, nsb = d.nsb, cm = d.cm;
  const int lane = threadIdx.x & (kWave - 1);
  const float f = R.f;
  // the WW entries of (k0, ux, uy) start here (fits an int: nt nsb^2 WW < 2^31 for every valid LUT)
  base[lane] = live ? (int)((((int64_t)R.k0 * nsb + ux) * nsb + uy) * WW) : -1;
  __builtin_amdgcn_wave_barrier();
  float dn[9];  // entries (a, b) in [-1, 1]^2, index (a + 1) * 3 + (b + 1)
  // stripe terms of every entry
  for (int ab0 = 0; ab0 < WW; ab0 += kCorrBatch) {
#pragma unroll
    for (int m = 0; m < kCorrBatch; ++m) {
      const int e = lane + kWave * m;
      const int sg = e / kCorrBatch, u = e - sg * kCorrBatch;
      const int b = base[sg];
      stage[e] = (b >= 0 && ab0 + u < WW) ? DNS[b + ab0 + u] : make_float2(0.f, 0.f);
    }
    // wave-local LDS: the stores and loads of one wave complete in program order; the wave
    // barriers keep the compiler from moving them across
    __builtin_amdgcn_wave_barrier();
    float2 D[kCorrBatch];
#pragma unroll
    for (int u = 0; u < kCorrBatch; ++u) {
      const int ab = ab0 + u;
      const int a = ab / W - d.ne, b = ab % W - d.ne;
      const bool on = ab < WW && (a != 0 || b != 0) && a >= -d.nn && a <= d.nn && b >= -d.nn && b <= d.nn;
      D[u] = on ? stage[lane * kCorrBatch + u] : make_float2(0.f, 0.f);
    }
    __builtin_amdgcn_wave_barrier();  // every lane has its entries before the next batch is staged
    if (!live) continue;
    float dl[kCorrBatch];  // all of the batch's values before its stores: no store between two uses
#pragma unroll
    for (int u = 0; u < kCorrBatch; ++u) {
      const int ab = ab0 + u;
      const int a = ab / W - d.ne, b = ab % W - d.ne;
      const bool on = ab < WW && (a != 0 || b != 0) && a >= -d.nn && a <= d.nn && b >= -d.nn && b <= d.nn;
      dl[u] = on ? R.ne * (D[u].x - f * D[u].y) : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kCorrBatch; ++u) {
      const int ab = ab0 + u;
      if (ab >= WW) break;
      const int a = ab / W - d.ne, b = ab % W - d.ne;
      const float delta = dl[u];
      if (a >= -1 && a <= 1 && b >= -1 && b <= 1) {
#pragma unroll
        for (int e = 0; e < 9; ++e)
          if (e == (a + 1) * 3 + (b + 1)) dn[e] = delta;
      } else {
        dtot[(int64_t)ab * n + i] = delta;
      }
    }
  }
  if (!live) return;
  const float2* DMk = DM + ((int64_t)R.k0 * d.nm + R.idx) * cm * cm;
  // diffusion bins of the idx template, one x-bin row of loads at a time
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int vx = ux + k - SYM;
    const int dx = vx < 0 ? -1 : (vx >= nsb ? 1 : 0);
    const float wxs = R.ne * R.wx[k];
    const int cx = offset_bin(vx - dx * nsb, 0, nsb);
    float2 D[NB];
    float c[NB];
    int e[NB];
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      const int vy = uy + jj - SYM;
      const int dy = vy < 0 ? -1 : (vy >= nsb ? 1 : 0);
      c[jj] = wxs * R.wy[jj];
      e[jj] = (dx + 1) * 3 + (dy + 1);
      D[jj] = c[jj] != 0.f ? DMk[cx * cm + offset_bin(vy - dy * nsb, 0, nsb)] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      if (c[jj] == 0.f) continue;
      const float v = c[jj] * (D[jj].x - f * D[jj].y);
#pragma unroll
      for (int q = 0; q < 9; ++q)
        if (q == e[jj]) dn[q] += v;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int a = q / 3 - 1, b = q % 3 - 1;
    dtot[(int64_t)((a + d.ne) * W + (b + d.ne)) * n + i] = dn[q];
  }
}

// The record of the segment with input index src at sorted position i, and its pair record.
__device__ __forceinline__ void record(const Dev& d, const SegIn& in, int64_t src, int64_t i, SegRec& r, int& ux, int& uy,
                                       SegRec* __restrict__ rec, PairRec* __restrict__ prs) {
  Phys ph;
  seg_physics(d, in, src, ph);
  double wx[kMaxNb], wy[kMaxNb];
  tran_weights(d, ph.x0, ph.st, wx, nullptr);
  tran_weights(d, ph.y0, ph.st, wy, nullptr);
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
  ux = pymod_i(ph.bx, d.nsb);
  uy = pymod_i(ph.by, d.nsb);
  r.uxy = ux | (uy << 16);
  rec[i] = r;
  PairRec P;
  P.S = d.nt - d.L - ph.k0;
  P.ne = r.ne;
  P.f = r.f;
  P.flags = ((P.S >= 2 && P.S + d.L <= d.nout) ? 1 : 0) | (ux << 8) | (uy << 16);
  P.U = ux * d.ly * kWinPad * 4;
  P.V = uy * kWinPad * 4;
  P.U0 = offset_bin(ux, 0, d.nsb) * d.ly * kWinPad * 4;
  P.V0 = offset_bin(uy, 0, d.nsb) * kWinPad * 4;
  prs[i] = P;
}

// One thread per segment in sorted order (kRecThreads per workgroup): the records and the boundary
// corrections.  Threads past the end or of dropped segments stay to stage their wave's loads.
template <int NB>
__global__ __launch_bounds__(kRecThreads) void k_records(Dev d, SegIn in, const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals, int64_t n,
                                                         const float2* __restrict__ DNS, const float2* __restrict__ DM,
                                                         SegRec* __restrict__ rec, PairRec* __restrict__ prs,
                                                         float* __restrict__ dtot) {
  __shared__ float2 stage[kRecThreads / kWave][kWave * kCorrBatch];
  __shared__ int sbase[kRecThreads / kWave][kWave];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wid = threadIdx.x / kWave;
  const bool live = i < n && keys[i] < (uint32_t)d.n_cells;
  SegRec r = {};
  int ux = 0, uy = 0;
  if (live) record(d, in, vals[i], i, r, ux, uy, rec, prs);
  seg_corrections<NB>(d, DNS, DM, r, live, ux, uy, i, n, stage[wid], sbase[wid], dtot);
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
  const float3* TM3;  // [nm][cm][cm][kWinPad] padded (T_{m-1}, T_m, T_{m+1}) windows of the diffusion bins
  const float* TM;    // [nm][cm][cm][L]   main-bin template windows
  const float* TN;    // [lx][ly][L]       raw response window (neighbours)
  const float* TNP;   // [lx][ly][kWinPad] the same, padded
  const float2* DM;   // [nt][nm][cm][cm]  (C[nt-L]-C[k], C[min(k+1,nt-1)]-C[k]) of template m (tick-major)
  const float2* DN;   // [nt][lx][ly]      same for the raw response
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

// NaN passes through both clamps, as through jnp.maximum / jnp.minimum (a waveform poisoned by an
// overflowing current keeps its NaN charge, as in the reference).
__device__ __forceinline__ double digitize(const Dev& d, double q) {
  double x = q * d.gain + d.vped - d.vcm;
  x = (x > 0 || x != x) ? x : 0.0;
  double y = x * d.adcc / (d.vref - d.vcm);
  return (y < d.adcc || y != y) ? y : d.adcc;
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

// Self-trigger of one row (fee_jax.py:168-279, noise off), then digitisation (fee_jax.py:55-71).
// The float64 cumulative charge Q stays in registers: thread t holds ticks t * TPER + u in q[u]
// (as finish_row leaves them) and takes the tick after its last one from the next lane (DPP, or
// LDS across waves once, then updated by the same resets).  Q0 (LDS, written only with a tape)
// is the unreset copy.  The passes' decisions go to LDS from the threads that own the ticks;
// wave 0 then writes the outputs and the tape with one lane per pass (and one lane per tape point
// for the clamp masks).
template <int NT>
__device__ void fee_row(const Dev& d, double (&q)[kMaxOut / NT], const double* Q0, int r, int plane,
                        const RowOut& out, int* sh_int) {
  constexpr int TPER = kMaxOut / NT;
  constexpr int NW = NT / kWave;
  __shared__ int s_pe[kMaxH], s_pg[kMaxH], s_pi[kMaxH];
  __shared__ double s_pa[kMaxH];
  __shared__ double s_qmax[NW], s_first[NW];
  __shared__ double s_qg, s_gnone;
  const int t = threadIdx.x;
  const int nout = d.nout;
  const int lane = t & (kWave - 1), wid = t / kWave;
  const int H = d.H;
  const int o0 = t * TPER;
  // Q[o] into v when this thread holds tick o (register indices fixed at compile time)
  auto held = [&](int o, double& v) {
    bool own = false;
#pragma unroll
    for (int u = 0; u < TPER; ++u)
      if (o0 + u == o) {
        v = q[u];
        own = true;
      }
    return own;
  };
  // reset index of a pass without a crossing (i* = nout - 2), and its window end
  int g_none = nout - 2 + 1 + d.interval + 1;
  if (g_none >= nout) g_none = nout - 1;
  int e_none = nout - 2 + 1 + d.interval;
  if (e_none >= nout) e_none = nout - 1;
  if (lane == 0) s_first[wid] = q[0];
  {
    double v;
    if (held(g_none, v)) s_gnone = v;
  }
  __syncthreads();
  double nx = __shfl_down(q[0], 1, kWave);  // Q[o0 + TPER]
  if (lane == kWave - 1) nx = wid + 1 < NW ? s_first[wid + 1] : 0.0;
  const double q_gnone = s_gnone;
  for (int k = 0; k < H; ++k) {
    int cand = 0x7fffffff;
    double qmax = -1.0e300;  // max of Q[1 .. nout-1], the values a crossing needs at or above thr
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int o = o0 + u;
      if (o <= nout - 2) {
        const double q1 = u + 1 < TPER ? q[u + 1] : nx;
        if (q1 >= d.thr && q[u] <= d.thr && o < cand) cand = o;
        qmax = q1 > qmax ? q1 : qmax;
      }
    }
    cand = wave_min_i(cand);
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
      const double v = __shfl_xor(qmax, m, kWave);
      qmax = v > qmax ? v : qmax;
    }
    if (lane == 0) {
      sh_int[wid] = cand;
      s_qmax[wid] = qmax;
    }
    __syncthreads();
    int istar = 0x7fffffff;
    double bmax = -1.0e300;
    for (int w = 0; w < NW; ++w) {
      istar = sh_int[w] < istar ? sh_int[w] : istar;
      bmax = s_qmax[w] > bmax ? s_qmax[w] : bmax;
    }
    if (bmax < d.thr && (k > 0 || q_gnone >= 0.0)) {
      // Nothing reaches the threshold: no crossing in this pass, and none later either.  A reset
      // Q <- max(Q - Q[g], 0) with Q[g] >= 0 never raises a value above max(Q, 0) < thr.  After
      // the first pass Q >= 0 everywhere, so Q[g] >= 0; in the first pass (the raw cumulative
      // charge may be negative) it is checked.  Every remaining pass ends at i* = nout-2 (no hit)
      // with e and g at their clamps, exactly as running it would; Q is not read again (the tape
      // follows Q0).
      double v;
      if (held(e_none, v)) {
        for (int kk = k; kk < H; ++kk) {
          s_pe[kk] = e_none;
          s_pg[kk] = g_none;
          s_pi[kk] = nout - 2;
          s_pa[kk] = v;
        }
      }
      __syncthreads();
      break;
    }
    if (istar == 0x7fffffff) istar = nout - 2;
    int e = istar + 1 + d.interval;
    if (e >= nout) e = nout - 1;
    int g = istar + 1 + d.interval + 1;
    if (g >= nout) g = nout - 1;
    {
      double v;
      if (held(e, v)) {
        s_pe[k] = e;
        s_pg[k] = g;
        s_pi[k] = istar;
        s_pa[k] = v;
      }
      if (held(g, v)) s_qg = v;
    }
    __syncthreads();
    const double Qg = s_qg;
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      if (o0 + u < nout) {
        const double x = q[u] - Qg;
        q[u] = x < 0 ? 0.0 : x;
      }
    }
    const double x = nx - Qg;  // the next thread's first tick, reset as its owner resets it
    nx = x < 0 ? 0.0 : x;
  }
  if (t >= kWave) return;
  // lane k < H: pass k's hit
  bool hit = false;
  float slope = 0.f;
  int tick = nout - 2;
  if (t < H) {
    const double a = s_pa[t];
    const int i = s_pi[t];
    const bool invalid = (a < d.thr) || (i == nout - 2);
    tick = invalid ? nout - 2 : i;
    out.adc[(int64_t)r * H + t] = (float)digitize(d, invalid ? 0.0 : a);
    out.ticks[(int64_t)r * H + t] = tick;
    hit = tick < nout - 3;
    slope = invalid ? 0.f : (float)digitize_slope(d, a);
  }
  const int hits = __popcll(__ballot(hit));
  if (t == 0) out.hit_cnt[r] = hits;
  if (!out.tape) return;
  // the clamp masks: lane j < 2H follows Q at point j (e_j, then g_{j-H}) through the passes
  FeeTape* tp = out.tape + r;
  double v = 0.0;
  if (t < 2 * H) {
    const int p = t < H ? s_pe[t] : s_pg[t - H];
    tp->pt[t] = (int16_t)p;
    v = Q0[p];
  }
  uint32_t mine = 0;
  for (int k = 0; k < H; ++k) {
    const double qg = __shfl(v, H + k, kWave);
    const double x = v - qg;
    const uint32_t m = (uint32_t)(__ballot(t < 2 * H && x >= 0) & 0xffffffffull);
    if (t == k) mine = m;
    v = x < 0 ? 0.0 : x;
  }
  if (t < H) {
    tp->mask[t] = mine;
    tp->slope[t] = slope;
    tp->tick[t] = tick;
  }
  if (t == 0) tp->plane = plane;
}

// Sum of the per-wave partial rows, waveform store, float64 cumulative charge: thread t holds
// Q[t * TPER + u] in q[u] (consecutive ticks per thread); with a tape Q0 receives a copy in LDS.
template <int NT>
__device__ void finish_row(const Dev& d, const float* acc, int nwaves, int rstride, double (&q)[kMaxOut / NT],
                           double* Q0, int r, const RowOut& out, double* shd) {
  constexpr int TPER = kMaxOut / NT;
  const int t = threadIdx.x;
  float row[TPER];
  double loc[TPER];
  double run = 0.0;
#pragma unroll
  for (int u = 0; u < TPER; ++u) {
    const int o = t * TPER + u;
    float v = 0.f;
    if (o < d.nout)
      for (int w = 0; w < nwaves; ++w) v += acc[w * rstride + o];
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
#pragma unroll
  for (int u = 0; u < TPER; ++u) q[u] = ex + loc[u];
  // the scan's barriers follow every read of the partial rows, which Q0 may now overwrite
  if (out.tape) {
#pragma unroll
    for (int u = 0; u < TPER; ++u) {
      const int o = t * TPER + u;
      if (o < d.nout) Q0[o] = q[u];
    }
  }
}

// Contributions of one segment to one pixel row.  The segment's stripes landing on the row share
// one time window: V[j] = sum over stripes n_st * X_st[j], j < L, placed as (1-f) V at output
// tick S+j and f V at S-1+j (S = nt - L - k0), plus the cumsum boundary correction Delta at
// output ticks S-1 ((1-f) Delta) and S-2 (f Delta) (sim_jax.py:170-261).  Lane l of chunk c
// owns output offset m = 63*c + l - 2 and holds V[m]; V[m+1] comes from lane l+1.  Window ticks
// are kept for 1 <= o <= nout-1, corrections for 0 <= o <= nout-2 (the reference's garbage
// column takes the rest).  Every per-pair quantity is wave uniform (scalar registers).
constexpr int kChunk = kWave - 1;
// Slack after each wave's partial row for the interior store: a window at S (2 <= S, S + L <= nout)
// is stored as three chunks of 63 lanes from index S - 2, so it writes up to S + 3 * 63 - 3 <=
// nout - L + 186.
__host__ __device__ inline int row_pad(int L) { return L < 3 * kChunk - 2 ? 3 * kChunk - 2 - L : 0; }

// lane i receives lane i+1's value (lane 63 receives 0): DPP wavefront shift, no LDS traffic
__device__ __forceinline__ float lane_next(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

// General store of one chunk with every bound checked (segments whose window reaches the ends).
__device__ __forceinline__ void scatter_chunk(float* accw, int lane, int m, float v, float vn, float omf, float f,
                                              float delta, int S, int L, int nout) {
  const int o = S + m;
  if (lane < kChunk && m < L && o >= 0 && o < nout) {
    float val = 0.f;
    if (o >= 1) val = omf * v + f * vn;
    if (o <= nout - 2) {
      if (m == -1) val += omf * delta;
      if (m == -2) val += f * delta;
    }
    accw[o] += val;
  }
}

// Store of a whole window held as three chunks (v0, v1, v2) of up to 63 outputs each, V = v
// scaled by the charge: ne_omf = scale*(1-f), ne_f = scale*f.  `interior` windows (the caller
// passes 2 <= S and S + L <= nout) need no per-lane tick checks: output m of chunk c goes to
// accw[S + m] with m = 63 c + lane - 2.
__device__ __forceinline__ void store_window(float* accw, int lane, float v0, float v1, float v2, float scale, float omf,
                                             float f, float delta, int S, int L, int nout, bool interior) {
  const float n0 = lane_next(v0), n1 = lane_next(v1), n2 = lane_next(v2);
  if (interior) {
    // all three chunks in one go (disjoint addresses); lanes past the window add +0 to the
    // row or to the row_pad slack after it, so no per-chunk branch is needed
    const float so = scale * omf, sf = scale * f;
    float* p = accw + (S - 2) + lane;
    float c0 = so * v0 + sf * n0;
    c0 += (lane == 0) ? f * delta : ((lane == 1) ? omf * delta : 0.f);
    const float c1 = (lane + kChunk - 2 < L) ? so * v1 + sf * n1 : 0.f;
    const float c2 = (lane + 2 * kChunk - 2 < L) ? so * v2 + sf * n2 : 0.f;
    if (lane < kChunk) {
      const float a0 = p[0], a1 = p[kChunk], a2 = p[2 * kChunk];
      p[0] = a0 + c0;
      p[kChunk] = a1 + c1;
      p[2 * kChunk] = a2 + c2;
    }
  } else {
    scatter_chunk(accw, lane, lane - 2, scale * v0, scale * n0, omf, f, delta, S, L, nout);
    scatter_chunk(accw, lane, lane - 2 + kChunk, scale * v1, scale * n1, omf, f, delta, S, L, nout);
    scatter_chunk(accw, lane, lane - 2 + 2 * kChunk, scale * v2, scale * n2, omf, f, delta, S, L, nout);
  }
}

// lane i receives lane i+1's value, lane 63 receives 0 (DPP wave shift; bound_ctrl supplies the
// zero, so no register needs clearing first)
__device__ __forceinline__ float lane_next0(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130 /* wave_shl:1 */, 0xf, 0xf, true));
}

// Store of a neighbour-stripe window with the pair's factors precomputed: so = n_e (1-f),
// sf = n_e f, fD = f Delta, omfD = (1-f) Delta.  v0..v2 come from a padded table row (zero past the
// window), so interior stores need no mask beyond lane < 63.  base = this wave's partial row - 2.
// m0, m1: 1 on lane 0 (resp. 1), else 0; fD m0 + omfD m1 is exactly the correction of the lane.
__device__ __forceinline__ void store_stripe(float* base, int lane, float m0, float m1, float v0, float v1, float v2,
                                             float so, float sf, float fD, float omfD, int S, int L, int nout,
                                             bool interior) {
  const float n0 = lane_next0(v0), n1 = lane_next0(v1), n2 = lane_next0(v2);
  if (interior) {
    float* p = base + S + lane;
    float c0 = so * v0 + sf * n0;
    c0 += fD * m0 + omfD * m1;
    const float c1 = so * v1 + sf * n1;
    const float c2 = so * v2 + sf * n2;
    if (lane < kChunk) {
      const float a0 = p[0], a1 = p[kChunk], a2 = p[2 * kChunk];
      p[0] = a0 + c0;
      p[kChunk] = a1 + c1;
      p[2 * kChunk] = a2 + c2;
    }
  } else {
    float* accw = base + 2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = c == 0 ? v0 : (c == 1 ? v1 : v2), vn = c == 0 ? n0 : (c == 1 ? n1 : n2);
      const int m = lane - 2 + c * kChunk;
      const int o = S + m;
      if (lane < kChunk && m < L && o >= 0 && o < nout) {
        float val = 0.f;
        if (o >= 1) val = so * v + sf * vn;
        if (o <= nout - 2) {
          if (m == -1) val += omfD;
          if (m == -2) val += fD;
        }
        accw[o] += val;
      }
    }
  }
}

// Window values of one pair (and, for the centre pairs, the pair's correction): the loads of a
// group of pairs are issued before any of their stores, so a wave keeps several independent pairs
// in flight.
struct FarLoads {
  float t0, t1, t2;
  float D;  // boundary correction Delta of the pair (charge included)
};

// Neighbour-stripe window loads through a buffer descriptor of the padded raw-window table: the
// pair's byte offset goes in the scalar offset field and the lane's in the vector one, so there
// is no per-pair address arithmetic.  0x00020000: word 3 of a raw (stride 0) buffer descriptor on
// gfx9-family GPUs.
constexpr int kBufDesc3 = 0x00020000;
__device__ __forceinline__ void far_load_buf(__amdgpu_buffer_rsrc_t r, int lane4, int off, FarLoads& q) {
  q.D = 0.f;
  q.t0 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane4, off, 0));
  q.t1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane4 + 4 * kChunk, off, 0));
  q.t2 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane4 + 8 * kChunk, off, 0));
}

// Distinct LUT cells of a segment's diffusion bins that land on pixel offset a along one axis, with
// the bins' weights summed per cell: bins on the same cell read the same template window and the
// bin weights factorise (n_bin = n_e wx wy), so each distinct (cx, cy) is visited once with weight
// W[i] W[j].  ok[i] marks the distinct entries (slot of the first bin of each cell).
template <int NB>
__device__ __forceinline__ void axis_merge(int u, int a, int nsb, const float* w, float scale, bool* ok, int* cell,
                                           float* W) {
  constexpr int SYM = (NB - 1) / 2;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    ok[i] = false;
    cell[i] = 0;
    W[i] = 0.f;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int v = u + i - SYM;
    const int dv = v < 0 ? -1 : (v >= nsb ? 1 : 0);
    if (dv != a) continue;
    const int cc = offset_bin(v - dv * nsb, 0, nsb);
    const float wi = scale * w[i];
    bool merged = false;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      if (!merged && ok[j] && cell[j] == cc) {
        W[j] += wi;
        merged = true;
      }
    }
    if (!merged) {
      ok[i] = true;
      cell[i] = cc;
      W[i] = wi;
    }
  }
}

// A segment's diffusion bins along one axis reach pixel offset a (|a| <= 1) when some bin
// position u + i - SYM, i < NB, lies in pixel a: always for a = 0, else only near the edge.
template <int NB>
__device__ __forceinline__ bool bins_reach(int u, int a, int nsb) {
  constexpr int SYM = (NB - 1) / 2;
  return a == 0 ? true : (a > 0 ? u + (NB - 1 - SYM) >= nsb : u - SYM < 0);
}

// Window of the diffusion bins of segment R that land on the pixel (a, b) from its main pixel
// (|a|, |b| <= 1), charge included: V[j] = sum_bins n_bin * (lb T_idx + la T_{idx-1} +
// lc T_{idx+1})[cx][cy][j] (sim_jax.py:186-194), at the padded positions lane + 63 k (the chunks of
// store_window), each distinct LUT cell once.  Bin i of the segment lies floor((u + i - SYM) / nsb)
// in {-1, 0, 1} pixels from the main pixel.
template <int NB>
__device__ __forceinline__ void pixel_bins(const Dev& d, const Tables& tb, const SegRec& R, int a, int b, int lane,
                                           float& v0, float& v1, float& v2) {
  const int nsb = d.nsb, cm = d.cm;
  const int ux = R.uxy & 0xffff, uy = R.uxy >> 16;
  bool okx[NB], oky[NB];
  int cxs[NB], cys[NB];
  float WX[NB], WY[NB];
  axis_merge<NB>(ux, a, nsb, R.wx, R.ne, okx, cxs, WX);
  axis_merge<NB>(uy, b, nsb, R.wy, 1.f, oky, cys, WY);
  const float la = R.la, lb = R.lb, lc = R.lc;
  const float3* T3 = tb.TM3 + (int64_t)R.idx * cm * cm * kWinPad + lane;
  v0 = v1 = v2 = 0.f;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (!okx[i]) continue;
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
      if (!oky[jj]) continue;
      const float c = WX[i] * WY[jj];
      const float3* tp = T3 + ((int64_t)cxs[i] * cm + cys[jj]) * kWinPad;
      const float3 q0 = tp[0], q1 = tp[kChunk], q2 = tp[2 * kChunk];
      v0 += c * (q0.y * lb + q0.x * la + q0.z * lc);
      v1 += c * (q1.y * lb + q1.x * la + q1.z * lc);
      v2 += c * (q2.y * lb + q2.x * la + q2.z * lc);
    }
  }
}

// Dynamic LDS of the waveform kernel: the per-wave partial rows (reused afterwards for the unreset
// float64 cumulative charge of the tape), then the neighbourhood's cell lists (ncell begins, ncell + 1
// cumulative counts, ncell offsets).
__host__ __device__ inline size_t rows_acc_bytes(int waves, int nout, int L) {
  const size_t acc = (size_t)waves * (nout + row_pad(L)) * 4, q = (size_t)nout * 8;
  return ((acc > q ? acc : q) + 15) & ~(size_t)15;
}
inline size_t rows_lds_bytes(int waves, int nout, int L, int ncell) {
  return rows_acc_bytes(waves, nout, L) + (size_t)(3 * ncell + 1) * 4;
}

constexpr int kMaxCells = 17 * 17;  // (2n+1)^2 for number_pix_neighbors <= 8

// The non-empty cells of a row's neighbourhood.  Cell i = (a, b), a, b in [-nn, nn] in row-major
// order, is the segment cell of the pixel (px - a, py - b); it is listed when cls(a, b) >= 0 and
// it holds segments, ordered by (cls, i); cnt(begin, end) gives how many of the cell's leading
// segments to list (a cell with none is left out).  One thread per cell loads the cell's segment range;
// then wave 0 ranks the listed cells, 64 at a time per class, with a ballot (the place) and a wave
// scan of the counts (the segments before each).  Writes c_beg, c_ab (a + 64 << 8 | b + 64), c_cum[0..k] (exclusive sums of the
// cells' segment counts) and ncls[c] = number of listed cells of class <= c, c < 3.  scratch
// holds 3 * kMaxCells ints; the call begins and ends with every thread of the block in it.
template <int NT, class Cls, class Cnt>
__device__ __forceinline__ void list_cells(const Dev& d, const int2* __restrict__ rng, int cell, Cls cls, Cnt cnt,
                                           int* scratch, int* c_beg, int* c_cum, int* c_ab, int* ncls) {
  const int t = threadIdx.x;
  const int side = 2 * d.ne + 1, ncell = side * side;
  const int px = cell % d.gx - d.halo;
  const int py = (cell / d.gx) % d.gy - d.halo;
  const int evp = cell / (d.gx * d.gy);
  int* s_key = scratch;
  int* s_cnt = scratch + kMaxCells;
  int* s_beg = scratch + 2 * kMaxCells;
  for (int i = t; i < ncell; i += NT) {
    const int a = i / side - d.ne, b = i % side - d.ne;
    const int c = cls(a, b);
    const int qx = px - a, qy = py - b;
    int key = -1, n = 0, beg = 0;
    if (c >= 0 && qx >= -d.halo && qx < d.nx + d.halo && qy >= -d.halo && qy < d.ny + d.halo) {
      const int2 rg = rng[((int64_t)evp * d.gy + (qy + d.halo)) * d.gx + (qx + d.halo)];
      const int m = rg.y > rg.x ? cnt(rg.x, rg.y) : 0;
      if (m > 0) {
        key = c * ncell + i;
        n = m;
        beg = rg.x;
      }
    }
    s_key[i] = key;
    s_cnt[i] = n;
    s_beg[i] = beg;
  }
  __syncthreads();
  // ranks in (class, i) order: wave 0 walks the cells class by class, 64 at a time; a ballot gives
  // each listed cell its rank and a wave scan of the counts the segments before it
  if (t < kWave) {
    const int lane = t;
    int rank0 = 0, before0 = 0;
    for (int c = 0; c < 3; ++c) {
      for (int i0 = 0; i0 < ncell; i0 += kWave) {
        const int i = i0 + lane;
        const int key = i < ncell ? s_key[i] : -1;
        const bool on = key >= 0 && key / ncell == c;
        const uint64_t m = __ballot(on);
        if (m == 0) continue;
        const int n = on ? s_cnt[i] : 0;
        int inc = n;
#pragma unroll
        for (int dd = 1; dd < kWave; dd <<= 1) {
          const int u = __shfl_up(inc, dd, kWave);
          if (lane >= dd) inc += u;
        }
        if (on) {
          const int rk = rank0 + __popcll(m & ((1ull << lane) - 1));
          const int a = i / side - d.ne, b = i % side - d.ne;
          c_beg[rk] = s_beg[i];
          c_cum[rk] = before0 + (inc - n);
          c_ab[rk] = ((a + 64) << 8) | (b + 64);
        }
        rank0 += __popcll(m);
        before0 += __shfl(inc, kWave - 1, kWave);
      }
      if (lane == 0) ncls[c] = rank0;
    }
    if (lane == 0) c_cum[rank0] = before0;
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------
// work-balanced XCD bands of pixel rows for the row kernels
// ------------------------------------------------------------------------------------------
//
// Workgroup i of a launch runs on XCD i % kXcd, so each XCD is given one contiguous band of rows
// (its records and tables stay in that XCD's L2).  Rows differ in cost (the segments in their
// neighbourhood), and equal row counts per band leave some XCDs with up to 1.4x the mean work on
// uneven batches.  The bands are therefore cut at equal shares of the estimated work: segments in
// the row's (2e+1)^2 neighbourhood cells plus a fixed per-row cost (row set-up, sum, cumulative
// charge and self-trigger) in the same units.

// Processing order of the pixel rows: by tile of `tile` x `tile` cells of one event plane, in row
// order within a tile (a stable sort on the tile id).  Consecutive workgroups of an XCD then work
// on neighbouring pixels, whose neighbourhoods share most of their segments, so the segment
// records and corrections a band reads stay in that XCD's L2.  The order changes which workgroup
// computes a row and when, never what it computes (tile 1: plain row order).
inline int64_t row_tiles(const Dev& d, int tile) {
  const int64_t ntx = (d.gx + tile - 1) / tile, nty = (d.gy + tile - 1) / tile;
  return d.n_cells / ((int64_t)d.gx * d.gy) * ntx * nty;
}

// Sort keys of the rows: the tile id, or `past` for the rows P..max_rows-1 (after every tile).
__global__ void k_tile_keys(Dev d, int tile, const int* __restrict__ row_cell, const int* __restrict__ n_rows_dev,
                            int64_t max_rows, uint32_t past, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= max_rows) return;
  const int64_t P = *n_rows_dev < max_rows ? *n_rows_dev : max_rows;
  vals[r] = (uint32_t)r;
  if (r >= P) {
    keys[r] = past;
    return;
  }
  const int cell = row_cell[r];
  const int cx = cell % d.gx, cy = (cell / d.gx) % d.gy, evp = cell / (d.gx * d.gy);
  const uint32_t ntx = (uint32_t)((d.gx + tile - 1) / tile), nty = (uint32_t)((d.gy + tile - 1) / tile);
  keys[r] = ((uint32_t)evp * nty + (uint32_t)(cy / tile)) * ntx + (uint32_t)(cx / tile);
}

// Estimated work of every row, in processing order (work[s] for the row order[s]): one thread per
// slot.  Rows P..max_rows-1 hold no pixel (and keep their place in the order): their hit count is
// zeroed here for the hit scan over max_rows (the row kernel never visits them).
__global__ void k_row_work(Dev d, const int2* __restrict__ rng, const int* __restrict__ row_cell,
                           const int* __restrict__ order, const int* __restrict__ n_rows_dev, int64_t max_rows,
                           int row_cost, int* __restrict__ work, int* __restrict__ hit_cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = *n_rows_dev < max_rows ? *n_rows_dev : max_rows;
  if (s >= P) {
    if (s < max_rows) hit_cnt[s] = 0;
    return;
  }
  const int cell = row_cell[order[s]];
  const int px = cell % d.gx - d.halo;
  const int py = (cell / d.gx) % d.gy - d.halo;
  const int evp = cell / (d.gx * d.gy);
  int n = row_cost;
  for (int a = -d.ne; a <= d.ne; ++a) {
    const int qx = px - a;
    if (qx < -d.halo || qx >= d.nx + d.halo) continue;
    for (int b = -d.ne; b <= d.ne; ++b) {
      const int qy = py - b;
      if (qy < -d.halo || qy >= d.ny + d.halo) continue;
      const int2 rg = rng[((int64_t)evp * d.gy + (qy + d.halo)) * d.gx + (qx + d.halo)];
      n += rg.y - rg.x;
    }
  }
  work[s] = n;
}

// Band starts band[0..kXcd] (band[kXcd] = P) at equal shares of the total work; one workgroup of
// kBandThreads threads.  The equal-work start of band x is the first row whose exclusive work
// prefix reaches x/kXcd of the total (every row's work is >= 1, so the prefix is strictly
// increasing).  Each band then holds at most `slots` rows (the workgroups per XCD of the row
// kernel, kXcd * slots >= P): start_x is clamped to [max(start_{x-1}, P - (kXcd - x) slots),
// start_{x-1} + slots], which keeps the starts non-decreasing and leaves room for the later bands.
constexpr int kBandThreads = 1024;

__global__ __launch_bounds__(kBandThreads) void k_bands(const int* __restrict__ work, const int* __restrict__ n_rows_dev,
                                                      int64_t max_rows, int64_t slots, int* __restrict__ band) {
  __shared__ int64_t part[kBandThreads];
  __shared__ int64_t cut[kXcd + 1];
  const int t = threadIdx.x;
  const int64_t P = *n_rows_dev < max_rows ? *n_rows_dev : max_rows;
  const int64_t per = (P + kBandThreads - 1) / kBandThreads;
  const int64_t r0 = (int64_t)t * per < P ? (int64_t)t * per : P;
  const int64_t r1 = r0 + per < P ? r0 + per : P;
  int64_t s = 0;
  for (int64_t r = r0; r < r1; ++r) s += work[r];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kBandThreads; off <<= 1) {  // inclusive Hillis-Steele scan
    const int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  const int64_t total = part[kBandThreads - 1];
  const int64_t ex = part[t] - s;
  for (int x = 1; x < kXcd; ++x) {
    const int64_t target = total * x / kXcd;
    if (ex < target && target <= ex + s) {
      int64_t r = r0, e = ex;
      while (e < target) e += work[r++];
      cut[x] = r;
    }
    if (target == 0 && t == 0) cut[x] = 0;
  }
  __syncthreads();
  if (t == 0) {
    int64_t prev = 0;
    band[0] = 0;
    for (int x = 1; x < kXcd; ++x) {
      int64_t lo = P - (int64_t)(kXcd - x) * slots;
      lo = lo > prev ? lo : prev;
      const int64_t hi = prev + slots;
      int64_t v = cut[x] < lo ? lo : cut[x];
      v = v > hi ? hi : v;
      band[x] = (int)v;
      prev = v;
    }
    band[kXcd] = (int)P;
  }
}

#ifndef LSIM_FAR_GROUP
#define LSIM_FAR_GROUP 8
#endif
constexpr int kFarGroup = LSIM_FAR_GROUP;  // independent neighbour-only pairs in flight per wave

// XCD-aware, work-balanced mapping of workgroups onto rows: workgroups are dispatched round-robin
// over the XCDs, so workgroup i runs on XCD x = i % kXcd; it takes slot i / kXcd of band x (k_bands:
// contiguous slots of the tile-major processing order, cut at equal shares of the estimated work,
// so a band's records and tables stay in that XCD's L2 and every XCD finishes at about the same
// time), and computes row order[slot].  The grid holds kXcd times the
// longest band allowed (rows_grid); the workgroups past their band's end exit at once.
template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) __attribute__((amdgpu_waves_per_eu(WAVES == 2 ? 5 : 1))) void k_rows_fwd(Dev d, Tables tb, const SegRec* __restrict__ rec,
                                                            const int2* __restrict__ rng, const int* __restrict__ row_cell,
                                                            const int* __restrict__ n_rows_dev, int64_t max_rows,
                                                            const int* __restrict__ band, const int* __restrict__ order,
                                                            const float* __restrict__ dtot,
                                                            const PairRec* __restrict__ prs, int64_t nseg, RowOut out) {
  constexpr int NT = WAVES * kWave;
  extern __shared__ __align__(16) char smem[];
  __shared__ double shd[NT / kWave];
  __shared__ int shi[NT / kWave];
  __shared__ int ncls[3];  // far cells, + the centre cell, + the 8 side cells

  const int P0 = *n_rows_dev;
  const int P = P0 < max_rows ? P0 : (int)max_rows;
  const int xcd = (int)(blockIdx.x % kXcd);
  const int slot = band[xcd] + (int)(blockIdx.x / kXcd);
  if (slot >= band[xcd + 1] || slot >= P) return;  // past the band (the bound on P is a guard only)
  const int r = order[slot];
  const int nout = d.nout;
  float* acc = (float*)smem;
  // this row's neighbourhood cells: far cells first, then the 3x3 near cells, in (a, b) order
  const int ncell = (2 * d.ne + 1) * (2 * d.ne + 1);
  int* const c_beg = (int*)(smem + rows_acc_bytes(WAVES, nout, d.L));
  int* const c_cum = c_beg + ncell;
  int* const c_ab = c_cum + ncell + 1;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
  const int rstride = nout + row_pad(d.L);
  const int cell = row_cell[r];
  const int evp = cell / (d.gx * d.gy);  // ev * ntpc + plane
  const int plane = evp % d.ntpc;
  if (t == 0) {
    const int px = cell % d.gx - d.halo;
    const int py = (cell / d.gx) % d.gy - d.halo;
    out.pixel_ids[r] = ((int64_t)evp * d.ny + py) * d.nx + px;
  }
  // far cells first, then the 8 side cells, then the centre cell (the partial rows are scratch)
  list_cells<NT>(
      d, rng, cell,
      [](int a, int b) {
        const bool near = a >= -1 && a <= 1 && b >= -1 && b <= 1;
        return !near ? 0 : ((a == 0 && b == 0) ? 2 : 1);
      },
      [](int b, int e) { return e - b; }, reinterpret_cast<int*>(acc), c_beg, c_cum, c_ab, ncls);
  {
    // 16-byte stores (smem is 16-byte aligned), then the remainder
    const int nz = WAVES * rstride, n4 = nz / 4;
    for (int o = t; o < n4; o += NT) reinterpret_cast<float4*>(acc)[o] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int o = 4 * n4 + t; o < nz; o += NT) acc[o] = 0.f;
  }
  __syncthreads();

  float* accw = acc + wid * rstride;
  const int nsc = ncls[1], nct = ncls[2];  // far + side cells, all cells
  // Neighbour stripes of the far and side pairs, flattened over those cells in blocks of 64: wave
  // w takes blocks w, w + WAVES, ...  Lane l loads pair 64 b + l's record, boundary correction and
  // row offset into registers (one memory round trip per block instead of one per pair group),
  // then the wave walks the block kFarGroup pairs at a time, reading each pair's values from its
  // lane.  A side pair's correction covers its diffusion bins too; the bins that land on a side
  // row (segments near the pixel edge only) follow per block, picked out by a ballot.
  {
    const int nfar = c_cum[nsc];
    const int W4 = kWinPad * 4, nsb = d.nsb, side = 2 * d.ne + 1;
    const __amdgpu_buffer_rsrc_t tnp =
        __builtin_amdgcn_make_buffer_rsrc((void*)tb.TNP, 0, d.lx * d.ly * kWinPad * 4, kBufDesc3);
    // lane l's pair of the block at b0: its cell (a binary search of the cell list), record and
    // correction.  The loads of the wave's next block are issued before this block's stores, so
    // they arrive while the pairs are stored.
    struct BlockIn {
      PairRec P;
      float D;
      int sg, ab;
    };
    auto fetch = [&](int b0, BlockIn& B) {
      const int p = b0 + lane;
      if (p < nfar) {
        int lo = 0, hi = nsc - 1;  // the cell holding p: last c with c_cum[c] <= p
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (c_cum[mid] <= p) lo = mid;
          else hi = mid - 1;
        }
        B.sg = c_beg[lo] + (p - c_cum[lo]);
        B.ab = c_ab[lo];
        const int a = (B.ab >> 8) - 64, b = (B.ab & 255) - 64;
        B.P = prs[B.sg];
        B.D = dtot[(int64_t)((a + d.ne) * side + (b + d.ne)) * nseg + B.sg];
      }
    };
    BlockIn cur, nxt;
    if (wid * kWave < nfar) fetch(wid * kWave, cur);
    for (int b0 = wid * kWave; b0 < nfar; b0 += WAVES * kWave) {
      if (b0 + WAVES * kWave < nfar) fetch(b0 + WAVES * kWave, nxt);
      const int p = b0 + lane;
      // padding lanes: an interior window at S = 2 with zero charge and correction (adds +0)
      int rS = (2 << 1) | 1, roff = 0, rsg = 0, rab = 0;
      float rso = 0.f, rsf = 0.f, rfD = 0.f, romfD = 0.f;
      bool bins = false;
      if (p < nfar) {
        const int sg = cur.sg, ab = cur.ab;
        const int a = (ab >> 8) - 64, b = (ab & 255) - 64;
        // row offset = cell constant + ma U + za U0 + mb V + zb V0 (bytes)
        int kc = 0, ma = 0, za = 0, mb = 0, zb = 0;
        if (a > 0) { kc += (nsb * a + nsb / 2 - 1) * d.ly * W4; ma = -1; }
        else if (a < 0) { kc += (nsb * -a - nsb / 2) * d.ly * W4; ma = 1; }
        else za = 1;
        if (b > 0) { kc += (nsb * b + nsb / 2 - 1) * W4; mb = -1; }
        else if (b < 0) { kc += (nsb * -b - nsb / 2) * W4; mb = 1; }
        else zb = 1;
        const PairRec& P = cur.P;
        const float D = cur.D;
        // with no neighbour stripes (nn = 0) a side pixel gets its diffusion bins only
        const float st = a >= -d.nn && a <= d.nn && b >= -d.nn && b <= d.nn ? 1.f : 0.f;
        roff = kc + ma * P.U + za * P.U0 + mb * P.V + zb * P.V0;
        rS = (P.S << 1) | (P.flags & 1);
        const float omf = 1.f - P.f;
        rso = st * (P.ne * omf);
        rsf = st * (P.ne * P.f);
        rfD = P.f * D;
        romfD = omf * D;
        rsg = sg;
        rab = ab;
        bins = a >= -1 && a <= 1 && b >= -1 && b <= 1 && P.ne != 0.f && bins_reach<NB>((P.flags >> 8) & 255, a, nsb) &&
               bins_reach<NB>((P.flags >> 16) & 255, b, nsb);
      }
      const int n = __builtin_amdgcn_readfirstlane(nfar - b0 < kWave ? nfar - b0 : kWave);
      float* base = accw - 2;
      const float m0 = lane == 0 ? 1.f : 0.f, m1 = lane == 1 ? 1.f : 0.f;
      // pairs i .. i + kFarGroup - 1 of the block: their factors from the owning lanes, their window
      // loads issued (j < 64: i < n <= 64 and i is a multiple of kFarGroup)
      struct FarGroup {
        FarLoads q[kFarGroup];
        int S[kFarGroup];
        float so[kFarGroup], sf[kFarGroup], fD[kFarGroup], omfD[kFarGroup];
      };
      auto issue = [&](int i, FarGroup& G) {
#pragma unroll
        for (int g = 0; g < kFarGroup; ++g) {
          const int j = i + g;
          G.S[g] = __builtin_amdgcn_readlane(rS, j);
          G.so[g] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rso), j));
          G.sf[g] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rsf), j));
          G.fD[g] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rfD), j));
          G.omfD[g] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(romfD), j));
          far_load_buf(tnp, lane * 4, __builtin_amdgcn_readlane(roff, j), G.q[g]);
        }
      };
      // unconditional stores: a branch on the charge would let the compiler sink a pair's loads
      // into it and wait on them there instead of overlapping them with the other pairs'
      auto drain = [&](const FarGroup& G) {
#pragma unroll
        for (int g = 0; g < kFarGroup; ++g)
          store_stripe(base, lane, m0, m1, G.q[g].t0, G.q[g].t1, G.q[g].t2, G.so[g], G.sf[g], G.fD[g], G.omfD[g],
                       G.S[g] >> 1, d.L, d.nout, (G.S[g] & 1) != 0);
      };
      // (issuing the next group's loads before this group's stores measured no faster on the
      // MI355X: the loads are not what the waves wait on; 8-pair groups are 1 % faster than 4)
      for (int i = 0; i < n; i += kFarGroup) {
        FarGroup A;
        issue(i, A);
        drain(A);
      }
      for (uint64_t m = __ballot(bins); m != 0; m &= m - 1) {
        const int j = __builtin_ctzll(m);
        const int ab = __builtin_amdgcn_readlane(rab, j);
        const SegRec R = rec[__builtin_amdgcn_readlane(rsg, j)];
        float v0, v1, v2;
        pixel_bins<NB>(d, tb, R, (ab >> 8) - 64, (ab & 255) - 64, lane, v0, v1, v2);
        const int S = d.nt - d.L - R.k0;
        store_window(accw, lane, v0, v1, v2, 1.f, 1.f - R.f, R.f, 0.f, S, d.L, d.nout, S >= 2 && S + d.L <= d.nout);
      }
      cur = nxt;
    }
  }
  // the centre cell: the window of each segment's bins on this pixel, from the interpolated
  // templates of its distinct LUT cells, one segment per wave at a time
  const float* dtc = dtot + (int64_t)(d.ne * (2 * d.ne + 1) + d.ne) * nseg;
  for (int c = nsc; c < nct; ++c) {
    const int beg = __builtin_amdgcn_readfirstlane(c_beg[c]);
    const int cnt = __builtin_amdgcn_readfirstlane(c_cum[c + 1] - c_cum[c]);
    for (int k = wid; k < cnt; k += WAVES) {
      const int sg = beg + k;
      float v0, v1, v2;
      pixel_bins<NB>(d, tb, rec[sg], 0, 0, lane, v0, v1, v2);
      const PairRec P = prs[sg];
      store_window(accw, lane, v0, v1, v2, 1.f, 1.f - P.f, P.f, dtc[sg], P.S, d.L, d.nout, (P.flags & 1) != 0);
    }
  }
  __syncthreads();
  double q[kMaxOut / NT];
  double* Q0 = (double*)smem;
  finish_row<NT>(d, acc, WAVES, rstride, q, Q0, r, out, shd);
  fee_row<NT>(d, q, Q0, r, plane, out, shi);
}

// Self-trigger on caller waveforms (simulate_stochastic on an arbitrary wfs): one workgroup per
// row, the same finish/fee code as the fused path.
template <int NT>
__global__ __launch_bounds__(NT) void k_fee_rows(Dev d, const float* __restrict__ wfs, const int64_t* __restrict__ pids,
                                                 int64_t n_rows, RowOut out) {
  __shared__ __align__(16) char smem[kMaxOut * 8];
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
  double q[kMaxOut / NT];
  double* Q0 = (double*)smem;
  RowOut o2 = out;
  o2.wfs = nullptr;
  finish_row<NT>(d, accs, 1, d.nout, q, Q0, (int)r, o2, shd);
  fee_row<NT>(d, q, Q0, (int)r, plane, o2, shi);
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
  if (pos < 0 || pos >= P * H) return;  // offsets come from a scan of per-row counts <= H
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
// 4b. negative pixel ids of the reference's pixel set
// ------------------------------------------------------------------------------------------

// The reference's pixel set is the sorted unique main-pixel ids plus an appended -1
// (sim_jax.py:716-721).  Besides the ids >= 0 (the rows built here) it holds negative ids:
// a padding segment (eventID -1, optimize/dataio.py:347-359) on the pixel plane gets the id
// pixel2id packs for event -1 (detsim_jax.py:230-244), and a segment whose main pixel is off the
// plane gets -1.  Those rows carry no charge and no hits, so only their ids are produced here.
// Every such id lies in [-ntpc ny nx, -1]; bit (id + ntpc ny nx) of `bits` marks it (integer
// atomics: the bit set does not depend on the order).  eventID < -1 is invalid input and skipped.
__global__ void k_neg_mark(Dev d, SegIn in, uint32_t* __restrict__ bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in.n) return;
  const int ev = to_int_clamped((double)ldcol(in.ev, i, in.stride));
  if (ev < -1) return;
  Phys ph;
  seg_physics(d, in, i, ph);
  const bool on_plane = ph.mpx >= 0 && ph.mpx < d.nx && ph.mpy >= 0 && ph.mpy < d.ny;
  if (ev >= 0 && on_plane) return;
  const int64_t span = (int64_t)d.ntpc * d.ny * d.nx;
  const int64_t b = on_plane ? pixel_id(d, -1, ph.plane, ph.mpx, ph.mpy) + span : span - 1;
  if (b < 0 || b >= span) return;
  atomicOr(&bits[b >> 5], 1u << (b & 31));
}

// The marked ids in ascending order (one workgroup of kPadThreads threads: each takes a
// contiguous run of bitmap words, counts its bits, and writes its ids after a block scan).
constexpr int kPadThreads = 1024;

__global__ __launch_bounds__(kPadThreads) void k_pad_list(const uint32_t* __restrict__ bits, int64_t span,
                                                         int64_t* __restrict__ ids, int64_t max_ids,
                                                         int64_t* __restrict__ n_out) {
  __shared__ int64_t part[kPadThreads];
  const int t = threadIdx.x;
  const int64_t words = (span + 31) >> 5;
  const int64_t per = (words + kPadThreads - 1) / kPadThreads;
  const int64_t w0 = (int64_t)t * per, w1 = w0 + per < words ? w0 + per : words;
  int64_t cnt = 0;
  for (int64_t w = w0; w < w1; ++w) cnt += __popc(bits[w]);
  part[t] = cnt;
  __syncthreads();
  for (int off = 1; off < kPadThreads; off <<= 1) {  // inclusive Hillis-Steele scan
    const int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t pos = part[t] - cnt;
  for (int64_t w = w0; w < w1; ++w) {
    uint32_t m = bits[w];
    while (m) {
      const int k = __ffs(m) - 1;
      m &= m - 1;
      if (pos < max_ids) ids[pos] = w * 32 + k - span;
      ++pos;
    }
  }
  if (t == kPadThreads - 1) *n_out = part[t];
}

// ------------------------------------------------------------------------------------------
// 5. response tables (consts_jax.py:427-447), float64 arithmetic
// ------------------------------------------------------------------------------------------

// One workgroup per (template m, cx, cy) series: template values over the full time range,
// prefix sums, then the window, boundary-difference and window-prefix tables.
// main != 0: m in [0, nm), (cx, cy) in [0, cm)^2 -> TM, DM, CTM; else the raw response over all
// (cx, cy) -> TN, DN, CTN.
__global__ __launch_bounds__(256) void k_lut_tables(Dev d, const float* __restrict__ raw, int main_tables, float* __restrict__ Tw,
                                                    float* __restrict__ Twp, float2* __restrict__ Dt,
                                                    double* __restrict__ CT) {
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
  if (Twp)
    for (int p = threadIdx.x; p < kWinPad; p += blockDim.x)
      Twp[sbase * kWinPad + p] = (p >= 2 && p < L + 2) ? win[p - 2] : 0.f;
  const double cref = T[nt - L];
  for (int k = threadIdx.x; k < nt; k += blockDim.x) {
    const double c0 = T[k];
    const double c1 = T[k + 1 < nt ? k + 1 : nt - 1];
    Dt[(int64_t)k * gridDim.x + sbase] = make_float2((float)(cref - c0), (float)(c1 - c0));  // tick-major
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

// Padded (T_{m-1}, T_m, T_{m+1}) windows for m in [1, nm-2]; other slabs are zero.
// DNS[k][ux][uy][ab] = DN[k][offset_bin(ux, a)][offset_bin(uy, b)] for the (2 ne + 1)^2 row offsets
// (a, b) = (ab / W - ne, ab % W - ne) of a segment in sub-bin (ux, uy): the raw response's D entries
// of one segment's neighbour stripes, contiguous.
__global__ void k_lut_dns(Dev d, const float2* __restrict__ DN, float2* __restrict__ DNS) {
  const int W = 2 * d.ne + 1, WW = W * W, nsb = d.nsb;
  const int64_t total = (int64_t)d.nt * nsb * nsb * WW;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
    const int ab = (int)(j % WW);
    const int64_t q = j / WW;
    const int uy = (int)(q % nsb), ux = (int)((q / nsb) % nsb);
    const int64_t k = q / ((int64_t)nsb * nsb);
    const int a = ab / W - d.ne, b = ab % W - d.ne;
    DNS[j] = DN[(k * d.lx + offset_bin(ux, a, nsb)) * d.ly + offset_bin(uy, b, nsb)];
  }
}

__global__ void k_lut_tm3(Dev d, const float* __restrict__ TM, float3* __restrict__ TM3) {
  const int64_t cc = (int64_t)d.cm * d.cm;
  const int64_t n = cc * d.nm * kWinPad;
  const int64_t per = cc * d.L;  // one template slab of TM
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(i % kWinPad);
    const int64_t series = i / kWinPad;  // m * cm * cm + cx * cm + cy
    const int64_t m = series / cc;
    float3 v = make_float3(0.f, 0.f, 0.f);
    if (m >= 1 && m <= d.nm - 2 && p >= 2 && p < d.L + 2) {
      const int64_t t = series * d.L + (p - 2);
      v = make_float3(TM[t - per], TM[t], TM[t + per]);
    }
    TM3[i] = v;
  }
}

}  // namespace lsim
