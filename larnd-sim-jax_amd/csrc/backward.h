// Adjoint of the LUT chain for the 7 physics parameters (Ab, kb, eField, lifetime, vdrift,
// long_diff, tran_diff).  Integer decisions are held at their forward values, as reverse-mode
// AD through the reference does (floor/astype/searchsorted/argmax carry no gradient).
//
// Stage 1 (one thread per pixel row): the self-trigger adjoint.  The ADC value of pass k is
// Q^(k)[e_k]; pass k then sets Q^(k+1) = max(Q^(k) - Q^(k)[g_k], 0).  Reverse mode over the 2H
// tracked points gives the adjoint of Q^(0), and because Q^(0) is the cumulative sum of
// t_sampling * wfs, d loss / d wfs is piecewise constant with at most 2H steps.
// Stage 2 (one workgroup per pixel row, like the forward's waveform kernel): for every segment
// whose stripes reach the row, inner products of the row's step function with the segment's
// template windows come from window prefix sums; they give d loss / d (charge, transverse width,
// fractional tick, longitudinal width) of the pair, which the per-segment chain factors
// (k_adj_records) turn into the 7 gradients; a fixed-order block reduction writes per-row
// partial gradients.
// Stage 3: fixed-order reduction of the partials (and of the hit-z term) to 7 values.
#pragma once
#include "forward.h"

namespace lsim {

constexpr int kNGrad = 7;
#ifndef LSIM_ADJ_WAVES
#define LSIM_ADJ_WAVES 4
#endif
constexpr int kAdjWaves = LSIM_ADJ_WAVES;  // waves per pixel-row workgroup of the adjoint

__device__ __forceinline__ double readlane_d(double v, int j) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
  return __hiloint2double(hi, lo);
}

// One wave per row (kFeeRows rows per workgroup).  Lane j < 2H follows tracked point j: the reverse
// pass over the H self-trigger passes keeps A[j] = d loss / d Q at point j in lane j, each pass's
// masked sum taken in point order; the points are then merged by tick (lane-parallel: the first
// lane of a tick owns it and sums the lanes of that tick in lane order), and lane 0 turns the
// per-tick jumps into the step function (suffix sums, equal neighbours merged, zero tail dropped).
constexpr int kFeeRows = 4;

__global__ __launch_bounds__(kFeeRows * kWave) void k_fee_adjoint(Dev d, const FeeTape* __restrict__ tape,
                                                                  const float* __restrict__ d_adc,
                                                                  const float* __restrict__ d_z,
                                                                  const int64_t* __restrict__ n_rows_dev, int64_t max_rows,
                                                                  Pieces* __restrict__ pieces, double* __restrict__ dz_part) {
  __shared__ int spos[kFeeRows][2 * kMaxH];
  __shared__ double sval[kFeeRows][2 * kMaxH];
  __shared__ int sx[kFeeRows][2 * kMaxH];
  __shared__ float sc[kFeeRows][2 * kMaxH];
  __shared__ int sn[kFeeRows];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int64_t r = (int64_t)blockIdx.x * kFeeRows + wv;
  int64_t P = *n_rows_dev;
  if (P > max_rows) P = max_rows;
  if (r >= max_rows) return;
  if (r >= P) {
    if (lane == 0) dz_part[r] = 0.0;
    return;
  }
  const FeeTape* tp = tape + r;
  const int H = d.H, np2 = 2 * H;
  const bool act = lane < np2;
  // everything the passes read, loaded at once: lane k < H holds pass k's mask and ADC term
  const int p = act ? (int)tp->pt[lane] : 0x7fffffff;
  const int mk = lane < H ? (int)tp->mask[lane] : 0;
  const double gk = lane < H ? (double)d_adc[r * H + lane] * (double)tp->slope[lane] : 0.0;
  double zk = 0.0;
  if (d_z && lane < H) {
    const int pl = tp->plane;
    const double za = d.tpc[pl][2][0], zc = d.tpc[pl][2][1];
    const double sgn = (zc - za) > 0 ? 1.0 : ((zc - za) < 0 ? -1.0 : 0.0);
    zk = (double)d_z[r * H + lane] * (double)tp->tick[lane] * d.ts * sgn;
  }
  double A = 0.0;
  for (int k = H - 1; k >= 0; --k) {
    const uint32_t m = (uint32_t)__builtin_amdgcn_readlane(mk, k);
    double s = 0.0;
    for (int j = 0; j < np2; ++j) {
      const double aj = readlane_d(A, j);
      if ((m >> j) & 1u) s += aj;
    }
    if (act && !((m >> lane) & 1u)) A = 0.0;
    if (lane == H + k) A -= s;
    if (lane == k) A += gk;
  }
  double dz = 0.0;  // d loss / d vdrift_eff through the hit z positions, summed in pass order
  for (int k = 0; k < H; ++k) dz += readlane_d(zk, k);
  // merge the points by tick
  bool owner = act;
  double val = 0.0;
  for (int j = 0; j < np2; ++j) {
    const int pj = __builtin_amdgcn_readlane(p, j);
    const double aj = readlane_d(A, j);
    if (act && pj == p) {
      val += aj;
      if (j < lane) owner = false;
    }
  }
  int rank = 0;
  for (int j = 0; j < np2; ++j) {
    const int pj = __builtin_amdgcn_readlane(p, j);
    const int oj = __builtin_amdgcn_readlane(owner ? 1 : 0, j);
    if (oj && pj < p) ++rank;
  }
  if (owner) {
    spos[wv][rank] = p;
    sval[wv][rank] = val;
  }
  const int n = __popcll(__ballot(owner));
  // wave-local LDS: stores and loads of one wave complete in program order; the wave barriers
  // keep the compiler from moving them across
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    // step values G on (pos[q-1], pos[q]]: suffix sums of the jumps
    double suf = 0.0;
    for (int q = n - 1; q >= 0; --q) {
      suf += sval[wv][q];
      sc[wv][q] = (float)(suf * d.ts);
    }
    // keep only the steps where G changes, then drop a zero tail
    int m = 0;
    float last = 0.f;
    for (int q = 0; q < n; ++q) {
      const float cq = sc[wv][q];
      if (m > 0 && cq == last) {
        sx[wv][m - 1] = spos[wv][q];  // same value: extend the previous piece
      } else {
        sx[wv][m] = spos[wv][q];
        sc[wv][m] = cq;
        last = cq;
        ++m;
      }
    }
    while (m > 0 && sc[wv][m - 1] == 0.f) --m;
    sn[wv] = m;
    dz_part[r] = dz;
  }
  __builtin_amdgcn_wave_barrier();
  const int mm = __builtin_amdgcn_readfirstlane(sn[wv]);
  Pieces* pc = pieces + r;
  if (lane == 0) {
    pc->n = mm;
    pc->pad[0] = pc->pad[1] = pc->pad[2] = 0;
  }
  if (lane < 2 * kMaxH) {
    pc->x[lane] = lane < mm ? sx[wv][lane] : 0x7fffffff;
    pc->c[lane] = lane < mm ? sc[wv][lane] : 0.f;
  }
}

// Per-segment factors of the chain rule, in the same (sorted) order as the SegRec records.
struct __align__(16) AdjRec {
  double fN[4];          // d n_e / d (Ab, kb, eField, lifetime)
  double fS[2];          // d sigma_T / d (eField, tran_diff)
  double fL[2];          // d lambda / d (eField, long_diff)
  double fF;             // d f / d eField
  double dl[3];          // d (a, b, c) / d lambda of the Lagrange weights
  float dwx[5], dwy[5];  // d w / d sigma_T of the transverse weights
  float pad[2];
};
static_assert(sizeof(AdjRec) == 144, "AdjRec layout");

// One thread per record: recompute the segment's physics (float64) and store the chain factors.
__global__ __launch_bounds__(256) void k_adj_records(Dev d, SegIn in, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, int64_t n,
                                                     AdjRec* __restrict__ adj) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || keys[s] >= (uint32_t)d.n_cells) return;
  Phys ph;
  seg_physics(d, in, vals[s], ph);
  double wx[kMaxNb], wy[kMaxNb], dwx[kMaxNb], dwy[kMaxNb];
  tran_weights(d, ph.x0, ph.st, wx, dwx);
  tran_weights(d, ph.y0, ph.st, wy, dwy);
  const double v = d.v, dv = d.dv;
  const double att = exp(-ph.td / d.lifetime) * (double)ph.mask;
  AdjRec A;
  A.fN[0] = ph.dq_dAb * att;
  A.fN[1] = ph.dq_dkb * att;
  A.fN[2] = ph.dq_dE_rec * att + ph.ne * ph.td * dv / (v * d.lifetime);
  A.fN[3] = ph.ne * ph.td / (d.lifetime * d.lifetime);
  A.fS[0] = -ph.st * dv / (2 * v);
  A.fS[1] = d.tran_diff > 0 ? ph.st / (2 * d.tran_diff) : 0.0;
  A.fL[0] = -1.5 * ph.lam * dv / v;
  A.fL[1] = d.long_diff > 0 ? ph.lam / (2 * d.long_diff) : 0.0;
  A.fF = -ph.tau * dv / v;
  const double s0 = d.sig[ph.idx - 1], s1 = d.sig[ph.idx], s2 = d.sig[ph.idx + 1];
  const double l = ph.lam;
  A.dl[0] = (2 * l - s1 - s2) / ((s0 - s1) * (s0 - s2));
  A.dl[1] = (2 * l - s0 - s2) / ((s1 - s0) * (s1 - s2));
  A.dl[2] = (2 * l - s0 - s1) / ((s2 - s0) * (s2 - s1));
#pragma unroll
  for (int k = 0; k < kMaxNb; ++k) {
    A.dwx[k] = k < d.nb ? (float)dwx[k] : 0.f;
    A.dwy[k] = k < d.nb ? (float)dwy[k] : 0.f;
  }
  A.pad[0] = A.pad[1] = 0.f;
  adj[s] = A;
}

// The row's step function G in LDS: breakpoints x[q] (ascending) and values c[q], G(o) = c[q] on
// (x[q-1], x[q]] and 0 past x[np-1]; jumps dc[q] = c[q] - c[q+1] (c[np] = 0); and qlut[o] = the first
// q with x[q] >= o, for 0 <= o < nout (np when none).
struct RowG {
  const int* x;
  const float* c;
  const float* dc;
  const uint16_t* qlut;
  int np;
  __device__ __forceinline__ double at(int o) const {  // G(o), 0 <= o < nout
    const int q = qlut[o];
    return q < np ? (double)c[q] : 0.0;
  }
};

// Window placement of one (segment, row) pair (sim_jax.py:170-261): (1-f) V at o = S + j, f V at
// o = S - 1 + j, (1-f) Delta at S - 1 and f Delta at S - 2, S = nt - L - k0; G's values at the two
// correction ticks and the tick ranges of the two windows kept in the row.
struct PairCtx {
  int S, lo1, hi1, lo0, hi0;
  bool interior;  // both windows whole: lo1 = S, hi1 = S + L - 1, lo0 = S - 1, hi0 = S + L - 2
  double f, ne, gc0, gc1, gc;
};

// Window dot products of one (segment, row) pair, telescoped over G's breakpoints.  With
// CX[k] = sum_{j<k} X[j] and F(o) = CX[o - S + sh + 1],
//   sum_{o=lo}^{hi} G(o) X[o - S + sh] = G(hi) F(hi) - G(lo) F(lo - 1) + sum_{lo <= x_q < hi} dc_q F(x_q),
// so each breakpoint inside the window costs one prefix-sum load per table.  y1: the (1-f) window
// (sh = 0, o in [lo1, hi1]); y0: the f window (sh = 1, o in [lo0, hi0]).  NT tables at stride ts.
// Whole windows take fewer loads: F(lo - 1) = CX[0] = 0, both F(hi) are CX[L], and a breakpoint
// needs the adjacent CX[x_q - S + 1] (y1) and CX[x_q - S + 2] (y0).
template <int NT>
__device__ __forceinline__ void g_dots(const RowG& G, const double* __restrict__ CX, int64_t ts, const PairCtx& c, int L,
                                       double* y1, double* y0) {
  const int S = c.S, lo1 = c.lo1, hi1 = c.hi1, lo0 = c.lo0, hi0 = c.hi0;
  if (c.interior) {
    const double gh1 = G.at(hi1), gh0 = G.at(hi0);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const double cl = CX[k * ts + L];
      y1[k] = gh1 * cl;
      y0[k] = gh0 * cl;
    }
    const int q1 = G.qlut[hi1];  // breakpoints x_q < hi1
    for (int q = G.qlut[lo0]; q < q1; ++q) {
      const int xq = G.x[q];
      const double dq = (double)G.dc[q];
      const int i = xq - S + 1;  // 0 <= i <= L - 1
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const double v1 = CX[k * ts + i], v0 = CX[k * ts + i + 1];
        if (xq >= lo1) y1[k] += dq * v1;
        if (xq < hi0) y0[k] += dq * v0;
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < NT; ++k) y1[k] = y0[k] = 0.0;
  const bool w1 = lo1 <= hi1, w0 = lo0 <= hi0;
  if (w1) {
    const double gh = G.at(hi1), gl = G.at(lo1);
#pragma unroll
    for (int k = 0; k < NT; ++k) y1[k] = gh * CX[k * ts + hi1 - S + 1] - gl * CX[k * ts + lo1 - S];
  }
  if (w0) {
    const double gh = G.at(hi0), gl = G.at(lo0);
#pragma unroll
    for (int k = 0; k < NT; ++k) y0[k] = gh * CX[k * ts + hi0 - S + 2] - gl * CX[k * ts + lo0 - S + 1];
  }
  if (!w1 && !w0) return;
  const int lo = lo0 < lo1 ? lo0 : lo1;
  const int q1 = G.qlut[hi1 > 0 ? hi1 : 0];  // breakpoints x_q < hi1 (hi0 <= hi1)
  for (int q = G.qlut[lo]; q < q1; ++q) {
    const int xq = G.x[q];
    const double dq = (double)G.dc[q];
    if (w1 && xq >= lo1) {
#pragma unroll
      for (int k = 0; k < NT; ++k) y1[k] += dq * CX[k * ts + xq - S + 1];
    }
    if (w0 && xq >= lo0 && xq < hi0) {
#pragma unroll
      for (int k = 0; k < NT; ++k) y0[k] += dq * CX[k * ts + xq - S + 2];
    }
  }
}

// Distinct LUT cells of the diffusion bins that land on the row along one axis (pixel offset a):
// bins with the same cell have their weights summed (the bin weights factorise, w = wx wy, and
// every other factor depends on the cell only), so each distinct (cx, cy) is visited once.
template <int NB>
__device__ __forceinline__ void axis_cells(int u, int a, int nsb, const float* w, const float* dw, bool* ok, int* cell,
                                           float* W, float* DW) {
  constexpr int SYM = (NB - 1) / 2;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    ok[i] = false;
    cell[i] = 0;
    W[i] = 0.f;
    DW[i] = 0.f;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int v = u + i - SYM;
    const int dv = v < 0 ? -1 : (v >= nsb ? 1 : 0);
    if (dv != a) continue;
    const int cc = offset_bin(v - dv * nsb, 0, nsb);
    bool merged = false;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      if (!merged && ok[j] && cell[j] == cc) {
        W[j] += w[i];
        DW[j] += dw[i];
        merged = true;
      }
    }
    if (!merged) {
      ok[i] = true;
      cell[i] = cc;
      W[i] = w[i];
      DW[i] = dw[i];
    }
  }
}

// G vanishes from tick S - 2 on: the pair has no gradient
__device__ __forceinline__ bool pair_dead(int S, const RowG& G) { return G.qlut[S - 2 > 0 ? S - 2 : 0] >= G.np; }

// (window start S, charge, fractional tick) -> the pair's placement; G must not vanish at S - 2
__device__ __forceinline__ void pair_ctx(const Dev& d, int S, float ne, float f, const RowG& G, PairCtx& c) {
  const int L = d.L, nout = d.nout;
  c.S = S;
  c.f = (double)f;
  c.ne = (double)ne;
  c.gc1 = (S - 2 >= 0 && S - 2 <= nout - 2) ? G.at(S - 2) : 0.0;
  c.gc0 = (S - 1 >= 0 && S - 1 <= nout - 2) ? G.at(S - 1) : 0.0;
  c.gc = (1.0 - c.f) * c.gc0 + c.f * c.gc1;
  c.lo1 = S > 1 ? S : 1;
  c.hi1 = (S + L - 1) < (nout - 1) ? (S + L - 1) : (nout - 1);
  c.lo0 = (S - 1) > 1 ? (S - 1) : 1;
  c.hi0 = (S + L - 2) < (nout - 1) ? (S + L - 2) : (nout - 1);
  c.interior = c.lo1 == S && c.hi1 == S + L - 1 && c.lo0 == S - 1 && c.hi0 == S + L - 2;
}

// The neighbour stripe of pixel offset (a, b) != (0, 0) of a segment at sub-bins (ux, uy):
// d loss / d n_e (gN) and / d f (gF).
__device__ __forceinline__ void stripe_adjoint(const Dev& d, const Tables& tb, int ux, int uy, const PairCtx& c, int a,
                                               int b, const RowG& G, double& gN, double& gF) {
  const int L = d.L, nsb = d.nsb;
  const int64_t sidx = (int64_t)offset_bin(ux, a, nsb) * d.ly + offset_bin(uy, b, nsb);
  double y1, y0;
  g_dots<1>(G, tb.CTN + sidx * (L + 1), 0, c, L, &y1, &y0);
  const float2 D = tb.DN[(int64_t)(d.nt - L - c.S) * d.lx * d.ly + sidx];  // tick-major, k0 = nt - L - S
  const double corr = (double)D.x - c.f * (double)D.y, dy = (double)D.y;
  gN += (1.0 - c.f) * y1 + c.f * y0 + c.gc * corr;
  gF += c.ne * (y0 - y1 + (c.gc1 - c.gc0) * corr - c.gc * dy);
}

// The diffusion bins of x-bin i (and every later x-bin with the same LUT cell, merged into it)
// landing on pixel offset (a, b), |a|, |b| <= 1: d loss / d (n_e, f, sigma_T, lambda) summed over
// the y bins (only y-bin jsel when jsel >= 0).  false (nothing added) when bin i misses column a
// or an earlier x-bin has its cell.
template <int NB>
__device__ __forceinline__ bool xbin_adjoint(const Dev& d, const Tables& tb, const SegRec& R, const AdjRec* __restrict__ Ap,
                                             const PairCtx& c, int a, int b, int i, int jsel, const RowG& G, double& gN,
                                             double& gF, double& gS, double& gL) {
  constexpr int SYM = (NB - 1) / 2;
  const int L = d.L, nsb = d.nsb, cm = d.cm;
  const int ux = R.uxy & 0xffff, uy = R.uxy >> 16;
  const int vi = ux + i - SYM;
  const int di = vi < 0 ? -1 : (vi >= nsb ? 1 : 0);
  if (di != a) return false;
  const int cx = offset_bin(vi - di * nsb, 0, nsb);
  float WXi = 0.f, DWXi = 0.f;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int v = ux + k - SYM;
    const int dv = v < 0 ? -1 : (v >= nsb ? 1 : 0);
    if (dv != a || offset_bin(v - dv * nsb, 0, nsb) != cx) continue;
    if (k < i) return false;  // merged into x-bin k
    WXi += R.wx[k];
    DWXi += Ap->dwx[k];
  }
  bool oky[NB];
  int cyv[NB];
  float WY[NB], DWY[NB];
  axis_cells<NB>(uy, b, nsb, R.wy, Ap->dwy, oky, cyv, WY, DWY);
  const int64_t tstride = (int64_t)cm * cm * (L + 1);
  const double la = (double)R.la, lb = (double)R.lb, lc = (double)R.lc;
  const double da = Ap->dl[0], db = Ap->dl[1], dc = Ap->dl[2];
  const double f = c.f, ne = c.ne;
#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    if (!oky[jj] || (jsel >= 0 && jj != jsel)) continue;
    const int64_t cell = ((int64_t)R.idx * cm + cx) * cm + cyv[jj];
    double y1[3], y0[3];
    g_dots<3>(G, tb.CTM + (cell - (int64_t)cm * cm) * (L + 1), tstride, c, L, y1, y0);
    const float2 D = tb.DM[(int64_t)R.k0 * d.nm * cm * cm + cell];  // tick-major
    const double corr = (double)D.x - f * (double)D.y, dy = (double)D.y;
    const double wxy = (double)WXi * (double)WY[jj];
    const double dwxy = (double)DWXi * (double)WY[jj] + (double)WXi * (double)DWY[jj];
    const double Y1 = la * y1[0] + lb * y1[1] + lc * y1[2];
    const double Y0 = la * y0[0] + lb * y0[1] + lc * y0[2];
    const double gn = (1.0 - f) * Y1 + f * Y0 + c.gc * corr;  // d loss / d n_bin per unit weight
    gN += gn * wxy;
    gS += gn * ne * dwxy;
    gF += ne * wxy * (Y0 - Y1 + (c.gc1 - c.gc0) * corr - c.gc * dy);
    gL += ne * wxy * ((1.0 - f) * (da * y1[0] + db * y1[1] + dc * y1[2]) + f * (da * y0[0] + db * y0[1] + dc * y0[2]));
  }
  return true;
}

// The row of this workgroup in the adjoint row kernels (-1: none): slot blockIdx / kXcd of XCD
// blockIdx % kXcd's band of the processing order, or with no bands (band == nullptr) the XCD's
// equal share of the rows in row order.
__device__ __forceinline__ int adj_row(const int* __restrict__ band, const int* __restrict__ order, int P) {
  const int xcd = (int)(blockIdx.x % kXcd);
  if (band) {
    const int slot = band[xcd] + (int)(blockIdx.x / kXcd);
    return (slot >= band[xcd + 1] || slot >= P) ? -1 : order[slot];
  }
  if ((int64_t)blockIdx.x >= P) return -1;
  const int q8 = P / kXcd, rem = P % kXcd;
  return xcd * q8 + (xcd < rem ? xcd : rem) + (int)(blockIdx.x / kXcd);
}

// Chain rule to the 7 parameters from the pair's d loss / d (n_e, f, sigma_T, lambda).
__device__ __forceinline__ void add_chain(const AdjRec* __restrict__ Ap, double gN, double gF, double gS, double gL,
                                          bool near, double* g) {
  g[0] += gN * Ap->fN[0];
  g[1] += gN * Ap->fN[1];
  g[2] += gN * Ap->fN[2] + gF * Ap->fF;
  g[3] += gN * Ap->fN[3];
  if (near) {
    g[2] += gS * Ap->fS[0] + gL * Ap->fL[0];
    g[5] += gL * Ap->fL[1];
    g[6] += gS * Ap->fS[1];
  }
}

#ifndef LSIM_ADJ_BIN_ITEMS
#define LSIM_ADJ_BIN_ITEMS 0
#endif
constexpr bool kAdjBinItems = LSIM_ADJ_BIN_ITEMS != 0;  // one work item per diffusion bin (else per x-bin row)

// Stage 2, row-owned: one workgroup per pixel row holds the row's step function in LDS and walks
// the segments of its neighbourhood cells (the same pairs the forward's waveform kernel visits),
// every thread taking pairs from the flattened list; fixed-order block reduction to the row's 7
// partial gradients.  Two passes (separate launches, separate partials): NEAR = false takes the
// neighbour-only cells (|a| or |b| > 1; little register state, high occupancy), NEAR = true the
// 3x3 cells with the diffusion bins.  Rows without gradient (no valid hit) exit at once.  (A
// waveform cotangent takes k_rows_adj_dense instead.)
template <int NB, int WAVES, bool NEAR>
__global__ __launch_bounds__(WAVES * kWave) __attribute__((amdgpu_waves_per_eu(4))) void k_rows_adj(Dev d, Tables tb, const SegRec* __restrict__ rec,
                                                            const PairRec* __restrict__ prs,
                                                            const AdjRec* __restrict__ adj, const int2* __restrict__ rng,
                                                            const int* __restrict__ row_cell,
                                                            const int* __restrict__ n_rows_dev, int64_t max_rows,
                                                            const int* __restrict__ band,
                                                            const int* __restrict__ order,
                                                            const Pieces* __restrict__ pieces,
                                                            double* __restrict__ partial) {
  constexpr int NT = WAVES * kWave;
  constexpr int kNP = 2 * kMaxH;  // breakpoints held in LDS
  __shared__ int c_beg[kMaxCells], c_cum[kMaxCells + 1], c_ab[kMaxCells];
  __shared__ int ncls[3];
  __shared__ int cell_scratch[3 * kMaxCells];
  __shared__ int spx[kNP];
  __shared__ float spc[kNP], sdc[kNP];
  __shared__ uint16_t qlut[kMaxOut];
  __shared__ double red[kNGrad][NT / kWave];
  const int t = threadIdx.x;
  const int P0 = *n_rows_dev;
  const int P = P0 < max_rows ? P0 : (int)max_rows;
  // workgroup i runs on XCD i % kXcd.  With bands (large launches): the forward's XCD bands of
  // equal estimated work in tile-major order, slot i / kXcd of XCD i % kXcd's band; without (small
  // launches, where rebuilding them costs more than it saves): equal row counts per XCD in row
  // order.  The partials of rows no workgroup takes (past P) were zeroed by the host.
  const int r = adj_row(band, order, P);
  if (r < 0) return;
  const Pieces* pr = pieces + r;
  const int np = pr->n;
  {
    if (np == 0) {
      if (t < kNGrad) partial[(int64_t)r * kNGrad + t] = 0.0;
      return;
    }
    if (t < np) {
      spx[t] = pr->x[t];
      spc[t] = pr->c[t];
      sdc[t] = pr->c[t] - (t + 1 < np ? pr->c[t + 1] : 0.f);
    }
    __syncthreads();
    // qlut[o] = number of breakpoints below o = first q with x[q] >= o (np from nout on: G = 0 there)
    for (int o = t; o < kMaxOut; o += NT) {
      int q = 0;
      for (int k = 0; k < np; ++k) q += spx[k] < o ? 1 : 0;
      qlut[o] = (uint16_t)q;
    }
  }
  __syncthreads();  // qlut complete
  // this pass's cells in (a, b) order, each with its live pairs only: window starts ascend within a
  // cell and a pair is dead from the first start at which G has vanished, so the live pairs lead
  const RowG G{spx, spc, sdc, qlut, np};
  list_cells<NT>(
      d, rng, row_cell[r],
      [](int a, int b) {
        const bool near = a >= -1 && a <= 1 && b >= -1 && b <= 1;
        return near == NEAR ? 0 : -1;
      },
      [&](int b, int e) {
        int lo = b, hi = e;  // the first dead pair of [b, e)
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (pair_dead(prs[mid].S, G)) hi = mid;
          else lo = mid + 1;
        }
        return lo - b;
      },
      cell_scratch, c_beg, c_cum, c_ab, ncls);
  double g[kNGrad];
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) g[k] = 0.0;
  const int nct = ncls[2];
  const int tot = c_cum[nct];
  // work items: a far pair is one item (its stripe); a near pair is NX + 1 items on consecutive
  // threads, its diffusion bins by x-bin row (NX = NB) or one by one (NX = NB * NB), and its stripe
  constexpr int NX = kAdjBinItems ? NB * NB : NB;
  constexpr int NI = NEAR ? NX + 1 : 1;
  int cur = 0;
  const int lane = t & (kWave - 1), wid = t / kWave;
  if (!NEAR) {
    // far pairs: the stripe only.  Each thread fetches its next pair's record (the first 16 bytes
    // of the pair record) and chain factors while the current pair's table loads are in flight.
    struct FarItem {
      int4 P;  // S, ne, f, flags
      double fN0, fN1, fN2, fN3, fF;
      int ab;
    };
    auto fetch = [&](int i, FarItem& x) {
      while (i >= c_cum[cur + 1]) ++cur;
      const int sg = c_beg[cur] + (i - c_cum[cur]);
      x.ab = c_ab[cur];
      x.P = *reinterpret_cast<const int4*>(prs + sg);
      const AdjRec* A = adj + sg;
      x.fN0 = A->fN[0];
      x.fN1 = A->fN[1];
      x.fN2 = A->fN[2];
      x.fN3 = A->fN[3];
      x.fF = A->fF;
    };
    FarItem x, y;
    if (t < tot) fetch(t, x);
    for (int it = t; it < tot; it += NT) {
      if (it + NT < tot) fetch(it + NT, y);
      const float ne = __int_as_float(x.P.y);
      if (ne != 0.f) {  // every listed pair is live (G does not vanish before its window)
        PairCtx pc;
        pair_ctx(d, x.P.x, ne, __int_as_float(x.P.z), G, pc);
        double gN = 0.0, gF = 0.0;
        stripe_adjoint(d, tb, (x.P.w >> 8) & 255, (x.P.w >> 16) & 255, pc, (x.ab >> 8) - 64, (x.ab & 255) - 64, G, gN,
                       gF);
        g[0] += gN * x.fN0;
        g[1] += gN * x.fN1;
        g[2] += gN * x.fN2 + gF * x.fF;
        g[3] += gN * x.fN3;
      }
      x = y;
    }
  }
  for (int it = t; NEAR && it < tot * NI; it += NT) {
    const int i = it / NI, part = it - i * NI;
    while (i >= c_cum[cur + 1]) ++cur;
    const int s = c_beg[cur] + (i - c_cum[cur]);
    const int ab = c_ab[cur];
    const int a = (ab >> 8) - 64, b = (ab & 255) - 64;
    double gN = 0.0, gF = 0.0, gS = 0.0, gL = 0.0;
    PairCtx pc;
    const SegRec R = rec[s];
    if (R.ne == 0.f) continue;
    const int S = d.nt - d.L - R.k0;
    if (pair_dead(S, G)) continue;
    pair_ctx(d, S, R.ne, R.f, G, pc);
    if (part == NX) {
      if ((a == 0 && b == 0) || a < -d.nn || a > d.nn || b < -d.nn || b > d.nn) continue;  // no stripe there
      stripe_adjoint(d, tb, R.uxy & 0xffff, R.uxy >> 16, pc, a, b, G, gN, gF);
    } else if (!xbin_adjoint<NB>(d, tb, R, adj + s, pc, a, b, kAdjBinItems ? part / NB : part,
                                 kAdjBinItems ? part % NB : -1, G, gN, gF, gS, gL)) {
      continue;
    }
    add_chain(adj + s, gN, gF, gS, gL, NEAR, g);
  }
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) {
    const double v = wave_sum(g[k]);
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (t < kNGrad) {
    double v = 0.0;
    for (int w = 0; w < NT / kWave; ++w) v += red[t][w];
    partial[(int64_t)r * kNGrad + t] = v;
  }
}

// ------------------------------------------------------------------------------------------
// Stage 2 with a waveform cotangent (lsim_backward_wfs)
// ------------------------------------------------------------------------------------------
//
// G(o) = gw[r][o] + the self-trigger's step function: one value per output tick, so the window
// sums are plain dot products, and they take the forward's access pattern with reads in place of
// stores.  A whole wave works on one (segment, row) pair: lane l holds the padded window samples
// p = l + 63 c (c < 3) of the pair (output tick o = S - 2 + p), G(o) comes from LDS and the next
// sample (the f window) from lane l + 1 by a DPP shift; lane 63 only supplies that next sample.
// The window sums of a pair enter the gradients linearly, so every lane scales its own share by
// the pair's chain factors and adds it to its 7 partial gradients: no reduction per pair, one
// per row at the end (fixed order: deterministic).
//   Stripes (the raw-window stripe of every pixel offset but the centre, |a|, |b| <= nn): blocks
//     of 64 pairs whose factors the lanes compute in parallel and stage in LDS; the wave then reads
//     them back by broadcast, 4 pairs' window loads in flight at a time.  The boundary-correction
//     terms (no window) are added by the lane that set the pair up.
//   Diffusion bins landing on the row (the 3x3 cells; side rows only for segments near the pixel
//     edge): one (segment, row) item per wave at a time, LUT cell by LUT cell from the interleaved
//     template windows (T_{m-1}, T_m, T_{m+1}) of TM3: V = la T_{m-1} + lb T_m + lc T_{m+1} for
//     n_e, f, sigma_T, and dV/dlambda = da (T_{m-1} - T_m) + dc (T_{m+1} - T_m) for lambda (the
//     three Lagrange derivatives sum to zero, so the differences lose no precision).

constexpr int kGOff = 192;                       // LDS index of output tick 0
constexpr int kGLen = kMaxOut + 2 * kGOff + 64;  // ticks -kGOff .. kMaxOut + kGOff + 63

// One stripe pair as staged for the wave (64 bytes).
struct __align__(16) DenseStripe {
  int S;             // window start, clamped to 2 when the window misses the row (factors 0 then)
  int off;           // byte offset of the pair's padded raw-window row in TNP
  float c1, c0, ne;  // 1 - f, f, charge (0 for padding and for windows that miss the row)
  int pad;
  double fN[4], fF;
};
static_assert(sizeof(DenseStripe) == 64, "DenseStripe layout");

// G at a correction tick (those are kept for 0 <= o <= nout - 2); Gw holds 0 at tick 0 (a window
// tick is kept from 1 on), G0 holds G(0).
__device__ __forceinline__ double g_corr(const float* Gw, float G0, int o, int nout) {
  return (o >= 0 && o <= nout - 2) ? (double)(o == 0 ? G0 : Gw[kGOff + o]) : 0.0;
}

// The window of a pair meets the output ticks 1 .. nout - 1 (window ticks S - 1 .. S + L - 1)
__device__ __forceinline__ bool window_live(int S, int L, int nout) { return S >= 2 - L && S <= nout; }

template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_rows_adj_dense(Dev d, Tables tb, const SegRec* __restrict__ rec,
                                                                  const PairRec* __restrict__ prs,
                                                                  const AdjRec* __restrict__ adj,
                                                                  const int2* __restrict__ rng,
                                                                  const int* __restrict__ row_cell,
                                                                  const int* __restrict__ n_rows_dev, int64_t max_rows,
                                                                  const int* __restrict__ band,
                                                                  const int* __restrict__ order,
                                                                  const Pieces* __restrict__ pieces,
                                                                  const float* __restrict__ gw,
                                                                  double* __restrict__ partial) {
  constexpr int NT = WAVES * kWave;
  constexpr int SYM = (NB - 1) / 2;
  __shared__ float Gw[kGLen];
  __shared__ DenseStripe stage[WAVES][kWave];
  __shared__ int c_beg[kMaxCells], c_cum[kMaxCells + 1], c_ab[kMaxCells];
  __shared__ int ncls[3];
  __shared__ int cell_scratch[3 * kMaxCells];
  __shared__ double red[kNGrad][WAVES];
  __shared__ float sG0;
  __shared__ int s_any;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1), wid = t / kWave;
  const int P0 = *n_rows_dev;
  const int P = P0 < max_rows ? P0 : (int)max_rows;
  double* out_near = partial + max_rows * kNGrad;  // the second half of the partials: zero here
  const int r = adj_row(band, order, P);  // as in k_rows_adj (partials of the rows past P pre-zeroed)
  if (r < 0) return;
  const int nout = d.nout, L = d.L;
  // G over the padded tick range: window ticks 1 .. nout - 1, zero elsewhere
  if (t == 0) s_any = 0;
  __syncthreads();
  {
    const Pieces* pr = pieces + r;
    const float* gr = gw + (int64_t)r * nout;
    const int nf = pr->n;
    bool any = false;
    for (int i = t; i < kGLen; i += NT) {
      const int o = i - kGOff;
      float v = 0.f;
      if (o >= 0 && o < nout) {
        float cf = 0.f;
        for (int k = nf - 1; k >= 0; --k)
          if (pr->x[k] >= o) cf = pr->c[k];
        v = gr[o] + cf;
        any = any || v != 0.f;
        if (o == 0) sG0 = v;
      }
      Gw[i] = (o >= 1 && o < nout) ? v : 0.f;
    }
    if (any) s_any = 1;
  }
  __syncthreads();
  if (!s_any) {  // no cotangent on this row
    if (t < kNGrad) {
      partial[(int64_t)r * kNGrad + t] = 0.0;
      out_near[(int64_t)r * kNGrad + t] = 0.0;
    }
    return;
  }
  const float G0 = sG0;
  // cells: far (stripes), side (stripes and bins), centre (bins)
  list_cells<NT>(
      d, rng, row_cell[r],
      [](int a, int b) {
        const bool near = a >= -1 && a <= 1 && b >= -1 && b <= 1;
        return !near ? 0 : ((a == 0 && b == 0) ? 2 : 1);
      },
      [](int b, int e) { return e - b; }, cell_scratch, c_beg, c_cum, c_ab, ncls);
  double g[kNGrad];
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) g[k] = 0.0;
  const int nsb = d.nsb, lxly = d.lx * d.ly;
  const int n_stripe_cells = d.nn >= 1 ? ncls[1] : ncls[0];
  const int nstripe = c_cum[n_stripe_cells];
  const __amdgpu_buffer_rsrc_t tnp =
      __builtin_amdgcn_make_buffer_rsrc((void*)tb.TNP, 0, d.lx * d.ly * kWinPad * 4, kBufDesc3);
  const float gmask = lane < kChunk ? 1.f : 0.f;  // lane 63 only supplies the next sample
  // ---- stripes ----
  for (int b0 = wid * kWave; b0 < nstripe; b0 += WAVES * kWave) {
    const int p = b0 + lane;
    DenseStripe z;
    z.S = 2;
    z.off = 0;
    z.c1 = z.c0 = z.ne = 0.f;
    z.pad = 0;
    z.fN[0] = z.fN[1] = z.fN[2] = z.fN[3] = z.fF = 0.0;
    if (p < nstripe) {
      int lo = 0, hi = n_stripe_cells - 1;  // the cell holding p: last c with c_cum[c] <= p
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c_cum[mid] <= p) lo = mid;
        else hi = mid - 1;
      }
      const int sg = c_beg[lo] + (p - c_cum[lo]);
      const int ab = c_ab[lo];
      const int a = (ab >> 8) - 64, b = (ab & 255) - 64;
      const PairRec R = prs[sg];
      if (R.ne != 0.f) {
        const int S = R.S;
        const double f = (double)R.f, ne = (double)R.ne;
        const int ux = (R.flags >> 8) & 255, uy = (R.flags >> 16) & 255;
        const int sidx = offset_bin(ux, a, nsb) * d.ly + offset_bin(uy, b, nsb);
        const float2 D = tb.DN[(int64_t)(d.nt - L - S) * lxly + sidx];  // tick-major, k0 = nt - L - S
        const double corr = (double)D.x - f * (double)D.y, dy = (double)D.y;
        const double gc0 = g_corr(Gw, G0, S - 1, nout), gc1 = g_corr(Gw, G0, S - 2, nout);
        const double gc = (1.0 - f) * gc0 + f * gc1;
        const AdjRec* A = adj + sg;
        const double fN0 = A->fN[0], fN1 = A->fN[1], fN2 = A->fN[2], fN3 = A->fN[3], fF = A->fF;
        // the correction terms of the pair (no window)
        const double gNc = gc * corr, gFc = ne * ((gc1 - gc0) * corr - gc * dy);
        g[0] += gNc * fN0;
        g[1] += gNc * fN1;
        g[2] += gNc * fN2 + gFc * fF;
        g[3] += gNc * fN3;
        if (window_live(S, L, nout)) {
          z.S = S;
          z.off = sidx * kWinPad * 4;
          z.c1 = (float)(1.0 - f);
          z.c0 = R.f;
          z.ne = R.ne;
          z.fN[0] = fN0;
          z.fN[1] = fN1;
          z.fN[2] = fN2;
          z.fN[3] = fN3;
          z.fF = fF;
        }
      }
    }
    stage[wid][lane] = z;
    __builtin_amdgcn_wave_barrier();
    const int n = __builtin_amdgcn_readfirstlane(nstripe - b0 < kWave ? nstripe - b0 : kWave);
    for (int i = 0; i < n; i += 4) {
      // pairs i .. i + 3 (entries past n are padding: factors 0, a window at S = 2)
      DenseStripe q[4];
      float x[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        q[k] = stage[wid][i + k];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          x[k][c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(tnp, lane * 4 + c * 4 * kChunk + q[k].off, 0, 0));
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float u1 = 0.f, u0 = 0.f;
        const float* gp = Gw + kGOff + q[k].S - 2 + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float gv = gmask * gp[c * kChunk];
          const float xn = lane_next0(x[k][c]);
          u1 += gv * x[k][c];
          u0 += gv * xn;
        }
        const double aa = (double)(q[k].c1 * u1 + q[k].c0 * u0), bb = (double)(q[k].ne * (u0 - u1));
        g[0] += q[k].fN[0] * aa;
        g[1] += q[k].fN[1] * aa;
        g[2] += q[k].fN[2] * aa + q[k].fF * bb;
        g[3] += q[k].fN[3] * aa;
      }
    }
    __builtin_amdgcn_wave_barrier();  // the stage is rewritten by the next block
  }
  // ---- diffusion bins of the 3x3 cells ----
  const int kb0 = c_cum[ncls[0]], kb1 = c_cum[ncls[2]];
  int cur = ncls[0];
  const double lead = lane == 0 ? 1.0 : 0.0;  // terms without a window: once per item
  for (int k = kb0 + wid; k < kb1; k += WAVES) {
    while (k >= c_cum[cur + 1]) ++cur;
    const int s = __builtin_amdgcn_readfirstlane(c_beg[cur] + (k - c_cum[cur]));
    const int ab = __builtin_amdgcn_readfirstlane(c_ab[cur]);
    const int a = (ab >> 8) - 64, b = (ab & 255) - 64;
    const SegRec R = rec[s];
    if (R.ne == 0.f) continue;
    const int ux = R.uxy & 0xffff, uy = R.uxy >> 16;
    if (!bins_reach<NB>(ux, a, nsb) || !bins_reach<NB>(uy, b, nsb)) continue;
    const AdjRec* A = adj + s;
    const int S = d.nt - L - R.k0;
    const double f = (double)R.f, ne = (double)R.ne;
    const double gc0 = g_corr(Gw, G0, S - 1, nout), gc1 = g_corr(Gw, G0, S - 2, nout);
    const double gc = (1.0 - f) * gc0 + f * gc1;
    const bool live = window_live(S, L, nout);
    const float* gp = Gw + kGOff + (live ? S : 2) - 2 + lane;
    float gv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gv[c] = live ? gmask * gp[c * kChunk] : 0.f;
    bool oky[NB];
    int cyv[NB];
    float WY[NB], DWY[NB];
    axis_cells<NB>(uy, b, nsb, R.wy, A->dwy, oky, cyv, WY, DWY);
    const float la = R.la, lb = R.lb, lc = R.lc;
    const float da = (float)A->dl[0], dc = (float)A->dl[2];
    const int cm = d.cm;
    double gN = 0.0, gF = 0.0, gS = 0.0, gL = 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int vi = ux + i - SYM;
      const int di = vi < 0 ? -1 : (vi >= nsb ? 1 : 0);
      if (di != a) continue;
      const int cx = offset_bin(vi - di * nsb, 0, nsb);
      // x-bins on the same LUT cell are merged into the first of them
      float WXi = 0.f, DWXi = 0.f;
      bool first = true;
#pragma unroll
      for (int kk = 0; kk < NB; ++kk) {
        const int v = ux + kk - SYM;
        const int dv = v < 0 ? -1 : (v >= nsb ? 1 : 0);
        if (dv != a || offset_bin(v - dv * nsb, 0, nsb) != cx) continue;
        if (kk < i) first = false;
        WXi += R.wx[kk];
        DWXi += A->dwx[kk];
      }
      if (!first) continue;
      // the y cells of this x cell: all window loads first, then the sums
      float3 q[NB][3];
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        const int cy = oky[jj] ? cyv[jj] : cyv[0];
        const float3* T3 = tb.TM3 + (((int64_t)R.idx * cm + cx) * cm + cy) * kWinPad + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) q[jj][c] = T3[c * kChunk];
      }
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        if (!oky[jj]) continue;
        const int64_t cell = ((int64_t)R.idx * cm + cx) * cm + cyv[jj];
        float u1 = 0.f, u0 = 0.f, ul1 = 0.f, ul0 = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float V = q[jj][c].y * lb + q[jj][c].x * la + q[jj][c].z * lc;
          const float VL = da * (q[jj][c].x - q[jj][c].y) + dc * (q[jj][c].z - q[jj][c].y);
          const float Vn = lane_next0(V), VLn = lane_next0(VL);
          u1 += gv[c] * V;
          u0 += gv[c] * Vn;
          ul1 += gv[c] * VL;
          ul0 += gv[c] * VLn;
        }
        const float2 D = tb.DM[(int64_t)R.k0 * d.nm * cm * cm + cell];  // tick-major
        const double corr = lead * ((double)D.x - f * (double)D.y), dy = lead * (double)D.y;
        const double wxy = (double)WXi * (double)WY[jj];
        const double dwxy = (double)DWXi * (double)WY[jj] + (double)WXi * (double)DWY[jj];
        const double gn = (1.0 - f) * (double)u1 + f * (double)u0 + gc * corr;
        gN += gn * wxy;
        gS += gn * ne * dwxy;
        gF += ne * wxy * ((double)u0 - (double)u1 + (gc1 - gc0) * corr - gc * dy);
        gL += ne * wxy * ((1.0 - f) * (double)ul1 + f * (double)ul0);
      }
    }
    add_chain(A, gN, gF, gS, gL, true, g);
  }
  // fixed-order block reduction of the lanes' partial gradients
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) {
    const double v = wave_sum(g[k]);
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (t < kNGrad) {
    double v = 0.0;
    for (int w = 0; w < WAVES; ++w) v += red[t][w];
    partial[(int64_t)r * kNGrad + t] = v;
    out_near[(int64_t)r * kNGrad + t] = 0.0;
  }
}

// One pass over the per-row partials (all 7 gradients and the hit-z term per thread), then a
// fixed-order combine: butterfly within each wave, waves in index order.  Deterministic.
constexpr int kReduceThreads = 1024;

__global__ __launch_bounds__(kReduceThreads) void k_grad_reduce(Dev d, const double* __restrict__ partial, int64_t nblocks,
                                                                const double* __restrict__ dz_part,
                                                                const int64_t* __restrict__ n_rows_dev, int64_t max_rows,
                                                                double* __restrict__ out) {
  __shared__ double sh[kNGrad + 1][kReduceThreads / kWave];
  int64_t P = *n_rows_dev;
  if (P > max_rows) P = max_rows;
  double acc[kNGrad + 1];
#pragma unroll
  for (int k = 0; k <= kNGrad; ++k) acc[k] = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < kNGrad; ++k) acc[k] += partial[b * kNGrad + k];
  }
  for (int64_t r = threadIdx.x; r < P; r += blockDim.x) acc[kNGrad] += dz_part[r];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k <= kNGrad; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) sh[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[kNGrad + 1];
    for (int k = 0; k <= kNGrad; ++k) {
      tot[k] = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) tot[k] += sh[k][w];
    }
    for (int k = 0; k < kNGrad; ++k) out[k] = tot[k];
    out[2] += tot[kNGrad] * d.dv;  // hit z depends on eField through the drift velocity
  }
}

}  // namespace lsim
