// Adjoint of the LUT chain for the 7 physics parameters (Ab, kb, eField, lifetime, vdrift,
// long_diff, tran_diff).  Integer decisions are held at their forward values, as reverse-mode
// AD through the reference does (floor/astype/searchsorted/argmax carry no gradient).
//
// Stage 1 (one thread per pixel row): the self-trigger adjoint.  The ADC value of pass k is
// Q^(k)[e_k]; pass k then sets Q^(k+1) = max(Q^(k) - Q^(k)[g_k], 0).  Reverse mode over the 2H
// tracked points gives the adjoint of Q^(0), and because Q^(0) is the cumulative sum of
// t_sampling * wfs, d loss / d wfs is piecewise constant with at most 2H steps.
// Stage 2 (one thread per segment): for every pixel row the segment reaches, inner products of
// that step function with the segment's template windows come from window prefix sums, which
// yields d loss / d (charge, transverse width, fractional tick, longitudinal width) of the
// segment; the chain rule to the parameters follows and a fixed-order block reduction writes
// per-block partial gradients.
// Stage 3: fixed-order reduction of the partials (and of the hit-z term) to 7 values.
#pragma once
#include "forward.h"

namespace lsim {

constexpr int kNGrad = 7;

__global__ void k_fee_adjoint(Dev d, const FeeTape* __restrict__ tape, const float* __restrict__ d_adc,
                              const float* __restrict__ d_z, const int64_t* __restrict__ n_rows_dev, int64_t max_rows,
                              Pieces* __restrict__ pieces, double* __restrict__ dz_part) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t P = *n_rows_dev;
  if (P > max_rows) P = max_rows;
  if (r >= max_rows) return;
  if (r >= P) {
    dz_part[r] = 0.0;
    return;
  }
  const FeeTape tp = tape[r];
  const int H = d.H;
  double A[2 * kMaxH];
  for (int j = 0; j < 2 * H; ++j) A[j] = 0.0;
  for (int k = H - 1; k >= 0; --k) {
    const uint32_t m = tp.mask[k];
    double s = 0.0;
    for (int j = 0; j < 2 * H; ++j) {
      if ((m >> j) & 1u) s += A[j];
      else A[j] = 0.0;
    }
    A[H + k] -= s;
    A[k] += (double)d_adc[r * H + k] * (double)tp.slope[k];
  }
  // merge points by position (ascending), suffix sums -> step values
  int pos[2 * kMaxH];
  double val[2 * kMaxH];
  int n = 0;
  for (int j = 0; j < 2 * H; ++j) {
    const int p = tp.pt[j];
    int q = 0;
    while (q < n && pos[q] < p) ++q;
    if (q < n && pos[q] == p) {
      val[q] += A[j];
    } else {
      for (int u = n; u > q; --u) {
        pos[u] = pos[u - 1];
        val[u] = val[u - 1];
      }
      pos[q] = p;
      val[q] = A[j];
      ++n;
    }
  }
  // step values G on (pos[q-1], pos[q]]; keep only the steps where G changes, drop a zero tail
  float cv[2 * kMaxH];
  double suf = 0.0;
  for (int q = n - 1; q >= 0; --q) {
    suf += val[q];
    cv[q] = (float)(suf * d.ts);
  }
  Pieces pc;
  int m = 0;
  for (int q = 0; q < n; ++q) {
    if (m > 0 && cv[q] == pc.c[m - 1]) {
      pc.x[m - 1] = pos[q];  // same value: extend the previous piece
    } else {
      pc.x[m] = pos[q];
      pc.c[m] = cv[q];
      ++m;
    }
  }
  while (m > 0 && pc.c[m - 1] == 0.f) --m;
  pc.n = m;
  pc.pad[0] = pc.pad[1] = pc.pad[2] = 0;
  for (int q = m; q < 2 * kMaxH; ++q) {
    pc.x[q] = 0x7fffffff;
    pc.c[q] = 0.f;
  }
  pieces[r] = pc;
  double dz = 0.0;
  if (d_z) {
    const int pl = tp.plane;
    const double za = d.tpc[pl][2][0], zc = d.tpc[pl][2][1];
    const double sgn = (zc - za) > 0 ? 1.0 : ((zc - za) < 0 ? -1.0 : 0.0);
    for (int k = 0; k < H; ++k) dz += (double)d_z[r * H + k] * (double)tp.tick[k] * d.ts * sgn;
  }
  dz_part[r] = dz;  // d loss / d vdrift_eff through the hit z positions
}

// G(o) of a row: c[q] for the first piece with x[q] >= o (0 past the last piece).
__device__ __forceinline__ double g_at(const Pieces* __restrict__ pc, int n, int q0, int o) {
  for (int q = q0; q < n; ++q)
    if (pc->x[q] >= o) return (double)pc->c[q];
  return 0.0;
}

// sum_{o=lo..hi} G(o) * X[o - S + sh] using the window prefix sums CX (L+1 entries); pieces
// before q0 end below lo.
__device__ __forceinline__ double g_dot(const Pieces* __restrict__ pc, int n, int q0, const double* __restrict__ CX, int lo,
                                        int hi, int S, int sh) {
  double acc = 0.0;
  int prev = q0 > 0 ? pc->x[q0 - 1] : -1;  // pieces cover (prev, x_q]
  for (int q = q0; q < n && prev < hi; ++q) {
    const int xq = pc->x[q];
    const int a = (prev + 1) > lo ? (prev + 1) : lo;
    const int b = xq < hi ? xq : hi;
    if (a <= b) acc += (double)pc->c[q] * (CX[b - S + sh + 1] - CX[a - S + sh]);
    prev = xq;
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_seg_adjoint(Dev d, SegIn in, Tables tb, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, int64_t n,
                                                     const int* __restrict__ cell_row, const Pieces* __restrict__ pieces,
                                                     int64_t max_rows, double* __restrict__ partial) {
  __shared__ double red[kNGrad][256 / kWave];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double gr[kNGrad];
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) gr[k] = 0.0;
  if (i < n && keys[i] < (uint32_t)d.n_cells) {
    Phys ph;
    seg_physics(d, in, vals[i], ph);
    if (ph.ne != 0.0) {
      double wx[kMaxNb], wy[kMaxNb], dwx[kMaxNb], dwy[kMaxNb];
      tran_weights(d, ph.x0, ph.st, wx, dwx);
      tran_weights(d, ph.y0, ph.st, wy, dwy);
      const double s0 = d.sig[ph.idx - 1], s1 = d.sig[ph.idx], s2 = d.sig[ph.idx + 1];
      const double l = ph.lam;
      const double da = (2 * l - s1 - s2) / ((s0 - s1) * (s0 - s2));
      const double db = (2 * l - s0 - s2) / ((s1 - s0) * (s1 - s2));
      const double dc = (2 * l - s0 - s1) / ((s2 - s0) * (s2 - s1));
      // forward used float32 copies of these; the adjoint linearises the same quantities
      const double ne = (double)(float)ph.ne;
      const double f = (double)(float)ph.f;
      const double la = (double)(float)ph.la, lb = (double)(float)ph.lb, lc = (double)(float)ph.lc;
      const int L = d.L, nt = d.nt, nsb = d.nsb, cm = d.cm;
      const int S = nt - L - ph.k0;
      const int ev = ph.ev, plane = ph.plane;
      const int ux = pymod_i(ph.bx, nsb), uy = pymod_i(ph.by, nsb);
      const int64_t tstride = (int64_t)cm * cm * (L + 1);
      double gN = 0.0, gS = 0.0, gF = 0.0, gL = 0.0;
      const int nout = d.nout;
      const int lo1 = S > 1 ? S : 1;                                   // (1-f) window: o = S + j
      const int hi1 = (S + L - 1) < (nout - 1) ? (S + L - 1) : (nout - 1);
      const int lo0 = (S - 1) > 1 ? (S - 1) : 1;                       // f window: o = S - 1 + j
      const int hi0 = (S + L - 2) < (nout - 1) ? (S + L - 2) : (nout - 1);
      for (int a = -d.nn; a <= d.nn; ++a) {
        const int tx = ph.mpx + a;
        if (tx < 0 || tx >= d.nx) continue;
        for (int b = -d.nn; b <= d.nn; ++b) {
          const int ty = ph.mpy + b;
          if (ty < 0 || ty >= d.ny) continue;
          const int64_t c = cell_of(d, ev, plane, tx, ty);
          const int r = cell_row[c];
          if (r < 0 || r >= max_rows) continue;
          const bool has_nb = (a != 0) || (b != 0);
          unsigned imask = 0, jmask = 0;
          if (a >= -1 && a <= 1 && b >= -1 && b <= 1) {
            for (int q = 0; q < d.nb; ++q) {
              if (floordiv_i(ph.bx + q - d.nbsym, nsb) == tx) imask |= 1u << q;
              if (floordiv_i(ph.by + q - d.nbsym, nsb) == ty) jmask |= 1u << q;
            }
          }
          const bool has_bins = imask && jmask;
          if (!has_nb && !has_bins) continue;
          const Pieces* pc = pieces + r;
          const int np = pc->n;
          if (np == 0) continue;
          // first piece reaching the earliest tick this segment touches (S - 2)
          int q0 = 0;
          while (q0 < np && pc->x[q0] < S - 2) ++q0;
          if (q0 == np) continue;
          // correction cotangents: (1-f) Delta at o = S-1, f Delta at o = S-2 (valid for o >= 0)
          const double gc0 = (S - 1 >= 0 && S - 1 <= nout - 2) ? g_at(pc, np, q0, S - 1) : 0.0;
          const double gc1 = (S - 2 >= 0 && S - 2 <= nout - 2) ? g_at(pc, np, q0, S - 2) : 0.0;
          const double gc = (1.0 - f) * gc0 + f * gc1;
          if (has_nb) {
            const int cxN = (int)fabsf((float)(ux - nsb / 2) + 0.5f - (float)(a * nsb));
            const int cyN = (int)fabsf((float)(uy - nsb / 2) + 0.5f - (float)(b * nsb));
            const int64_t sidx = (int64_t)cxN * d.ly + cyN;
            const double* CX = tb.CTN + sidx * (L + 1);
            const double y1 = g_dot(pc, np, q0, CX, lo1, hi1, S, 0);
            const double y0 = g_dot(pc, np, q0, CX, lo0, hi0, S, 1);
            const float2 D = tb.DN[sidx * nt + ph.k0];
            const double corr = (double)D.x - f * (double)D.y;
            gN += (1.0 - f) * y1 + f * y0 + gc * corr;
            gF += ne * (y0 - y1 + (gc1 - gc0) * corr - gc * (double)D.y);
          }
          if (has_bins) {
            for (int q = 0; q < d.nb; ++q) {
              if (!((imask >> q) & 1)) continue;
              const int cx = (int)fabsf((float)(pymod_i(ph.bx + q - d.nbsym, nsb) - nsb / 2) + 0.5f);
              for (int qq = 0; qq < d.nb; ++qq) {
                if (!((jmask >> qq) & 1)) continue;
                const int cy = (int)fabsf((float)(pymod_i(ph.by + qq - d.nbsym, nsb) - nsb / 2) + 0.5f);
                const int64_t base = (((int64_t)(ph.idx - 1) * cm + cx) * cm + cy) * (L + 1);
                const double* C1 = tb.CTM + base;
                const double* C2 = C1 + tstride;
                const double* C3 = C2 + tstride;
                const double y1a = g_dot(pc, np, q0, C1, lo1, hi1, S, 0), y0a = g_dot(pc, np, q0, C1, lo0, hi0, S, 1);
                const double y1b = g_dot(pc, np, q0, C2, lo1, hi1, S, 0), y0b = g_dot(pc, np, q0, C2, lo0, hi0, S, 1);
                const double y1c = g_dot(pc, np, q0, C3, lo1, hi1, S, 0), y0c = g_dot(pc, np, q0, C3, lo0, hi0, S, 1);
                const double y1 = lb * y1b + la * y1a + lc * y1c;
                const double y0 = lb * y0b + la * y0a + lc * y0c;
                const float2 D = tb.DM[(((int64_t)ph.idx * cm + cx) * cm + cy) * nt + ph.k0];
                const double corr = (double)D.x - f * (double)D.y;
                const double wxy = (double)(float)wx[q] * (double)(float)wy[qq];
                const double nbin = ne * wxy;
                const double gn = (1.0 - f) * y1 + f * y0 + gc * corr;  // d loss / d n_bin
                gN += gn * wxy;
                gS += gn * ne * (dwx[q] * wy[qq] + wx[q] * dwy[qq]);
                gF += nbin * (y0 - y1 + (gc1 - gc0) * corr - gc * (double)D.y);
                gL += nbin * ((1.0 - f) * (da * y1a + db * y1b + dc * y1c) + f * (da * y0a + db * y0b + dc * y0c));
              }
            }
          }
        }
      }
      // chain rule to the parameters
      const double v = d.v, dv = d.dv;
      const double att = exp(-ph.td / d.lifetime) * (double)ph.mask;
      const double dne_dAb = ph.dq_dAb * att;
      const double dne_dkb = ph.dq_dkb * att;
      const double dne_dE = ph.dq_dE_rec * att + ph.ne * ph.td * dv / (v * d.lifetime);
      const double dne_dtau = ph.ne * ph.td / (d.lifetime * d.lifetime);
      const double dst_dDT = d.tran_diff > 0 ? ph.st / (2 * d.tran_diff) : 0.0;
      const double dst_dE = -ph.st * dv / (2 * v);
      const double dlam_dDL = d.long_diff > 0 ? ph.lam / (2 * d.long_diff) : 0.0;
      const double dlam_dE = -1.5 * ph.lam * dv / v;
      const double df_dE = -ph.tau * dv / v;
      gr[0] = gN * dne_dAb;
      gr[1] = gN * dne_dkb;
      gr[2] = gN * dne_dE + gS * dst_dE + gF * df_dE + gL * dlam_dE;
      gr[3] = gN * dne_dtau;
      gr[4] = 0.0;  // vdrift is never read by the simulation (drift speed comes from eField)
      gr[5] = gL * dlam_dDL;
      gr[6] = gS * dst_dDT;
    }
  }
  // fixed-order block reduction
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) {
    const double s = wave_sum(gr[k]);
    if (lane == 0) red[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < kNGrad) {
    double s = 0.0;
    for (int w = 0; w < 256 / kWave; ++w) s += red[threadIdx.x][w];
    partial[(int64_t)blockIdx.x * kNGrad + threadIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void k_grad_reduce(Dev d, const double* __restrict__ partial, int64_t nblocks,
                                                     const double* __restrict__ dz_part, const int64_t* __restrict__ n_rows_dev,
                                                     int64_t max_rows, double* __restrict__ out) {
  __shared__ double sh[4];
  int64_t P = *n_rows_dev;
  if (P > max_rows) P = max_rows;
  for (int k = 0; k < kNGrad + 1; ++k) {
    double s = 0.0;
    if (k < kNGrad) {
      for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) s += partial[b * kNGrad + k];
    } else {
      for (int64_t r = threadIdx.x; r < P; r += blockDim.x) s += dz_part[r];
    }
    double tot;
    (void)block_exclusive_scan<double>(s, sh, tot);
    if (threadIdx.x == 0) {
      if (k < kNGrad) out[k] = tot;
      else out[2] += tot * d.dv;  // hit z depends on eField through the drift velocity
    }
    __syncthreads();
  }
}

}  // namespace lsim
