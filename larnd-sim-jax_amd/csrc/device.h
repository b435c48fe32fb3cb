// Device-side configuration, the per-segment record and the per-segment physics of the LUT
// simulation chain (quench, drift, sub-bin/pixel assignment, transverse weights, template
// interpolation).  All index-deciding arithmetic is float64 so that bins, pixels, ticks and
// template indices agree bit-for-bit with the float64 oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lsim.h"

namespace lsim {

constexpr int kWave = 64;
constexpr int kMaxNb = 5;          // diffusion bins per axis supported by the segment record
constexpr int kMaxOut = 2048;      // output ticks held per row in LDS
constexpr int kMaxH = LSIM_MAX_ADC;

// Everything a kernel needs, passed by value (kernel arguments are captured by hipGraphs).
struct Dev {
  // geometry
  int ntpc, nx, ny;
  int nsb, nb, nbsym, nn;  // sub-bins per pixel, diffusion bins, neighbour half width
  int ne;                 // half width of the cell grid walked per row: max(nn, 1) (diffusion bins reach one pixel)
  int halo, gx, gy;        // cell grid = pixel grid + halo on each side
  int n_events;
  int64_t n_cells;
  double pitch, w;
  double tpc[LSIM_MAX_TPC][3][2];
  // time / LUT
  int L, nt, nticks, nout, lx, ly, cm, nm, ext, H, interval;
  double sig[LSIM_MAX_TEMPLATES];
  // parameters
  double eField, Ab, kb, lifetime, long_diff, tran_diff, rho, alpha, beta, R_param, mev;
  double v, dv;  // drift velocity and d v / d eField
  double shift[3], margin, ts;
  double thr, gain, vped, vcm, vref, adcc;
  int recomb;
};

// Per-segment record consumed by the waveform kernel, stored in main-pixel (cell) order.
struct __align__(16) SegRec {
  int k0;                // cathode tick (integer part of t0 / t_sampling)
  int uxy;               // sub-bin position inside the main pixel: ux | uy << 16
  float ne, f;           // charge, fractional tick (k0..f: all a neighbour stripe reads)
  int bx, by, idx;       // sub-bin indices, template index
  float la, lb, lc;      // Lagrange weights
  float wx[5], wy[5];    // transverse weights along x and y
};

static_assert(sizeof(SegRec) == 80, "SegRec layout");

// Self-trigger decisions of one pixel row, kept for the adjoint.  Point j < H is the
// integration end e_j of pass j, point H + j its reset index g_j.  mask[k] bit j records
// whether Q^(k)[point j] - Q^(k)[g_k] >= 0 (the clamp of pass k let the value through).
struct __align__(16) FeeTape {
  int16_t pt[2 * kMaxH];
  uint32_t mask[kMaxH];
  float slope[kMaxH];   // d adc_out / d Q[e] (0 for a pass without a valid hit)
  int32_t tick[kMaxH];
  int32_t plane;
  int32_t pad[3];
};

// Piecewise-constant d loss / d wfs of one row: G(o) = c[i] for o in (x[i-1], x[i]], 0 past x[n-1].
struct __align__(16) Pieces {
  int32_t n;
  int32_t pad[3];
  int32_t x[2 * kMaxH];
  float c[2 * kMaxH];
};

// ------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ int floordiv_i(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
__device__ __forceinline__ int pymod_i(int a, int b) {
  int r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}

// numpy's floor_divide / remainder for float64 (npy_divmod).
__device__ __forceinline__ void py_divmod(double a, double b, double& fl, double& mod) {
  mod = fmod(a, b);
  double div = (a - mod) / b;
  if (mod != 0.0) {
    if ((b < 0) != (mod < 0)) {
      mod += b;
      div -= 1.0;
    }
  } else {
    mod = copysign(0.0, b);
  }
  if (div != 0.0) {
    fl = floor(div);
    if (div - fl > 0.5) fl += 1.0;
  } else {
    fl = copysign(0.0, a / b);
  }
}

// float64 -> int like numpy astype(int) for values that fit; far out-of-range values map to a
// sentinel that lands outside every pixel grid.
__device__ __forceinline__ int to_int_clamped(double v) {
  if (!(v > -1.0e9 && v < 1.0e9)) return -1000000000;
  return (int)v;
}

// LUT bin of a sub-bin position u (0 <= u < nsb) seen from a pixel `a` pitches away:
// floor(|u - nsb/2 + 0.5 - nsb*a|) (sim_jax.py:435, 443), in integer arithmetic (nsb even).
__device__ __forceinline__ int offset_bin(int u, int a, int nsb) {
  int t = 2 * u - nsb + 1 - 2 * nsb * a;
  t = t < 0 ? -t : t;
  return (t - 1) >> 1;
}

// ------------------------------------------------------------------------------------------
// per-segment physics
// ------------------------------------------------------------------------------------------

struct SegIn {
  const float *x, *y, *z, *dE, *dEdx, *ev, *zs, *ze, *dx;
  int64_t stride;
  int64_t n;
};

struct Phys {
  int ev, plane, mask;
  int bx, by;            // absolute sub-bin indices in the plane
  int mpx, mpy;          // main pixel
  double x0, y0;         // position inside the sub-bin
  double n0, ne;         // charge after recombination, after lifetime and mask
  double td, sl, st;     // drift time, longitudinal / transverse widths (cm)
  double t0, tau, lam;   // time to cathode (us), in ticks, longitudinal width in ticks
  int k0, idx;
  double f, la, lb, lc;
  double dq_dE_rec;      // d n0 / d eField through recombination (0 for BIRKS-only terms)
  double dq_dAb, dq_dkb; // d n0 / d Ab, d kb
};

// Recombination (quenching_jax.py:18-75) and its eField / Ab / kb derivatives.
__device__ inline void recombination(const Dev& d, double dEdx, double cosphi, double& R, double& dR_dE, double& dR_dAb,
                                     double& dR_dkb) {
  const double u = dEdx / (d.eField * d.rho);
  dR_dAb = dR_dkb = dR_dE = 0.0;
  if (d.recomb == 2) {  // BIRKS
    const double den = 1 + d.kb * dEdx / (d.eField * d.rho);
    R = d.Ab / den;
    dR_dAb = 1.0 / den;
    dR_dkb = -d.Ab * u / (den * den);
    dR_dE = d.Ab * d.kb * u / (d.eField * den * den);
  } else if (d.recomb == 1) {  // BOX
    const double csi = d.beta * dEdx / (d.eField * d.rho);
    const double lg = log(d.alpha + csi);
    const double r = lg / csi;
    R = r > 0 ? r : 0.0;
    if (r > 0) {
      const double dR_dcsi = (csi / (d.alpha + csi) - lg) / (csi * csi);
      dR_dE = dR_dcsi * (-csi / d.eField);
    }
  } else {  // ELLIPSOID
    const double bphi = d.beta / sqrt(1 - cosphi * cosphi + (1.0 / (d.R_param * d.R_param)) * cosphi * cosphi);
    const double csi = bphi * dEdx / (d.eField * d.rho);
    const double lg = log(d.alpha + csi);
    const double r = lg / (csi + 1e-10);
    R = r > 0 ? r : 0.0;
    if (r > 0) {
      const double den = csi + 1e-10;
      const double dR_dcsi = (den / (d.alpha + csi) - lg) / (den * den);
      dR_dE = dR_dcsi * (-csi / d.eField);
    }
  }
}

__device__ __forceinline__ float ldcol(const float* p, int64_t i, int64_t stride) { return p[i * stride]; }

// Quench + drift + bins + template interpolation of one segment (sim_jax.py:374-453,
// drifting_jax.py:93-132, detsim_jax.py:493-512, sim_jax.py:157-168).
__device__ inline void seg_physics(const Dev& d, const SegIn& in, int64_t i, Phys& o) {
  const int64_t s = in.stride;
  const double x = (double)ldcol(in.x, i, s) - d.shift[0];
  const double y = (double)ldcol(in.y, i, s) - d.shift[1];
  const double z = (double)ldcol(in.z, i, s) - d.shift[2];
  const double dE = (double)ldcol(in.dE, i, s);
  const double dEdx = (double)ldcol(in.dEdx, i, s);
  o.ev = to_int_clamped((double)ldcol(in.ev, i, s));
  double cosphi = 0.0;
  if (d.recomb == 3) {
    const double zs = (double)ldcol(in.zs, i, s) - d.shift[2];
    const double ze = (double)ldcol(in.ze, i, s) - d.shift[2];
    cosphi = fabs(ze - zs) / ((double)ldcol(in.dx, i, s) + 1e-10);
  }
  double R, dR_dE, dR_dAb, dR_dkb;
  recombination(d, dEdx, cosphi, R, dR_dE, dR_dAb, dR_dkb);
  o.n0 = R * dE * d.mev;
  o.dq_dE_rec = dR_dE * dE * d.mev;
  o.dq_dAb = dR_dAb * dE * d.mev;
  o.dq_dkb = dR_dkb * dE * d.mev;

  // TPC choice: first TPC whose (margin-extended) box contains the point, plane 0 if none.
  int plane = 0, mask = 0;
  for (int t = d.ntpc - 1; t >= 0; --t) {
    const double m = d.margin;
    const double zmin = fmin(d.tpc[t][2][1] - m, d.tpc[t][2][0] - m);
    const double zmax = fmax(d.tpc[t][2][1] + m, d.tpc[t][2][0] + m);
    const bool c = x >= d.tpc[t][0][0] - m && x <= d.tpc[t][0][1] + m && y >= d.tpc[t][1][0] - m &&
                   y <= d.tpc[t][1][1] + m && z >= zmin && z <= zmax;
    if (c) {
      plane = t;
      mask = 1;
    }
  }
  o.plane = plane;
  o.mask = mask;
  const double z_an = d.tpc[plane][2][0];
  o.td = (fabs(z - z_an) + 1e-6) / d.v;
  o.ne = o.n0 * exp(-o.td / d.lifetime) * (double)mask;
  o.sl = sqrt(o.td * 2 * d.long_diff);
  o.st = sqrt(o.td * 2 * d.tran_diff);

  double fx, fy;
  py_divmod(x - d.tpc[plane][0][0], d.w, fx, o.x0);
  py_divmod(y - d.tpc[plane][1][0], d.w, fy, o.y0);
  o.bx = to_int_clamped(fx);
  o.by = to_int_clamped(fy);
  o.mpx = floordiv_i(o.bx, d.nsb);
  o.mpy = floordiv_i(o.by, d.nsb);

  o.t0 = fabs(z - d.tpc[plane][2][1]) / d.v;
  o.tau = o.t0 / d.ts;
  o.lam = o.sl / d.v / d.ts;
  double fl = floor(o.tau);
  int k0 = fl < 0 ? 0 : (fl > (double)(d.nt - 1) ? d.nt - 1 : (int)fl);
  o.k0 = k0;
  o.f = o.tau - (double)k0;
  int cnt = 0;
  for (int m = 0; m < d.nm; ++m) cnt += (d.sig[m] < o.lam) ? 1 : 0;
  int idx = cnt < 1 ? 1 : (cnt > d.nm - 2 ? d.nm - 2 : cnt);
  o.idx = idx;
  const double s0 = d.sig[idx - 1], s1 = d.sig[idx], s2 = d.sig[idx + 1];
  const double l = o.lam;
  o.la = (l - s1) * (l - s2) / ((s0 - s1) * (s0 - s2));
  o.lb = (l - s0) * (l - s2) / ((s1 - s0) * (s1 - s2));
  o.lc = (l - s0) * (l - s1) / ((s2 - s0) * (s2 - s1));
}

// Transverse weights along one axis (detsim_jax.py:332-341) and their derivative in sigma.
__device__ inline void tran_weights(const Dev& d, double x0, double st, double* W, double* dW) {
  // every index below is a compile-time constant (the bin count nb <= kMaxNb only selects), so the
  // edge arrays stay in registers
  double E[kMaxNb + 1], dEs[kMaxNb + 1];
#pragma unroll
  for (int k = 0; k <= kMaxNb; ++k) {
    if (k == 0 || k >= d.nb) {
      E[k] = k == 0 ? -1.0 : 1.0;
      dEs[k] = 0.0;
    } else {
      const double e = (double)(k - d.nbsym) * d.w - x0;
      const double xi = e / (M_SQRT2 * st);
      E[k] = erf(xi);
      // d erf(e / (sqrt2 s)) / d s = 2/sqrt(pi) exp(-xi^2) * (-xi / s)
      dEs[k] = M_2_SQRTPI * exp(-xi * xi) * (-xi / st);
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxNb; ++k) {
    W[k] = k < d.nb ? 0.5 * (E[k + 1] - E[k]) : 0.0;
    if (dW) dW[k] = k < d.nb ? 0.5 * (dEs[k + 1] - dEs[k]) : 0.0;
  }
}

// Cell key of a segment's main pixel: (event, plane, y+halo, x+halo) in the halo-extended grid.
__device__ __forceinline__ int64_t cell_of(const Dev& d, int ev, int plane, int px, int py) {
  if (ev < 0 || ev >= d.n_events) return -1;
  if (px < -d.halo || px >= d.nx + d.halo || py < -d.halo || py >= d.ny + d.halo) return -1;
  return (((int64_t)ev * d.ntpc + plane) * d.gy + (py + d.halo)) * d.gx + (px + d.halo);
}

__device__ __forceinline__ int64_t pixel_id(const Dev& d, int ev, int plane, int px, int py) {
  return (((int64_t)ev * d.ntpc + plane) * d.ny + py) * d.nx + px;
}

}  // namespace lsim
