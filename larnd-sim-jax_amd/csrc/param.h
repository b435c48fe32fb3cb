// Parametrized current model (simulate_parametrized, sim_jax.py:339-372; detsim_jax.py:209-228,
// 322-330, 356-399, 441-639): electrons from the drifted segments, and for every (electron,
// pixel) pair of the (2n+1)^2 pixels around the electron's pixel an analytic induced current over
// the 51 sample times linspace(0, 5, 51) us, scattered into the pixel waveforms.  The current of a
// pair is a mix a I_b + (1 - a) I_c of two exponential tails with decay constants b, c (quadratic
// surfaces in the pair's distances from the pixel centre), each integrated over the tick around
// every sample time; with diffusion in the current the tails are exponentially modified Gaussians
// of the longitudinal spread.  All arithmetic is float64.
//
// Layout: electrons in (segment, sub-electron) order as SoA float64 arrays; pairs are (electron,
// offset (a, b) in row-major order); the rows (sorted unique pixel ids) list their pairs in pair
// order (a stable sort by row, on the host side), and one workgroup per row adds the currents of
// its pairs into per-wave float64 rows in LDS (lane k = sample k), summed in a fixed order:
// deterministic, no atomics.
//
// Adjoint (reverse mode, as jax.grad takes it through simulate_parametrized): the readout's
// step-function cotangent of every row (k_fee_adjoint) plus an optional cotangent on the
// waveforms; per pair, wave sums of the cotangent times the partial derivatives of the current
// in the pair's charge, distances, arrival time and longitudinal spread; per electron, factors
// that carry those to the 7 parameters (k_param_adj_electrons).  Integer decisions (pixel ids,
// the whole-tick split of t0, the readout's decisions) are held at their forward values.
#pragma once
#include "backward.h"

namespace lsim {

constexpr int kParamMaxSamples = kWave - 1;  // samples of one pair's current: one per lane, and a lane for the last tick boundary
constexpr int kParamWaves = 4;
constexpr int kParamMaxEdges = 17;    // transverse bin edges (16 bins)

struct ParamE {
  double *x, *y;  // electron position (cm, shifted)
  double* t0;     // drift time to the anode (us)
  double* ne;     // charge (e-)
  double* sig;    // longitudinal spread in time (us)
  int* plane;
};

// Electrons of one segment: mode 0 = the transverse bins of apply_tran_diff (detsim_jax.py:356-373:
// positions shifted by whole bin widths, charges times w_i w_j of the bins between `edges`, the
// outer edges at +-infinity); mode 1 = one electron spread by the caller's standard normals g
// (mc_diff, detsim_jax.py:377-399: x, y += g sigma_T; z += g sigma_L when the current carries no
// diffusion).  Then the pair pixel ids (detsim_jax.py:483-496, -1 outside the plane).
__global__ void k_param_electrons(Dev d, SegIn in, int mode, int diffusion, const float* __restrict__ normals,
                                  const double* __restrict__ edges, int n_edges, ParamE e, int64_t* __restrict__ pids) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= in.n) return;
  Phys ph;
  seg_physics(d, in, s, ph);
  const int64_t st = in.stride;
  double x = (double)ldcol(in.x, s, st) - d.shift[0];
  double y = (double)ldcol(in.y, s, st) - d.shift[1];
  double z = (double)ldcol(in.z, s, st) - d.shift[2];
  const int K = (2 * d.nn + 1) * (2 * d.nn + 1);
  const int nb = mode == 0 ? n_edges - 1 : 1;
  double w[kParamMaxEdges - 1];
  double sym = 0.0, bw = 0.0;
  if (mode == 0) {
    // weights of the bins between the edges (detsim_jax.py:322-330)
    double prev = -1.0;
    for (int k = 0; k < nb; ++k) {
      const double next = k + 1 == nb ? 1.0 : erf(edges[k + 1] / (M_SQRT2 * ph.st));
      w[k] = 0.5 * (next - prev);
      prev = next;
    }
    sym = (double)n_edges / 2.0 - 1.0;
    bw = edges[1] - edges[0];
  } else {
    x += (double)normals[s * 3 + 0] * ph.st;
    y += (double)normals[s * 3 + 1] * ph.st;
    if (!diffusion) z += (double)normals[s * 3 + 2] * ph.sl;
    w[0] = 1.0;
  }
  const double z_an = d.tpc[ph.plane][2][0];
  const double t0 = fabs(z - z_an) / d.v;
  const double sig = ph.sl / d.v;
  for (int i = 0; i < nb; ++i) {
    for (int j = 0; j < nb; ++j) {
      const int64_t el = s * nb * nb + i * nb + j;
      const double ex = mode == 0 ? x + (-sym + i) * bw : x;
      const double ey = mode == 0 ? y + (-sym + j) * bw : y;
      e.x[el] = ex;
      e.y[el] = ey;
      e.t0[el] = t0;
      e.ne[el] = mode == 0 ? ph.ne * (w[i] * w[j]) : ph.ne;
      e.sig[el] = sig;
      e.plane[el] = ph.plane;
      const double fx = floor((ex - d.tpc[ph.plane][0][0]) / d.pitch), fy = floor((ey - d.tpc[ph.plane][1][0]) / d.pitch);
      const int px = to_int_clamped(fx), py = to_int_clamped(fy);
      int k = 0;
      for (int a = -d.nn; a <= d.nn; ++a)
        for (int b = -d.nn; b <= d.nn; ++b, ++k) {
          const int qx = px + a, qy = py + b;
          const bool out = qx < 0 || qy < 0 || qx >= d.nx || qy >= d.ny;
          pids[el * K + k] = out ? (int64_t)-1 : pixel_id(d, ph.ev, ph.plane, qx, qy);
        }
    }
  }
}

// 0.5 erf((X - mu) / (sqrt2 sigma)) - emg(X; mu, sigma, lambda) / lambda: the exponentially modified
// Gaussian's CDF minus 1/2 (detsim_jax.py:441-476), whose differences are the tick integrals.
__device__ __forceinline__ double emg_cdf_half(double X, double mu, double sigma, double lambd) {
  const double coeff = lambd / 2;
  const double ex = coeff * (2 * mu + lambd * sigma * sigma - 2 * X);
  const double pdf = coeff * exp(ex) * erfc((mu + lambd * sigma * sigma - X) / (M_SQRT2 * sigma));
  return 0.5 * erf((X - mu) / (M_SQRT2 * sigma)) - pdf / lambd;
}

// One wave: sample k < T of one tail (decay `scale`) of a pair's current, loc = -t0 shifted,
// integrated over the tick around sample time t_k: the integral runs between the boundaries
// g_{k+1} = -t_k - dt/2 and g_k = -t_k + dt/2 (g_j = -t_j + dt/2 for j < T, g_T = -t_{T-1} - dt/2),
// so lane j evaluates the tail's CDF (integrated_expon_diff, with diffusion) or its exponential
// (integrated_expon) once at g_j and each value serves two samples (the next lane's by a shift).
__device__ __forceinline__ double param_tail(double gj, double loc, double scale, double sigma, double dt, int T,
                                             bool diffusion) {
  if (diffusion) {
    const double F = emg_cdf_half(gj, loc, sigma, 1.0 / scale);
    const double Fn = __shfl_down(F, 1, kWave);
    const double F1 = __shfl(F, 1, kWave), FT = __shfl(F, T - 1, kWave);  // lo of sample 0, up of T - 1
    return (F - Fn) / (F1 - FT) / dt;
  }
  const double G = exp(fmin(0.0, (loc - gj) / scale));
  const double Gn = __shfl_down(G, 1, kWave);
  return (Gn - G + exp(fmin(0.0, (loc - dt / 2) / scale)) / T) / dt;
}

// Both tails of a pair's current (decays sb and sc) at sample k: param_tail of each, with the
// Gaussian part 1/2 erf((g - mu) / (sqrt2 sigma)) of the two CDFs evaluated once (the same operations
// as emg_cdf_half, so the values are the same).
__device__ __forceinline__ void param_tails(double gj, double loc, double sb, double sc, double sigma, double dt, int T,
                                            bool diffusion, double& tb, double& tc) {
  if (!diffusion) {
    tb = param_tail(gj, loc, sb, sigma, dt, T, false);
    tc = param_tail(gj, loc, sc, sigma, dt, T, false);
    return;
  }
  const double half = 0.5 * erf((gj - loc) / (M_SQRT2 * sigma));
  double F[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double lambd = 1.0 / (k == 0 ? sb : sc);
    const double coeff = lambd / 2;
    const double ex = coeff * (2 * loc + lambd * sigma * sigma - 2 * gj);
    const double pdf = coeff * exp(ex) * erfc((loc + lambd * sigma * sigma - gj) / (M_SQRT2 * sigma));
    F[k] = half - pdf / lambd;
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double Fn = __shfl_down(F[k], 1, kWave);
    const double F1 = __shfl(F[k], 1, kWave), FT = __shfl(F[k], T - 1, kWave);
    (k == 0 ? tb : tc) = (F[k] - Fn) / (F1 - FT) / dt;
  }
}

// One workgroup per row (sorted unique pixel id upix[r]): the pairs order[offs[r] .. offs[r+1]) in
// pair order, pair i going to wave i % kParamWaves; lane k adds sample k of the pair's current
// times its charge at waveform tick t0_tick - T + k (detsim_jax.py:209-228: the ticks outside
// [0, nout) go to the reference's dropped column 0 and are left out here).  Rows of id < 0 are
// zero.  wfs [P][nout] float32.
__global__ __launch_bounds__(kParamWaves* kWave) void k_param_rows(Dev d, ParamE e, const int64_t* __restrict__ upix,
                                                                    const int* __restrict__ order,
                                                                    const int* __restrict__ offs, int K, int T,
                                                                    int diffusion, float* __restrict__ wfs) {
  __shared__ double acc[kParamWaves][kMaxOut];
  const int r = blockIdx.x;
  const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
  const int nout = d.nout;
  for (int o = t; o < kParamWaves * kMaxOut; o += kParamWaves * kWave) (&acc[0][0])[o] = 0.0;
  __syncthreads();
  const int64_t pid = upix[r];
  if (pid >= 0) {
    // pixel centre (detsim_jax.py:297-306)
    const int64_t nx = d.nx, ny = d.ny;
    const int px = (int)(pid % nx), py = (int)((pid / nx) % ny), plane = (int)((pid / (nx * ny)) % d.ntpc);
    const double cx = (double)px * d.pitch + d.tpc[plane][0][0] + d.pitch / 2;
    const double cy = (double)py * d.pitch + d.tpc[plane][1][0] + d.pitch / 2;
    const double dt = 5.0 / (double)(T - 1);
    const int kk = lane < T ? lane : T - 1;
    const double tk = kk == T - 1 ? 5.0 : (double)kk * dt;  // linspace(0, 5, T)
    const double gj = lane < T ? -tk + dt / 2 : -5.0 - dt / 2;  // tick boundaries (lane T: the last)
    const int b0 = offs[r], b1 = offs[r + 1];
    for (int i = b0 + wid; i < b1; i += kParamWaves) {
      const int pr = order[i];
      const int64_t el = pr / K;
      const double xd = fabs(e.x[el] - cx), yd = fabs(e.y[el] - cy);
      // detsim_jax.py:542-545, 585-588
      double a = 1.060 - 0.909 * xd - 0.909 * yd + 5.856 * xd * yd + 0.207 * xd * xd + 0.207 * yd * yd;
      const double b = 0.679 - 1.083 * xd - 1.083 * yd + 8.772 * xd * yd - 5.521 * xd * xd - 5.521 * yd * yd;
      const double c = 2.644 - 9.174 * xd - 9.174 * yd + 13.483 * xd * yd + 45.887 * xd * xd + 45.887 * yd * yd;
      const double ts0 = 2.948 - 2.705 * xd - 2.705 * yd + 4.825 * xd * yd + 20.814 * xd * xd + 20.814 * yd * yd;
      a = fmin(a, 1.0);
      // t0 split into whole ticks and the remainder (detsim_jax.py:629-633)
      const double t0 = e.t0[el];
      const int t0_tick = (int)(t0 / d.ts + 0.5);
      const double st0 = (t0 - t0_tick * d.ts) + ts0;
      const double sig = e.sig[el];
      double tlb, tlc;
      param_tails(gj, -st0, b, c, sig, dt, T, diffusion != 0, tlb, tlc);
      const double cur = a * tlb + (1 - a) * tlc;
      const int o = t0_tick - T + kk;
      if (lane < T && o >= 0 && o < nout) acc[wid][o] += cur * e.ne[el];
    }
  }
  __syncthreads();
  for (int o = t; o < nout; o += kParamWaves * kWave) {
    double v = 0.0;
    for (int w = 0; w < kParamWaves; ++w) v += acc[w][o];
    wfs[(int64_t)r * nout + o] = (float)v;
  }
}

// ------------------------------------------------------------------------------------------
// Adjoint
// ------------------------------------------------------------------------------------------

// Chain factors of one electron: the derivatives of its charge, position, arrival time and
// longitudinal spread in the parameters that move them.
struct __align__(16) ParamAdj {
  double dne[5];   // d ne / d (Ab, kb, eField, lifetime, tran_diff)
  double dx[2];    // d x / d (eField, tran_diff)
  double dy[2];    // d y / d (eField, tran_diff)
  double dt0[2];   // d t0 / d (eField, long_diff)
  double dsig[2];  // d sigma / d (eField, long_diff)
  double pad[3];
};
static_assert(sizeof(ParamAdj) == 128, "ParamAdj layout");

// One thread per segment, in the electron order of k_param_electrons: the same physics (float64),
// then the factors.  Charge: quench, lifetime and drift time as in the LUT chain; with transverse
// bins times the bin weights w_i w_j, whose derivative in sigma_T is that of the erf at each inner
// edge.  With normal draws the position moves by g sigma_T (and z by g sigma_L when the current
// carries no diffusion).  t0 = |z - z_anode| / v and sigma = sigma_L / v move with v (eField).
__global__ void k_param_adj_electrons(Dev d, SegIn in, int mode, int diffusion, const float* __restrict__ normals,
                                      const double* __restrict__ edges, int n_edges, ParamAdj* __restrict__ adj) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= in.n) return;
  Phys ph;
  seg_physics(d, in, s, ph);
  const double v = d.v, dv = d.dv;
  const double att = exp(-ph.td / d.lifetime) * (double)ph.mask;
  const double dne_E = ph.dq_dE_rec * att + ph.ne * ph.td * dv / (v * d.lifetime);
  const double dst_E = -ph.st * dv / (2 * v), dst_T = d.tran_diff > 0 ? ph.st / (2 * d.tran_diff) : 0.0;
  const double dsl_E = -ph.sl * dv / (2 * v), dsl_L = d.long_diff > 0 ? ph.sl / (2 * d.long_diff) : 0.0;
  const int nb = mode == 0 ? n_edges - 1 : 1;
  double w[kParamMaxEdges - 1], dw[kParamMaxEdges - 1];
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (mode == 0) {
    double prev = -1.0, dprev = 0.0;
    for (int k = 0; k < nb; ++k) {
      double next = 1.0, dnext = 0.0;
      if (k + 1 < nb) {
        const double xi = edges[k + 1] / (M_SQRT2 * ph.st);
        next = erf(xi);
        dnext = M_2_SQRTPI * exp(-xi * xi) * (-xi / ph.st);
      }
      w[k] = 0.5 * (next - prev);
      dw[k] = 0.5 * (dnext - dprev);
      prev = next;
      dprev = dnext;
    }
  } else {
    gx = (double)normals[s * 3 + 0];
    gy = (double)normals[s * 3 + 1];
    if (!diffusion) gz = (double)normals[s * 3 + 2];
    w[0] = 1.0;
    dw[0] = 0.0;
  }
  const double z = (double)ldcol(in.z, s, in.stride) - d.shift[2] + gz * ph.sl;
  const double z_an = d.tpc[ph.plane][2][0];
  const double t0 = fabs(z - z_an) / v;
  const double sz = z > z_an ? 1.0 : (z < z_an ? -1.0 : 0.0);
  const double sig = ph.sl / v;
  ParamAdj A;
  A.dx[0] = gx * dst_E;
  A.dx[1] = gx * dst_T;
  A.dy[0] = gy * dst_E;
  A.dy[1] = gy * dst_T;
  A.dt0[0] = -t0 * dv / v + sz * gz * dsl_E / v;
  A.dt0[1] = sz * gz * dsl_L / v;
  A.dsig[0] = dsl_E / v - sig * dv / v;
  A.dsig[1] = dsl_L / v;
  A.pad[0] = A.pad[1] = A.pad[2] = 0.0;
  for (int i = 0; i < nb; ++i) {
    for (int j = 0; j < nb; ++j) {
      const double W = w[i] * w[j], dW = dw[i] * w[j] + w[i] * dw[j];
      A.dne[0] = ph.dq_dAb * att * W;
      A.dne[1] = ph.dq_dkb * att * W;
      A.dne[2] = dne_E * W + ph.ne * dW * dst_E;
      A.dne[3] = ph.ne * ph.td / (d.lifetime * d.lifetime) * W;
      A.dne[4] = ph.ne * dW * dst_T;
      adj[s * nb * nb + i * nb + j] = A;
    }
  }
}

// One tail of a pair's current at sample k (param_tail) and its partial derivatives in the
// tail's location, decay constant and longitudinal spread.  With diffusion the tail is
// (F_k - F_{k+1}) / (F_1 - F_{T-1}) / dt of H(g) = 1/2 erf(u) - P, u = (g - mu) / (sqrt2 sigma),
// P = 1/2 exp(E) erfc(z), E = lambda (mu - g) + lambda^2 sigma^2 / 2,
// z = (mu + lambda sigma^2 - g) / (sqrt2 sigma), lambda = 1 / scale; since E - z^2 = -u^2:
//   dH/dmu = -lambda P,  dH/dsigma = lambda (phi(u) / (2 sqrt2) - lambda sigma P),
//   dH/dlambda = phi(u) sigma / (2 sqrt2) - (lambda sigma^2 - (g - mu)) P,  phi(u) = 2/sqrt(pi) exp(-u^2),
// and the quotient rule over the neighbouring lanes' values.  Every lane divides its H, P and phi
// by the shared denominator F_1 - F_{T-1} before the partials are formed: near the float64 limit
// (P ~ 1e306 for a decay constant near zero) the tail is finite while lambda P alone would not be;
// past it (an infinite denominator) the samples are 0 and their partials are taken as 0.
// Without diffusion the tail is (G_{k+1} - G_k + G_e / T) / dt with G = exp(min(0, (mu - g) / scale)).
struct TailD {
  double v, dloc, dscale, dsig;
};

__device__ __forceinline__ TailD param_tail_d(double gj, double loc, double scale, double sigma, double dt, int T,
                                              bool diffusion) {
  TailD r;
  if (diffusion) {
    const double lam = 1.0 / scale;
    const double D = gj - loc;
    const double u = D / (M_SQRT2 * sigma);
    const double E = lam * (-D) + 0.5 * lam * lam * sigma * sigma;
    const double z = (loc + lam * sigma * sigma - gj) / (M_SQRT2 * sigma);
    const double P = 0.5 * exp(E) * erfc(z);
    const double F = 0.5 * erf(u) - P;
    const double den = __shfl(F, 1, kWave) - __shfl(F, T - 1, kWave);
    const double p = P / den, f = F / den;
    const double ph = M_2_SQRTPI * exp(-u * u) * (0.5 * M_SQRT1_2) / den;  // phi(u) / (2 sqrt2) / den
    const double Fm = -lam * p;
    const double Fs = lam * (ph - lam * sigma * p);
    const double Fc = -lam * (lam * (ph * sigma - (lam * sigma * sigma - D) * p));
    const double fn = __shfl_down(f, 1, kWave), Fmn = __shfl_down(Fm, 1, kWave);
    const double Fsn = __shfl_down(Fs, 1, kWave), Fcn = __shfl_down(Fc, 1, kWave);
    const double Fm1 = __shfl(Fm, 1, kWave), FmT = __shfl(Fm, T - 1, kWave);
    const double Fs1 = __shfl(Fs, 1, kWave), FsT = __shfl(Fs, T - 1, kWave);
    const double Fc1 = __shfl(Fc, 1, kWave), FcT = __shfl(Fc, T - 1, kWave);
    const double q = f - fn;
    r.v = q / dt;
    if (!isfinite(den)) {
      // |den| beyond the float64 range: the samples are 0 (or NaN, and the row with them), and so
      // are their derivatives in the limit; the overflowing lanes would give NaN here
      r.dloc = r.dsig = r.dscale = 0.0;
      return r;
    }
    r.dloc = ((Fm - Fmn) - q * (Fm1 - FmT)) / dt;
    r.dsig = ((Fs - Fsn) - q * (Fs1 - FsT)) / dt;
    r.dscale = ((Fc - Fcn) - q * (Fc1 - FcT)) / dt;
    return r;
  }
  const double y = (loc - gj) / scale;
  const double G = exp(fmin(0.0, y));
  const double Gl = y < 0 ? G / scale : 0.0;
  const double Gc = y < 0 ? -G * y / scale : 0.0;
  const double ye = (loc - dt / 2) / scale;
  const double Ge = exp(fmin(0.0, ye));
  const double Gel = ye < 0 ? Ge / scale : 0.0;
  const double Gec = ye < 0 ? -Ge * ye / scale : 0.0;
  const double Gn = __shfl_down(G, 1, kWave), Gln = __shfl_down(Gl, 1, kWave), Gcn = __shfl_down(Gc, 1, kWave);
  r.v = (Gn - G + Ge / T) / dt;
  r.dloc = (Gln - Gl + Gel / T) / dt;
  r.dscale = (Gcn - Gc + Gec / T) / dt;
  r.dsig = 0.0;
  return r;
}

// One workgroup per row, over the same pairs as k_param_rows: the row's cotangent G(o) (the
// readout's step function plus the optional waveform cotangent gw[r][o]) in LDS; per pair, lane k
// takes G at the pair's tick o and the partials of sample k of the current; wave sums give
// d loss / d (charge, x, y, t0, sigma) of the pair, and the electron's factors the 7 gradients.
// Lanes with G(o) = 0 add nothing (a pair that overflows to NaN then adds no NaN where the loss
// does not see it).  partial[r][7]: the row's sums, waves combined in index order.
__global__ __launch_bounds__(kParamWaves* kWave) void k_param_rows_adj(Dev d, ParamE e, const ParamAdj* __restrict__ adj,
                                                                        const int64_t* __restrict__ upix,
                                                                        const int* __restrict__ order,
                                                                        const int* __restrict__ offs, int K, int T,
                                                                        int diffusion, const Pieces* __restrict__ pieces,
                                                                        const float* __restrict__ gw,
                                                                        double* __restrict__ partial) {
  constexpr int NT = kParamWaves * kWave;
  __shared__ double G[kMaxOut];
  __shared__ double red[kNGrad][kParamWaves];
  __shared__ int s_any;
  const int r = blockIdx.x;
  const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
  const int nout = d.nout;
  if (t == 0) s_any = 0;
  __syncthreads();
  {
    const Pieces* pr = pieces + r;
    const int nf = pr->n;
    bool any = false;
    for (int o = t; o < nout; o += NT) {
      double cf = 0.0;
      for (int k = nf - 1; k >= 0; --k)
        if (pr->x[k] >= o) cf = (double)pr->c[k];
      const double v = cf + (gw ? (double)gw[(int64_t)r * nout + o] : 0.0);
      G[o] = v;
      any = any || v != 0.0;
    }
    if (any) s_any = 1;
  }
  __syncthreads();
  double g[kNGrad];
#pragma unroll
  for (int k = 0; k < kNGrad; ++k) g[k] = 0.0;
  const int64_t pid = upix[r];
  if (s_any && pid >= 0) {
    const int64_t nx = d.nx, ny = d.ny;
    const int px = (int)(pid % nx), py = (int)((pid / nx) % ny), plane = (int)((pid / (nx * ny)) % d.ntpc);
    const double cx = (double)px * d.pitch + d.tpc[plane][0][0] + d.pitch / 2;
    const double cy = (double)py * d.pitch + d.tpc[plane][1][0] + d.pitch / 2;
    const double dt = 5.0 / (double)(T - 1);
    const int kk = lane < T ? lane : T - 1;
    const double tk = kk == T - 1 ? 5.0 : (double)kk * dt;
    const double gj = lane < T ? -tk + dt / 2 : -5.0 - dt / 2;
    const int b0 = offs[r], b1 = offs[r + 1];
    for (int i = b0 + wid; i < b1; i += kParamWaves) {
      const int pr = order[i];
      const int64_t el = pr / K;
      const double t0 = e.t0[el];
      const int t0_tick = (int)(t0 / d.ts + 0.5);
      const int o = t0_tick - T + kk;
      const double gv = (lane < T && o >= 0 && o < nout) ? G[o] : 0.0;
      if (__ballot(gv != 0.0) == 0) continue;  // the same pair in every lane
      const double ex = e.x[el] - cx, ey = e.y[el] - cy;
      const double xd = fabs(ex), yd = fabs(ey);
      const double ar = 1.060 - 0.909 * xd - 0.909 * yd + 5.856 * xd * yd + 0.207 * xd * xd + 0.207 * yd * yd;
      const double b = 0.679 - 1.083 * xd - 1.083 * yd + 8.772 * xd * yd - 5.521 * xd * xd - 5.521 * yd * yd;
      const double c = 2.644 - 9.174 * xd - 9.174 * yd + 13.483 * xd * yd + 45.887 * xd * xd + 45.887 * yd * yd;
      const double ts0 = 2.948 - 2.705 * xd - 2.705 * yd + 4.825 * xd * yd + 20.814 * xd * xd + 20.814 * yd * yd;
      const bool clamp = ar > 1.0;
      const double a = clamp ? 1.0 : ar;
      const double ax = clamp ? 0.0 : -0.909 + 5.856 * yd + 2 * 0.207 * xd;
      const double ay = clamp ? 0.0 : -0.909 + 5.856 * xd + 2 * 0.207 * yd;
      const double bx = -1.083 + 8.772 * yd - 2 * 5.521 * xd, by = -1.083 + 8.772 * xd - 2 * 5.521 * yd;
      const double cxd = -9.174 + 13.483 * yd + 2 * 45.887 * xd, cyd = -9.174 + 13.483 * xd + 2 * 45.887 * yd;
      const double tx = -2.705 + 4.825 * yd + 2 * 20.814 * xd, ty = -2.705 + 4.825 * xd + 2 * 20.814 * yd;
      const double st0 = (t0 - t0_tick * d.ts) + ts0;
      const double sig = e.sig[el];
      const TailD tb = param_tail_d(gj, -st0, b, sig, dt, T, diffusion != 0);
      const TailD tc = param_tail_d(gj, -st0, c, sig, dt, T, diffusion != 0);
      const double cur = a * tb.v + (1 - a) * tc.v;
      const double c_t0 = -(a * tb.dloc + (1 - a) * tc.dloc);
      const double c_sig = a * tb.dsig + (1 - a) * tc.dsig;
      const double c_x = ax * (tb.v - tc.v) + a * tb.dscale * bx + (1 - a) * tc.dscale * cxd + c_t0 * tx;
      const double c_y = ay * (tb.v - tc.v) + a * tb.dscale * by + (1 - a) * tc.dscale * cyd + c_t0 * ty;
      const double ne = e.ne[el];
      const bool live = gv != 0.0;
      const double q_ne = wave_sum(live ? gv * cur : 0.0);
      const double q_x = wave_sum(live ? gv * ne * c_x : 0.0);
      const double q_y = wave_sum(live ? gv * ne * c_y : 0.0);
      const double q_t0 = wave_sum(live ? gv * ne * c_t0 : 0.0);
      const double q_sig = wave_sum(live ? gv * ne * c_sig : 0.0);
      // d |x - cx| / d x = sign(x - cx) (0 at 0)
      const double gx = ex > 0 ? q_x : (ex < 0 ? -q_x : 0.0);
      const double gy = ey > 0 ? q_y : (ey < 0 ? -q_y : 0.0);
      const ParamAdj* A = adj + el;
      g[0] += q_ne * A->dne[0];
      g[1] += q_ne * A->dne[1];
      g[2] += q_ne * A->dne[2] + gx * A->dx[0] + gy * A->dy[0] + q_t0 * A->dt0[0] + q_sig * A->dsig[0];
      g[3] += q_ne * A->dne[3];
      g[5] += q_t0 * A->dt0[1] + q_sig * A->dsig[1];
      g[6] += q_ne * A->dne[4] + gx * A->dx[1] + gy * A->dy[1];
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kNGrad; ++k) red[k][wid] = g[k];
  }
  __syncthreads();
  if (t < kNGrad) {
    double v = 0.0;
    for (int w = 0; w < kParamWaves; ++w) v += red[t][w];
    partial[(int64_t)r * kNGrad + t] = v;
  }
}

}  // namespace lsim
