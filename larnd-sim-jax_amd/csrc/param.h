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
#pragma once
#include "device.h"

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
      const double cur = a * param_tail(gj, -st0, b, sig, dt, T, diffusion != 0) +
                         (1 - a) * param_tail(gj, -st0, c, sig, dt, T, diffusion != 0);
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

}  // namespace lsim
