"""FP64 CPU oracle of the parametrized current model -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module (the product path never does).  It restates, in NumPy
float64, the reference's second simulation mode (``simulate_parametrized``, sim_jax.py:339-372):
electrons from the drifted segments, the (2n+1)^2 pixels around each electron, the analytic
induced current of each (electron, pixel) pair over 51 ticks (detsim_jax.py:441-639), the
scatter of those currents into pixel waveforms (detsim_jax.py:209-228) and the self-trigger
readout shared with the LUT path (``oracle.simulate_stochastic``).

The random electron positions of ``mc_diff`` (detsim_jax.py:377-399, jax.random.normal) are an
input here: the caller passes the standard-normal draws, so that the oracle and the GPU see the
same electrons (jax's threefry stream is not restated).

Parity pinning (tests/test_param_oracle.py): the tick integrals of the exponentially modified
Gaussian against scipy.stats.exponnorm (its CDF, independent of the reference's erf/erfc
formula), the no-diffusion current against its closed form, the sigma -> 0 limit of the
diffusion model, the transverse-bin weights summing to one, and charge conservation of the
scatter (every current value lands in exactly one waveform tick or the dropped column 0).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf, erfc

from . import oracle as O

# detsim_jax.py:542-545, 585-588: quadratic surfaces in the pair's distances (x, y) from the pixel
# centre: mixing weight a, the two decay constants b and c (us) and the t0 shift
B_PARAMS = (1.060, -0.909, -0.909, 5.856, 0.207, 0.207)
C_PARAMS = (0.679, -1.083, -1.083, 8.772, -5.521, -5.521)
D_PARAMS = (2.644, -9.174, -9.174, 13.483, 45.887, 45.887)
T0_PARAMS = (2.948, -2.705, -2.705, 4.825, 20.814, 20.814)


def _surface(c, x, y):
    return c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * x * x + c[5] * y * y


def current_ticks(p) -> np.ndarray:
    """detsim_jax.py:624-625 -- the 51 sample times linspace(0, 5, int(5 / t_sampling) + 1)."""
    n = int(5 / p.t_sampling) + 1
    return np.linspace(0, 5, n)


def emg_pdf(x, mu, sigma, lambd):
    """detsim_jax.py:441-460 -- exponentially modified Gaussian density."""
    coeff = lambd / 2
    exponent = coeff * (2 * mu + lambd * sigma ** 2 - 2 * x)
    return coeff * np.exp(exponent) * erfc((mu + lambd * sigma ** 2 - x) / (np.sqrt(2) * sigma))


def integrated_expon_diff(x, loc, scale, diff, dt):
    """detsim_jax.py:463-476 -- per-tick integrals of the EMG over [x - dt/2, x + dt/2], divided by
    (lower bound at the first tick - upper bound at the last tick), then by dt.  x (..., T)."""
    lambd = 1 / scale
    a, b = x - dt / 2, x + dt / 2
    upper = 0.5 * erf((b - loc) / (np.sqrt(2) * diff)) - emg_pdf(b, loc, diff, lambd) / lambd
    lower = 0.5 * erf((a - loc) / (np.sqrt(2) * diff)) - emg_pdf(a, loc, diff, lambd) / lambd
    tick = upper - lower
    tick = tick / (lower[..., 0] - upper[..., -1])[..., None]
    return tick / dt


def integrated_expon(x, loc, scale, dt):
    """detsim_jax.py:534-540 -- the no-diffusion tick integrals of a one-sided exponential."""
    return (np.exp(np.minimum(0.0, (loc - x + dt / 2) / scale))
            - np.exp(np.minimum(0.0, (loc - x - dt / 2) / scale))
            + np.exp(np.minimum(0.0, (loc - dt / 2) / scale)) / x.shape[-1]) / dt


def current_model(t, t0, x, y, dt, sigma=None):
    """detsim_jax.py:544-622 -- induced current of a charge arriving at t0 at distances (x, y) from
    the pixel centre, at times t (..., T): current_model without diffusion (sigma None),
    current_model_diff with the longitudinal spread sigma (us)."""
    a = np.minimum(_surface(B_PARAMS, x, y), 1)
    b = _surface(C_PARAMS, x, y)
    c = _surface(D_PARAMS, x, y)
    st0 = t0 + _surface(T0_PARAMS, x, y)
    if sigma is None:
        return a * integrated_expon(-t, -st0, b, dt) + (1 - a) * integrated_expon(-t, -st0, c, dt)
    return (a * integrated_expon_diff(-t, -st0, b, sigma, dt)
            + (1 - a) * integrated_expon_diff(-t, -st0, c, sigma, dt))


def gaussian_1d_integral(edges, std):
    """detsim_jax.py:322-330 -- weights of the bins between edges (outer edges at -/+ infinity)."""
    e = np.ones((std.shape[0], len(edges)))
    e[:, 0] = -1
    e[:, 1:-1] = erf(np.asarray(edges[1:-1])[None] / (np.sqrt(2) * std[:, None]))
    return 0.5 * (e[:, 1:] - e[:, :-1])


def electrons(p, tracks, fields, normals=None, bin_edges=None):
    """simulate_drift (sim_jax.py:123-139) up to the electrons: shift, quench, drift, then either
    the Monte-Carlo spread (mc_diff, detsim_jax.py:377-399: x, y += N(0,1) sigma_T, and z += N(0,1)
    sigma_L when the current model carries no diffusion) with the given draws ``normals`` (N, 3),
    or the transverse bins of apply_tran_diff (detsim_jax.py:356-373) on ``bin_edges``.
    Returns a dict of per-electron float64 arrays."""
    q = O.segment_quantities(p, tracks, fields)
    x, y, z, ne = q["x"], q["y"], q["z"], q["n_e"]
    sl, st, plane, ev = q["sigma_l"], q["sigma_t"], q["plane"], q["ev"]
    if p.mc_diff:
        if normals is None:
            raise ValueError("mc_diff needs the normal draws")
        g = np.asarray(normals, np.float64)
        x = x + g[:, 0] * st
        y = y + g[:, 1] * st
        if not p.diffusion_in_current_sim:
            z = z + g[:, 2] * sl
        return dict(x=x, y=y, z=z, ne=ne, sl=sl, plane=plane, ev=ev)
    if bin_edges is None:
        raise ValueError("apply_tran_diff needs tran_diff_bin_edges")
    edges = np.asarray(bin_edges, np.float64)
    w = gaussian_1d_integral(edges, st)
    nb = len(edges) - 1
    sym = len(edges) / 2 - 1
    sh = np.arange(-sym, sym + 1)
    bw = edges[1] - edges[0]
    S = len(sh)
    assert S == nb
    sx = np.repeat(sh, S)
    sy = np.tile(sh, S)
    rep = lambda a: np.repeat(a, S * S)  # noqa: E731
    W = (w[:, :, None] * w[:, None, :]).reshape(-1)
    return dict(x=rep(x) + np.tile(sx * bw, len(x)), y=rep(y) + np.tile(sy * bw, len(y)), z=rep(z),
                ne=rep(ne) * W, sl=rep(sl), plane=rep(plane), ev=rep(ev))


def get_pixels(p, e):
    """detsim_jax.py:483-496 -- the (2n+1)^2 pixel ids around each electron's pixel, in (a, b) order
    (-1 outside the plane)."""
    b = O._borders(p)[e["plane"]]
    px = np.floor_divide(e["x"] - b[:, 0, 0], p.pixel_pitch).astype(np.int64)
    py = np.floor_divide(e["y"] - b[:, 1, 0], p.pixel_pitch).astype(np.int64)
    n = int(p.number_pix_neighbors)
    a_off = np.repeat(np.arange(-n, n + 1), 2 * n + 1)
    b_off = np.tile(np.arange(-n, n + 1), 2 * n + 1)
    return O.pixel2id(p, px[:, None] + a_off[None], py[:, None] + b_off[None], e["plane"][:, None], e["ev"][:, None])


def current_mc(p, e, pids):
    """detsim_jax.py:623-639 -- (t0_tick (E,), signals (E, K, T)) of every (electron, pixel) pair."""
    t = current_ticks(p)
    T = len(t)
    X, Y, plane, _ = O.id2pixel(p, pids)
    cx, cy = O.get_pixel_coordinates(p, X, Y, plane)
    xd = np.abs(e["x"][:, None] - cx)
    yd = np.abs(e["y"][:, None] - cy)
    v = O.vdrift(p)
    z_an = O._borders(p)[e["plane"], 2, 0]
    t0 = np.abs(e["z"] - z_an) / v
    t0_tick = (t0 / p.t_sampling + 0.5).astype(np.int64)
    t0 = t0 - t0_tick * p.t_sampling
    dt = 5.0 / (T - 1)
    sigma = (e["sl"] / v)[:, None, None] if p.diffusion_in_current_sim else None
    cur = current_model(t[None, None, :], t0[:, None, None], xd[..., None], yd[..., None], dt, sigma)
    return t0_tick, cur * e["ne"][:, None, None]


def accumulate(p, U, pids, t0_tick, signals):
    """detsim_jax.py:209-228 + sim_jax.py:316-324 -- scatter-add into (P, nticks + 1) waveforms, the
    ticks outside [0, nticks) going to the dropped column 0, rows by searchsorted of the pair's
    pixel in U.  Pairs whose pixel is outside the plane (id -1) land on the first -1 row."""
    nt = int(p.time_interval[1] / p.t_sampling) + 1
    E, K, T = signals.shape
    start = t0_tick - T
    tt = start[:, None] + np.arange(T)[None]
    tt = np.where((tt < 0) | (tt >= nt - 1), 0, tt + 1)
    rows = np.searchsorted(U, pids)
    flat = (rows[:, :, None] * nt + tt[:, None, :]).ravel()
    wfs = np.bincount(flat, weights=signals.ravel(), minlength=len(U) * nt)
    return wfs.reshape(len(U), nt)


def simulate_parametrized(p, tracks, fields, normals=None, bin_edges=None):
    """sim_jax.py:339-372 (noise off): the compacted hits, the sorted padded pixel set and the
    waveforms (the pixel set padded by pad_size with -1 in front, like the reference's sort)."""
    e = electrons(p, tracks, fields, normals, bin_edges)
    pids = get_pixels(p, e)
    U = np.unique(pids)
    t0_tick, sig = current_mc(p, e, pids)
    wfs = accumulate(p, U, pids, t0_tick, sig)
    out = O.simulate_stochastic(p, wfs[:, 1:], U)
    out.update(wfs=wfs, unique_pixels=U, electrons=e, pair_pids=pids, t0_tick=t0_tick, signals=sig)
    return out
