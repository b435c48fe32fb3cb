"""FP64 CPU oracle of the larnd-sim-jax LUT simulation chain -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the HIP implementation.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it; the product
path never does.

It restates the reference algorithm in NumPy, in float64 (``dtype`` can be lowered to float32
for the CPU-baseline timing, which then follows the reference's own fp32 arithmetic more
closely).  Each function cites the reference file:line it follows.  Structure follows the
reference: per-segment quantities (``simulate_drift_new``), the sorted unique main-pixel set
(``simulate_wfs``), materialised per-contribution values and indices summed with
``np.bincount`` (the ``segment_sum`` of ``simulate_signals``), the self-trigger loop of
``get_adc_values`` and the hit compaction of ``parse_output``.

Parity pinning (see tests/test_oracle.py): known answers from the reference's fixture
``output/jax_ref/output_lut_1.npz`` (digitize(0), pixel-centre x/y, hit z at tick 1998),
get_vdrift at the module0 field, closed forms (Birks, erf weights summing to 1, sigma->0
limits), brute-force pixel/bin enumeration, the independent scalar loop oracle
``simulate_wfs_loop`` on small events, linearity in charge, the charge-conservation identity
of the window + cumsum correction, and hand-built FEE waveforms with known crossings.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

# --------------------------------------------------------------------------------------
# constants and small helpers
# --------------------------------------------------------------------------------------


def vdrift(p) -> float:
    """consts_jax.py:192-216 -- mobility parameterisation, cm/us."""
    a0, a1, a2, a3, a4, a5 = p.ELECTRON_MOBILITY_PARAMS
    E = p.eField
    num = a0 + a1 * E + a2 * pow(E, 1.5) + a3 * pow(E, 2.5)
    denom = 1 + (a1 / a0) * E + a4 * pow(E, 2) + a5 * pow(E, 3)
    temp_corr = pow(p.temperature / 89, -1.5)
    return num / denom * temp_corr / 1000 * E


def _borders(p):
    return np.asarray(p.tpc_borders, dtype=np.float64)


def _sigmas(p):
    if p.long_diff_template is not None:
        return np.asarray(p.long_diff_template, dtype=np.float32).astype(np.float64)
    return np.linspace(0.001, 10, 100).astype(np.float32).astype(np.float64)


def pixel2id(p, px, py, plane, ev):
    """detsim_jax.py:230-244 -- packed pixel id, -1 outside the pixel plane (int64 here)."""
    nx, ny = int(p.n_pixels_x), int(p.n_pixels_y)
    ntpc = _borders(p).shape[0]
    px, py = np.asarray(px, np.int64), np.asarray(py, np.int64)
    outside = (px >= nx) | (py >= ny) | (px < 0) | (py < 0)
    pid = (np.asarray(ev, np.int64) * ntpc + np.asarray(plane, np.int64)) * ny + py
    pid = pid * nx + px
    return np.where(outside, np.int64(-1), pid)


def id2pixel(p, pid):
    """detsim_jax.py:265-278 -- (x, y, plane, event) with floor semantics for negatives."""
    nx, ny = int(p.n_pixels_x), int(p.n_pixels_y)
    ntpc = _borders(p).shape[0]
    pid = np.asarray(pid, np.int64)
    return pid % nx, (pid // nx) % ny, (pid // (nx * ny)) % ntpc, pid // (nx * ny * ntpc)


def get_pixel_coordinates(p, xpitch, ypitch, plane):
    """detsim_jax.py:296-306 -- pixel centre coordinates."""
    b = _borders(p)[np.asarray(plane, np.int64)]
    px = xpitch * p.pixel_pitch + b[..., 0, 0] + p.pixel_pitch / 2
    py = ypitch * p.pixel_pitch + b[..., 1, 0] + p.pixel_pitch / 2
    return px, py


def get_hit_z(p, ticks, plane):
    """detsim_jax.py:308-319 -- hit z from the tick (anode + tick*dt*v*sign)."""
    b = _borders(p)[np.asarray(plane, np.int64)]
    z_an, z_hi = b[..., 2, 0], b[..., 2, 1]
    return z_an + ticks * p.t_sampling * vdrift(p) * np.sign(z_hi - z_an)


# --------------------------------------------------------------------------------------
# per-segment quantities  (sim_jax.py:374-453)
# --------------------------------------------------------------------------------------


def _col(tracks, fields, name, dtype):
    return np.asarray(tracks[:, fields.index(name)], dtype=dtype)


def quench(p, dEdx, dE, cosphi=None):
    """quenching_jax.py:18-75 -- recombination factor and n_electrons."""
    mode = getattr(p.recombination_mode, "name", p.recombination_mode)
    if mode in ("BIRKS", 2):
        R = p.Ab / (1 + p.kb * dEdx / (p.eField * p.lArDensity))
    elif mode in ("BOX", 1):
        csi = p.beta * dEdx / (p.eField * p.lArDensity)
        R = np.maximum(0, np.log(p.alpha + csi) / csi)
    elif mode in ("ELLIPSOID", 3):
        b_phi = p.beta / np.sqrt(1 - cosphi ** 2 + (1.0 / p.R_param ** 2) * cosphi ** 2)
        csi = b_phi * dEdx / (p.eField * p.lArDensity)
        R = np.maximum(0, np.log(p.alpha + csi) / (csi + 1e-10))
    else:
        raise ValueError(f"Invalid recombination mode {p.recombination_mode}")
    return R * dE * p.MeVToElectrons


def drift(p, x, y, z, n_e):
    """drifting_jax.py:93-132 -- TPC choice, lifetime attenuation, diffusion widths."""
    b = _borders(p)
    m = p.size_margin
    zmin = np.minimum(b[:, 2, 1] - m, b[:, 2, 0] - m)
    zmax = np.maximum(b[:, 2, 1] + m, b[:, 2, 0] + m)
    cond = x[:, None] >= b[None, :, 0, 0] - m
    cond &= x[:, None] <= b[None, :, 0, 1] + m
    cond &= y[:, None] >= b[None, :, 1, 0] - m
    cond &= y[:, None] <= b[None, :, 1, 1] + m
    cond &= z[:, None] >= zmin[None]
    cond &= z[:, None] <= zmax[None]
    mask = cond.sum(axis=-1) >= 1
    plane = cond.astype(np.int64).argmax(axis=-1)
    z_an = b[plane, 2, 0]
    t_d = (np.abs(z - z_an) + 1e-6) / vdrift(p)
    n_e = n_e * np.exp(-t_d / p.lifetime) * mask
    sigma_l = np.sqrt(t_d * 2 * p.long_diff)
    sigma_t = np.sqrt(t_d * 2 * p.tran_diff)
    return plane, mask, n_e, t_d, sigma_l, sigma_t


def bin_edges(p):
    """sim_jax.py:398-403 -- diffusion bin edges relative to the sub-bin start."""
    nb = int(p.nb_tran_diff_bins)
    sym = (nb - 1) // 2
    w = p.pixel_pitch / p.nb_sampling_bins_per_pixel
    return np.linspace(-sym * w, (sym + 1) * w, nb + 1)


def gaussian_1d_integral(edges, x0, sigma):
    """detsim_jax.py:332-341 -- erf bin integrals; outer bins absorb the tails."""
    e = edges[None, :] - x0[:, None]
    E = np.ones_like(e)
    E[:, 0] = -1
    E[:, 1:-1] = erf(e[:, 1:-1] / (np.sqrt(2) * sigma[:, None]))
    return 0.5 * (E[:, 1:] - E[:, :-1])


def segment_quantities(p, tracks, fields, dtype=np.float64):
    """sim_jax.py:374-453 (+ quench, drift, get_bin_shifts, density_2d, pixel2id).

    Returns a dict of per-segment, per-diffusion-bin and per-neighbour arrays, plus the
    per-segment template interpolation quantities of simulate_signals (sim_jax.py:157-168).
    """
    c = lambda n: _col(tracks, fields, n, dtype)  # noqa: E731
    x = c("x") - p.shift_x
    y = c("y") - p.shift_y
    z = c("z") - p.shift_z
    ev = np.asarray(tracks[:, fields.index("eventID")]).astype(np.int64)
    cosphi = None
    mode = getattr(p.recombination_mode, "name", p.recombination_mode)
    if mode in ("ELLIPSOID", 3):
        zs = c("z_start") - p.shift_z
        ze = c("z_end") - p.shift_z
        cosphi = np.abs(ze - zs) / (c("dx") + 1e-10)
    n_e = quench(p, c("dEdx"), c("dE"), cosphi)
    plane, mask, n_e, t_d, sigma_l, sigma_t = drift(p, x, y, z, n_e)

    b = _borders(p)
    nsb = int(p.nb_sampling_bins_per_pixel)
    w = p.pixel_pitch / nsb
    X0, Y0 = b[plane, 0, 0], b[plane, 1, 0]
    bx = np.floor_divide(x - X0, w).astype(np.int64)
    by = np.floor_divide(y - Y0, w).astype(np.int64)
    x0 = np.remainder(x - X0, w)
    y0 = np.remainder(y - Y0, w)
    edges = bin_edges(p)
    wx = gaussian_1d_integral(edges, x0, sigma_t)
    wy = gaussian_1d_integral(edges, y0, sigma_t)
    W = np.einsum("ij,ik->ijk", wx, wy)
    nb = int(p.nb_tran_diff_bins)
    sym = (nb - 1) // 2
    sh = np.arange(-sym, sym + 1)
    sx = bx[:, None, None] + sh[None, :, None] + 0 * sh[None, None, :]
    sy = by[:, None, None] + 0 * sh[None, :, None] + sh[None, None, :]
    pix_x, pix_y = np.floor_divide(sx, nsb), np.floor_divide(sy, nsb)
    pid = pixel2id(p, pix_x, pix_y, plane[:, None, None], ev[:, None, None])
    cx = np.abs(np.remainder(sx, nsb) - nsb // 2 + 0.5).astype(np.int64)
    cy = np.abs(np.remainder(sy, nsb) - nsb // 2 + 0.5).astype(np.int64)

    v = vdrift(p)
    t0 = np.abs(z - b[plane, 2, 1]) / v
    lam = sigma_l / v / p.t_sampling

    nn = int(p.number_pix_neighbors)
    a_off = np.repeat(np.arange(-nn, nn + 1), 2 * nn + 1)
    b_off = np.tile(np.arange(-nn, nn + 1), 2 * nn + 1)
    mpx, mpy = np.floor_divide(bx, nsb), np.floor_divide(by, nsb)
    pid_nb = pixel2id(p, mpx[:, None] + a_off[None], mpy[:, None] + b_off[None], plane[:, None], ev[:, None])
    pid_nb[:, (2 * nn + 1) * nn + nn] = -999
    cx_nb = np.abs(np.remainder(bx, nsb)[:, None] - nsb // 2 + 0.5 - a_off[None] * nsb).astype(np.int64)
    cy_nb = np.abs(np.remainder(by, nsb)[:, None] - nsb // 2 + 0.5 - b_off[None] * nsb).astype(np.int64)

    # template interpolation (sim_jax.py:157-168)
    s = _sigmas(p)
    idx = np.clip(np.searchsorted(s, lam, side="left"), 1, len(s) - 2)
    s0, s1, s2 = s[idx - 1], s[idx], s[idx + 1]
    la = (lam - s1) * (lam - s2) / ((s0 - s1) * (s0 - s2))
    lb = (lam - s0) * (lam - s2) / ((s1 - s0) * (s1 - s2))
    lc = (lam - s0) * (lam - s1) / ((s2 - s0) * (s2 - s1))
    nt = None
    tau = t0 / p.t_sampling
    return dict(
        x=x, y=y, z=z, ev=ev, plane=plane, mask=mask, n_e=n_e, t_d=t_d, sigma_l=sigma_l, sigma_t=sigma_t,
        bx=bx, by=by, x0=x0, y0=y0, wx=wx, wy=wy, W=W, n_bin=(W * n_e[:, None, None]),
        pid=pid, cx=cx, cy=cy, main_pid=pid[:, sym, sym], t0=t0, lam=lam,
        pid_nb=pid_nb, cx_nb=cx_nb, cy_nb=cy_nb, idx=idx, la=la, lb=lb, lc=lc, tau=tau, nt=nt,
    )


# --------------------------------------------------------------------------------------
# response templates  (consts_jax.py:427-447)
# --------------------------------------------------------------------------------------


def gaussian_kernels(p):
    """consts_jax.py:427-428 -- normalised Gaussians over -extent..extent, one per template."""
    ext = int(p.long_diff_extent)
    k = np.arange(-ext, ext + 1, dtype=np.float64)
    s = _sigmas(p)
    g = np.exp(-0.5 * (k[None, :] / s[:, None]) ** 2) / (np.sqrt(2 * np.pi) * s[:, None])
    return g / g.sum(axis=1, keepdims=True)


def smear(series, g):
    """'same'-mode convolution along the last axis with zero padding (jnp.convolve,
    consts_jax.py:431-445): out[t] = sum_d series[t+d] * g[ext+d].  series (..., Nt), g (M, 2e+1)
    -> (M, ..., Nt)."""
    ext = (g.shape[1] - 1) // 2
    nt = series.shape[-1]
    pad = np.zeros(series.shape[:-1] + (nt + 2 * ext,), dtype=np.float64)
    pad[..., ext:ext + nt] = series
    win = np.lib.stride_tricks.sliding_window_view(pad, 2 * ext + 1, axis=-1)  # (..., Nt, 2e+1)
    out = np.tensordot(win, g, axes=([-1], [1]))  # (..., Nt, M); g is symmetric
    return np.moveaxis(out, -1, 0)


def build_templates(p, raw, main_extent=None):
    """consts_jax.py:427-447 restricted to the entries the LUT path reads.

    Diffusion bins only address (cx, cy) in [0, nsb//2) (``|s mod nsb - nsb//2 + 0.5|`` floors to
    at most nsb//2 - 1), so templates m >= 1 are materialised for those only; neighbours read
    template 0 = the raw response (``sim_jax.py:222, 251-255``).  Returns float64 arrays:
    T_main (M, c, c, Nt), C_main = cumsum_t T_main, R (Nx, Ny, Nt), C_R = cumsum_t R.
    """
    raw = np.asarray(raw, dtype=np.float64)
    c = main_extent or int(p.nb_sampling_bins_per_pixel) // 2
    g = gaussian_kernels(p)
    T_main = smear(raw[:c, :c], g)
    T_main[0] = raw[:c, :c]
    return dict(T_main=T_main, C_main=np.cumsum(T_main, axis=-1), R=raw, C_R=np.cumsum(raw, axis=-1))


# --------------------------------------------------------------------------------------
# pixel set and signals  (sim_jax.py:141-286, 689-736)
# --------------------------------------------------------------------------------------


def main_pixel_ids(p, tracks, fields, chunk=2_000_000):
    """The main-pixel id of every segment alone (sim_jax.py:394-429: shift, drift's TPC choice,
    bin pitches, pixel2id of the centre bin), chunked, for pixel sets of many-million-segment
    batches where ``segment_quantities`` would materialise every bin.  Pinned by equality with
    ``segment_quantities(...)["main_pid"]``."""
    out = np.empty(len(tracks), np.int64)
    b = _borders(p)
    nsb = int(p.nb_sampling_bins_per_pixel)
    w = p.pixel_pitch / nsb
    m = p.size_margin
    for lo in range(0, len(tracks), chunk):
        t = tracks[lo:lo + chunk]
        x = _col(t, fields, "x", np.float64) - p.shift_x
        y = _col(t, fields, "y", np.float64) - p.shift_y
        z = _col(t, fields, "z", np.float64) - p.shift_z
        ev = np.asarray(t[:, fields.index("eventID")]).astype(np.int64)
        zmin = np.minimum(b[:, 2, 1] - m, b[:, 2, 0] - m)
        zmax = np.maximum(b[:, 2, 1] + m, b[:, 2, 0] + m)
        cond = ((x[:, None] >= b[None, :, 0, 0] - m) & (x[:, None] <= b[None, :, 0, 1] + m)
                & (y[:, None] >= b[None, :, 1, 0] - m) & (y[:, None] <= b[None, :, 1, 1] + m)
                & (z[:, None] >= zmin[None]) & (z[:, None] <= zmax[None]))
        plane = cond.astype(np.int64).argmax(axis=1)
        bx = np.floor_divide(x - b[plane, 0, 0], w).astype(np.int64)
        by = np.floor_divide(y - b[plane, 1, 0], w).astype(np.int64)
        out[lo:lo + chunk] = pixel2id(p, np.floor_divide(bx, nsb), np.floor_divide(by, nsb), plane, ev)
    return out


def unique_pixels(q):
    """sim_jax.py:716-721 without the recompilation padding: sorted unique main pids with -1."""
    return np.unique(np.append(q["main_pid"].ravel(), -1)).astype(np.int64)


def _rows(U, pids):
    r = np.searchsorted(U, pids, side="left")
    rc = np.minimum(r, len(U) - 1)
    found = (r < len(U)) & (U[rc] == pids)
    return rc, found


def simulate_signals(p, q, tables, U, rows_subset=None, chunk=400_000, dtype=np.float64):
    """sim_jax.py:141-286 -- materialise window values + cumsum corrections, then segment_sum.

    Returns wfs of shape (len(U), Nticks) (column 0 is the garbage column).  With
    ``rows_subset`` only contributions landing on those rows are materialised (the sum over the
    other rows is skipped; the returned array still has len(U) rows, zero elsewhere).
    """
    nticks = int(p.time_interval[1] / p.t_sampling) + 1
    L = int(p.signal_length)
    T_main, C_main, R, C_R = (np.asarray(tables[k], dtype=dtype) for k in ("T_main", "C_main", "R", "C_R"))
    nt = R.shape[-1]
    N = len(q["n_e"])
    tau = q["tau"]
    k0 = np.clip(np.floor(tau).astype(np.int64), 0, nt - 1)
    f = (tau - k0).astype(dtype)

    # stripes: main diffusion bins (dropped if their pixel is not a row) ...
    rb, fb = _rows(U, q["pid"].reshape(N, -1))
    seg_b = np.repeat(np.arange(N), rb.shape[1]).reshape(N, -1)
    keep_b = fb
    # ... and neighbours (garbage row 0 if absent)
    rn, fn = _rows(U, q["pid_nb"])
    rn = np.where(fn, rn, 0)
    seg_n = np.repeat(np.arange(N), rn.shape[1]).reshape(N, -1)
    keep_n = np.ones_like(fn)
    if rows_subset is not None:
        sub = np.zeros(len(U), dtype=bool)
        sub[np.asarray(rows_subset)] = True
        keep_b &= sub[rb]
        keep_n &= sub[rn]

    wfs = np.zeros(len(U) * nticks, dtype=dtype)
    j = np.arange(L)
    tl = np.arange(nt - L, nt)

    def _acc(seg, row, vals_win, delta):
        start0 = nt - L - k0[seg]
        for s_off, wf in ((0, 1.0 - f[seg]), (1, f[seg])):
            tt = (start0 - s_off)[:, None] + j[None]
            tt = np.where((tt <= 0) | (tt >= nticks - 1), 0, tt + 1)
            idx = row[:, None] * nticks + tt
            wfs[:] += np.bincount(idx.ravel(), weights=(vals_win * wf[:, None]).ravel(), minlength=wfs.size)
            st = start0 - s_off
            cidx = np.where((st <= 0) | (st >= nticks - 1), 0, st) + row * nticks
            wfs[:] += np.bincount(cidx, weights=delta * wf, minlength=wfs.size)

    # main bins
    sb, cb = seg_b[keep_b], np.nonzero(keep_b.ravel())[0]
    rows_b = rb[keep_b]
    cxb = q["cx"].reshape(N, -1)[keep_b]
    cyb = q["cy"].reshape(N, -1)[keep_b]
    nbin = q["n_bin"].reshape(N, -1)[keep_b].astype(dtype)
    for lo in range(0, len(sb), max(1, chunk // L)):
        sl = slice(lo, lo + max(1, chunk // L))
        s_, r_, cx_, cy_, n_ = sb[sl], rows_b[sl], cxb[sl], cyb[sl], nbin[sl]
        i_ = q["idx"][s_]
        a_, b_, c_ = (q[k][s_].astype(dtype) for k in ("la", "lb", "lc"))
        win = (T_main[i_, cx_, cy_][:, tl] * b_[:, None] + T_main[i_ - 1, cx_, cy_][:, tl] * a_[:, None]
               + T_main[i_ + 1, cx_, cy_][:, tl] * c_[:, None]) * n_[:, None]
        C = C_main[i_, cx_, cy_]
        kk = k0[s_]
        interp = C[np.arange(len(s_)), kk] * (1.0 - f[s_]) + C[np.arange(len(s_)), np.minimum(kk + 1, nt - 1)] * f[s_]
        delta = (C[:, nt - L] - interp) * n_
        _acc(s_, r_, win, delta)

    # neighbours (template 0, undiffused charge)
    sn, rows_n = seg_n[keep_n], rn[keep_n]
    cxn, cyn = q["cx_nb"][keep_n], q["cy_nb"][keep_n]
    for lo in range(0, len(sn), max(1, chunk // L)):
        sl = slice(lo, lo + max(1, chunk // L))
        s_, r_, cx_, cy_ = sn[sl], rows_n[sl], cxn[sl], cyn[sl]
        n_ = q["n_e"][s_].astype(dtype)
        win = R[cx_, cy_][:, tl] * n_[:, None]
        C = C_R[cx_, cy_]
        kk = k0[s_]
        interp = C[np.arange(len(s_)), kk] * (1.0 - f[s_]) + C[np.arange(len(s_)), np.minimum(kk + 1, nt - 1)] * f[s_]
        delta = (C[:, nt - L] - interp) * n_
        _acc(s_, r_, win, delta)
    return wfs.reshape(len(U), nticks)


def simulate_wfs(p, raw_lut, tracks, fields, rows_subset_pids=None, tables=None, dtype=np.float64):
    """sim_jax.py:689-736 -- returns (wfs[:, 1:], unique_pixels) plus the per-segment dict.

    ``rows_subset_pids`` restricts the accumulation to those pixel ids (their rows are returned
    exact; other rows are zero).
    """
    q = segment_quantities(p, tracks, fields, dtype=np.float64)
    U = unique_pixels(q)
    if tables is None:
        tables = build_templates(p, raw_lut)
    sub = None
    if rows_subset_pids is not None:
        r, found = _rows(U, np.asarray(rows_subset_pids, np.int64))
        sub = r[found]
    wfs = simulate_signals(p, q, tables, U, rows_subset=sub, dtype=dtype)
    return wfs[:, 1:], U, q


# --------------------------------------------------------------------------------------
# front-end electronics  (fee_jax.py:55-71, 168-279) and readout (sim_jax.py:619-647, 738-769)
# --------------------------------------------------------------------------------------


def get_adc_values(p, wfs):
    """fee_jax.py:168-279 with noise off: ten sequential self-trigger passes per pixel.

    Returns (adc_charge (P, H), ticks (P, H), decisions) where decisions holds the raw
    crossing index i*, the integration end e, the reset index g and the validity per pass.
    """
    P, T = wfs.shape
    H = int(p.MAX_ADC_VALUES)
    thr = p.DISCRIMINATION_THRESHOLD
    interval = round((3 * p.CLOCK_CYCLE + p.ADC_HOLD_DELAY * p.CLOCK_CYCLE) / p.t_sampling)
    Q = np.cumsum(wfs * p.t_sampling, axis=-1)
    rows = np.arange(P)
    adcs = np.zeros((P, H))
    ticks = np.zeros((P, H))
    dec = dict(istar=np.zeros((P, H), np.int64), e=np.zeros((P, H), np.int64),
               g=np.zeros((P, H), np.int64), valid=np.zeros((P, H), bool))
    for k in range(H):
        cross = (Q[:, 1:] >= thr) & (Q[:, :-1] <= thr)
        istar = np.where(cross.any(axis=1), cross.argmax(axis=1), T - 2)
        e = istar + 1 + interval
        e = np.where(e >= T, T - 1, e)
        a = Q[rows, e]
        invalid = (a < thr) | (istar == T - 2)
        adcs[:, k] = np.where(invalid, 0, a)
        ticks[:, k] = np.where(invalid, T - 2, istar)
        g = istar + 1 + interval + 1
        g = np.where(g >= T, T - 1, g)
        Q = Q - Q[rows, g][:, None]
        Q = np.where(Q < 0, 0, Q)
        dec["istar"][:, k], dec["e"][:, k], dec["g"][:, k], dec["valid"][:, k] = istar, e, g, ~invalid
    return adcs, ticks, dec


def digitize(p, q):
    """fee_jax.py:55-71 -- ADC counts (continuous, clamped, not rounded)."""
    return np.minimum(np.maximum(q * p.GAIN + p.V_PEDESTAL - p.V_CM, 0) * p.ADC_COUNTS / (p.V_REF - p.V_CM), p.ADC_COUNTS)


def adc2charge(dw, p):
    """losses_jax.py:380-383 -- ADC counts to charge in ke-."""
    return (dw / p.ADC_COUNTS * (p.V_REF - p.V_CM) + p.V_CM - p.V_PEDESTAL) / p.GAIN * 1e-3


def parse_output(p, adcs, pixel_x, pixel_y, pixel_z, ticks, hit_prob, event, upix):
    """sim_jax.py:619-647 -- compact the (pixel, hit) grid in (row, k) order."""
    H = adcs.shape[1]
    mask = (hit_prob > p.hit_prob_threshold) & (event[:, None] >= 0) & (upix[:, None] >= 0)
    m = mask.ravel()
    rep = lambda a: np.repeat(a, H)[m]  # noqa: E731
    return dict(adcs=adcs.ravel()[m], pixel_x=rep(pixel_x), pixel_y=rep(pixel_y), pixel_z=pixel_z.ravel()[m],
                ticks=ticks.ravel()[m], hit_prob=hit_prob.ravel()[m], event=rep(event), hit_pixels=rep(upix))


def simulate_stochastic(p, wfs, upix):
    """sim_jax.py:738-769 (noise off) -- FEE, digitisation, coordinates, compaction."""
    integral, ticks, dec = get_adc_values(p, wfs)
    hit_prob = np.where(ticks < wfs.shape[1] - 3, 1.0, 0.0)
    adcs = digitize(p, integral)
    px, py, plane, event = id2pixel(p, upix)
    X, Y = get_pixel_coordinates(p, px, py, plane)
    Z = get_hit_z(p, ticks.ravel(), np.repeat(plane, adcs.shape[1])).reshape(ticks.shape)
    out = parse_output(p, adcs, X, Y, Z, ticks, hit_prob, event, upix)
    out["grid"] = dict(adcs=adcs, integral=integral, ticks=ticks, z=Z, hit_prob=hit_prob, dec=dec)
    return out


def simulate(p, raw_lut, tracks, fields, **kw):
    """LUTSimulation.predict (strategies.py:91-105) on the oracle."""
    wfs, U, q = simulate_wfs(p, raw_lut, tracks, fields, **kw)
    out = simulate_stochastic(p, wfs, U)
    out.update(wfs=wfs, unique_pixels=U, q=q)
    return out


# --------------------------------------------------------------------------------------
# losses  (losses_jax.py:13-80)
# --------------------------------------------------------------------------------------


def mse_loss(adcs, pids, adcs_ref, pids_ref):
    """losses_jax.py:41-56 -- per-pixel squared difference of the (P, H) hit arrays."""
    allp = np.unique(np.concatenate([pids, pids_ref]))
    sig = np.zeros((len(allp), adcs.shape[1]))
    np.add.at(sig, np.searchsorted(allp, pids), adcs)
    np.add.at(sig, np.searchsorted(allp, pids_ref), -adcs_ref)
    return float(np.sum(sig ** 2))


# --------------------------------------------------------------------------------------
# independent scalar loop oracle for tiny events (same specification, no vectorisation)
# --------------------------------------------------------------------------------------


def simulate_wfs_loop(p, raw_lut, tracks, fields):
    """Scalar re-derivation of simulate_wfs for a handful of segments (cross-checks the
    vectorised oracle's indexing).  Returns (wfs_out (P, Nticks-1), U) with the same row set."""
    b = _borders(p)
    ntpc = b.shape[0]
    nx, ny = int(p.n_pixels_x), int(p.n_pixels_y)
    nsb = int(p.nb_sampling_bins_per_pixel)
    nb = int(p.nb_tran_diff_bins)
    sym = (nb - 1) // 2
    nn = int(p.number_pix_neighbors)
    L = int(p.signal_length)
    raw = np.asarray(raw_lut, np.float64)
    nt = raw.shape[-1]
    nticks = int(p.time_interval[1] / p.t_sampling) + 1
    s = _sigmas(p)
    ext = int(p.long_diff_extent)
    v = vdrift(p)
    w = p.pixel_pitch / nsb

    def pid_of(px, py, plane, ev):
        if px < 0 or py < 0 or px >= nx or py >= ny:
            return -1
        return ((ev * ntpc + plane) * ny + py) * nx + px

    def gker(m):
        k = np.arange(-ext, ext + 1)
        g = np.array([math.exp(-0.5 * (kk / s[m]) ** 2) for kk in k])
        return g / g.sum()

    def template(m, cx, cy):
        if m == 0:
            return raw[cx, cy]
        g = gker(m)
        out = np.zeros(nt)
        for t in range(nt):
            acc = 0.0
            for d in range(-ext, ext + 1):
                if 0 <= t + d < nt:
                    acc += raw[cx, cy, t + d] * g[ext + d]
            out[t] = acc
        return out

    tcache = {}

    def T(m, cx, cy):
        if (m, cx, cy) not in tcache:
            tcache[(m, cx, cy)] = template(m, cx, cy)
        return tcache[(m, cx, cy)]

    segs = []
    for r in range(tracks.shape[0]):
        g_ = lambda n: float(tracks[r, fields.index(n)])  # noqa: E731
        x, y, z = g_("x") - p.shift_x, g_("y") - p.shift_y, g_("z") - p.shift_z
        ev = int(g_("eventID"))
        R = p.Ab / (1 + p.kb * g_("dEdx") / (p.eField * p.lArDensity))
        ne = R * g_("dE") * p.MeVToElectrons
        plane, inside = 0, False
        for t in range(ntpc):
            m = p.size_margin
            if (b[t, 0, 0] - m <= x <= b[t, 0, 1] + m and b[t, 1, 0] - m <= y <= b[t, 1, 1] + m
                    and min(b[t, 2, 0], b[t, 2, 1]) - m <= z <= max(b[t, 2, 0], b[t, 2, 1]) + m):
                plane, inside = t, True
                break
        td = (abs(z - b[plane, 2, 0]) + 1e-6) / v
        ne = ne * math.exp(-td / p.lifetime) * (1.0 if inside else 0.0)
        sl, st = math.sqrt(td * 2 * p.long_diff), math.sqrt(td * 2 * p.tran_diff)
        bx = math.floor((x - b[plane, 0, 0]) / w)
        by = math.floor((y - b[plane, 1, 0]) / w)
        x0 = (x - b[plane, 0, 0]) - bx * w
        y0 = (y - b[plane, 1, 0]) - by * w
        edges = [(k - sym) * w for k in range(nb + 1)]

        def wts(c0):
            E = [-1.0] + [math.erf((e - c0) / (math.sqrt(2) * st)) for e in edges[1:-1]] + [1.0]
            return [0.5 * (E[k + 1] - E[k]) for k in range(nb)]

        segs.append(dict(ev=ev, plane=plane, ne=ne, bx=bx, by=by, wx=wts(x0), wy=wts(y0),
                         t0=abs(z - b[plane, 2, 1]) / v, lam=sl / v / p.t_sampling))
    mains = sorted({pid_of(sg["bx"] // nsb, sg["by"] // nsb, sg["plane"], sg["ev"]) for sg in segs} | {-1})
    rowof = {pid: i for i, pid in enumerate(mains)}
    wfs = np.zeros((len(mains), nticks))

    def add_stripe(row, vals, delta, k0, f):
        start0 = nt - L - k0
        for sh, wf in ((0, 1 - f), (1, f)):
            for jj in range(L):
                t = start0 - sh + jj
                col = 0 if (t <= 0 or t >= nticks - 1) else t + 1
                wfs[row, col] += vals[jj] * wf
            st_ = start0 - sh
            col = 0 if (st_ <= 0 or st_ >= nticks - 1) else st_
            wfs[row, col] += delta * wf

    for sg in segs:
        tau = sg["t0"] / p.t_sampling
        k0 = min(max(math.floor(tau), 0), nt - 1)
        f = tau - k0
        lam = sg["lam"]
        idx = min(max(int(np.searchsorted(s, lam, side="left")), 1), len(s) - 2)
        s0, s1, s2 = s[idx - 1], s[idx], s[idx + 1]
        la = (lam - s1) * (lam - s2) / ((s0 - s1) * (s0 - s2))
        lb = (lam - s0) * (lam - s2) / ((s1 - s0) * (s1 - s2))
        lc = (lam - s0) * (lam - s1) / ((s2 - s0) * (s2 - s1))
        for i in range(nb):
            for jb in range(nb):
                sx, sy = sg["bx"] + i - sym, sg["by"] + jb - sym
                pid = pid_of(sx // nsb, sy // nsb, sg["plane"], sg["ev"])
                if pid not in rowof:
                    continue
                cx, cy = int(abs(sx % nsb - nsb // 2 + 0.5)), int(abs(sy % nsb - nsb // 2 + 0.5))
                n = sg["wx"][i] * sg["wy"][jb] * sg["ne"]
                Tm = la * T(idx - 1, cx, cy) + lb * T(idx, cx, cy) + lc * T(idx + 1, cx, cy)
                C = np.cumsum(T(idx, cx, cy))
                delta = n * (C[nt - L] - ((1 - f) * C[k0] + f * C[min(k0 + 1, nt - 1)]))
                add_stripe(rowof[pid], n * Tm[nt - L:], delta, k0, f)
        mpx, mpy = sg["bx"] // nsb, sg["by"] // nsb
        for a in range(-nn, nn + 1):
            for bb in range(-nn, nn + 1):
                # the centre entry carries id -999 (sim_jax.py:450): never a row -> garbage row 0
                pid = -999 if (a == 0 and bb == 0) else pid_of(mpx + a, mpy + bb, sg["plane"], sg["ev"])
                row = rowof.get(pid, 0)
                cx = int(abs(sg["bx"] % nsb - nsb // 2 + 0.5 - a * nsb))
                cy = int(abs(sg["by"] % nsb - nsb // 2 + 0.5 - bb * nsb))
                C = np.cumsum(raw[cx, cy])
                delta = sg["ne"] * (C[nt - L] - ((1 - f) * C[k0] + f * C[min(k0 + 1, nt - 1)]))
                add_stripe(row, sg["ne"] * raw[cx, cy, nt - L:], delta, k0, f)
    return wfs[:, 1:], np.array(mains, dtype=np.int64)
