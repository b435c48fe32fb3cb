"""torch-float64 autograd restatement of the parametrized current model -- TEST INFRASTRUCTURE ONLY.

The gradient checker for the hand-written adjoint of the parametrized model (lsim_param_backward).
Only ``tests/`` may import it.

Reverse-mode AD through the reference (``jax.value_and_grad`` over ParametrizedSimulation.predict,
optimize/strategies.py:136-149 -> simulate_parametrized, sim_jax.py:339-372) differentiates the
continuous quantities and holds every integer decision fixed: the pixel ids of the pairs
(detsim_jax.py:483-496), the whole-tick split of the arrival time (``astype(int)``,
detsim_jax.py:629-633), the scatter indices (detsim_jax.py:209-228) and the readout's decisions.
This module takes those decisions from the float64 NumPy oracle (``oracle.param``) and recomputes
every continuous quantity from the 7 parameters in torch float64: quench (quenching_jax.py:23-25,
BIRKS), drift (drifting_jax.py:107-124), the electrons (detsim_jax.py:322-330, 356-399), the
current (detsim_jax.py:441-639), the scatter, and the readout replay shared with the LUT chain's
checker (``ad_oracle.fee_replay``).

Pinning: ``tests/test_param_ad_oracle.py`` checks this module's forward against the NumPy oracle
and its gradients against central finite differences of the NumPy oracle itself (decisions free,
and checked unchanged over the difference interval).
"""
from __future__ import annotations

import numpy as np
import torch

from . import oracle as O
from . import param as PA
from .ad_oracle import GRAD_PARAMS, fee_replay, vdrift_t


def _surface_t(c, x, y):
    return c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * x * x + c[5] * y * y


def _emg_cdf_half(X, mu, sigma, lam):
    """0.5 erf((X - mu) / (sqrt2 sigma)) - emg_pdf(X) / lambda (detsim_jax.py:441-476), torch."""
    coeff = lam / 2
    ex = coeff * (2 * mu + lam * sigma ** 2 - 2 * X)
    pdf = coeff * torch.exp(ex) * torch.special.erfc((mu + lam * sigma ** 2 - X) / (np.sqrt(2) * sigma))
    return 0.5 * torch.erf((X - mu) / (np.sqrt(2) * sigma)) - pdf / lam


def _ticks_diff(x, loc, scale, sigma, dt):
    """integrated_expon_diff (detsim_jax.py:463-476), torch; x (..., T)."""
    lam = 1 / scale
    upper = _emg_cdf_half(x + dt / 2, loc, sigma, lam)
    lower = _emg_cdf_half(x - dt / 2, loc, sigma, lam)
    return (upper - lower) / (lower[..., :1] - upper[..., -1:]) / dt


def _ticks(x, loc, scale, dt):
    """integrated_expon (detsim_jax.py:534-540), torch; x (..., T)."""
    z = torch.zeros((), dtype=torch.float64)
    return (torch.exp(torch.minimum(z, (loc - x + dt / 2) / scale))
            - torch.exp(torch.minimum(z, (loc - x - dt / 2) / scale))
            + torch.exp(torch.minimum(z, (loc - dt / 2) / scale)) / x.shape[-1]) / dt


class FrozenParam:
    """Integer decisions and parameter-independent data of one event of the parametrized model,
    from the NumPy oracle.  ``normals`` (N, 3) for mc_diff, else ``bin_edges`` for the transverse
    bins.  BIRKS recombination."""

    def __init__(self, p, tracks, fields, normals=None, bin_edges=None):
        mode = getattr(p.recombination_mode, "name", p.recombination_mode)
        if mode not in ("BIRKS", 2):
            raise ValueError("the checker restates BIRKS recombination only")
        self.p = p
        with np.errstate(all="ignore"):
            o = PA.simulate_parametrized(p, tracks, fields, normals=normals, bin_edges=bin_edges)
        self.out = o
        q = O.segment_quantities(p, tracks, fields)
        f = lambda n: np.asarray(tracks[:, fields.index(n)], np.float64)  # noqa: E731
        self.x, self.y, self.z = f("x") - p.shift_x, f("y") - p.shift_y, f("z") - p.shift_z
        self.dE, self.dEdx = f("dE"), f("dEdx")
        b = np.asarray(p.tpc_borders, np.float64)
        self.plane = q["plane"]
        self.z_an = b[q["plane"], 2, 0]
        self.mask = q["mask"].astype(np.float64)
        self.mc = bool(p.mc_diff)
        self.normals = None if normals is None else np.asarray(normals, np.float64)
        self.edges = None if bin_edges is None else np.asarray(bin_edges, np.float64)
        e = o["electrons"]
        self.e = e
        self.per = 1 if self.mc else (len(self.edges) - 1) ** 2
        pids = o["pair_pids"]
        self.U = o["unique_pixels"]
        X, Y, plane, _ = O.id2pixel(p, pids)
        cx, cy = O.get_pixel_coordinates(p, X, Y, plane)
        self.cx, self.cy = cx, cy
        self.t0_tick = o["t0_tick"]
        nt = int(p.time_interval[1] / p.t_sampling) + 1
        self.nt = nt
        T = len(PA.current_ticks(p))
        tt = (self.t0_tick - T)[:, None] + np.arange(T)[None]
        tt = np.where((tt < 0) | (tt >= nt - 1), 0, tt + 1)
        rows = np.searchsorted(self.U, pids)
        self.flat = (rows[:, :, None] * nt + tt[:, None, :]).ravel()
        self.dec = o["grid"]["dec"]
        self.ticks = o["grid"]["ticks"]
        self.valid_rows = self.U >= 0

    def theta0(self):
        return torch.tensor([float(getattr(self.p, n)) for n in GRAD_PARAMS], dtype=torch.float64)

    def forward(self, theta):
        """theta: float64 tensor (7,) of (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff).
        Returns (adc grid (P, H), z grid (P, H), wfs (P, n_ticks - 1)) on the oracle's rows."""
        p = self.p
        Ab, kb, E, tau_l, _vd, DL, DT = (theta[i] for i in range(7))
        t = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
        v = vdrift_t(E, p)
        R = Ab / (1 + kb * t(self.dEdx) / (E * p.lArDensity))
        n0 = R * t(self.dE) * p.MeVToElectrons
        td = (torch.abs(t(self.z) - t(self.z_an)) + 1e-6) / v
        ne = n0 * torch.exp(-td / tau_l) * t(self.mask)
        sl = torch.sqrt(td * 2 * DL)
        st = torch.sqrt(td * 2 * DT)
        if self.mc:
            g = t(self.normals)
            ex = t(self.x) + g[:, 0] * st
            ey = t(self.y) + g[:, 1] * st
            ez = t(self.z) + g[:, 2] * sl if not p.diffusion_in_current_sim else t(self.z)
            ene, esl, z_an = ne, sl, t(self.z_an)
        else:
            edges = t(self.edges)
            inner = torch.erf(edges[None, 1:-1] / (np.sqrt(2) * st[:, None]))
            ones = torch.ones((len(st), 1), dtype=torch.float64)
            Ek = torch.cat([-ones, inner, ones], dim=1)
            w = 0.5 * (Ek[:, 1:] - Ek[:, :-1])
            W = (w[:, :, None] * w[:, None, :]).reshape(-1)
            rep = lambda a: torch.repeat_interleave(a, self.per)  # noqa: E731
            ex, ey = t(self.e["x"]), t(self.e["y"])
            ez = rep(t(self.z))
            ene, esl, z_an = rep(ne) * W, rep(sl), rep(t(self.z_an))
        # current_mc (detsim_jax.py:623-639) with the oracle's pixels and whole ticks
        ts = PA.current_ticks(p)
        Tn = len(ts)
        xd = torch.abs(ex[:, None] - t(self.cx))
        yd = torch.abs(ey[:, None] - t(self.cy))
        t0 = torch.abs(ez - z_an) / v
        t0 = t0 - t(self.t0_tick) * p.t_sampling
        dt = 5.0 / (Tn - 1)
        a = torch.minimum(_surface_t(PA.B_PARAMS, xd, yd), torch.ones((), dtype=torch.float64))
        b = _surface_t(PA.C_PARAMS, xd, yd)
        c = _surface_t(PA.D_PARAMS, xd, yd)
        st0 = t0[:, None] + _surface_t(PA.T0_PARAMS, xd, yd)
        x = -t(ts)[None, None, :]
        loc = -st0[..., None]
        if p.diffusion_in_current_sim:
            sig = (esl / v)[:, None, None]
            cur = a[..., None] * _ticks_diff(x, loc, b[..., None], sig, dt) + (1 - a[..., None]) * _ticks_diff(
                x, loc, c[..., None], sig, dt)
        else:
            cur = a[..., None] * _ticks(x, loc, b[..., None], dt) + (1 - a[..., None]) * _ticks(x, loc, c[..., None], dt)
        signals = cur * ene[:, None, None]
        P = len(self.U)
        wfs = torch.zeros(P * self.nt, dtype=torch.float64).index_add(0, torch.as_tensor(self.flat), signals.reshape(-1))
        wfs_out = wfs.reshape(P, self.nt)[:, 1:]
        adc, z = fee_replay(p, self.dec, self.ticks, self.U, wfs_out, v)
        return adc, z, wfs_out

    def vjp(self, w_adc, w_z, w_wfs=None):
        """Gradient of sum(w_adc adc) + sum(w_z z) over the valid rows (+ sum(w_wfs wfs) over all
        rows), w.r.t. the 7 parameters."""
        th = self.theta0().requires_grad_(True)
        adc, z, wfs = self.forward(th)
        vr = torch.as_tensor(self.valid_rows)
        L = (torch.as_tensor(w_adc) * adc[vr]).sum() + (torch.as_tensor(w_z) * z[vr]).sum()
        if w_wfs is not None:
            L = L + (torch.as_tensor(w_wfs) * wfs).sum()
        L.backward()
        return th.grad.detach().numpy(), float(L.detach())
