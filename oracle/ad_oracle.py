"""torch-float64 autograd restatement of the LUT chain -- TEST INFRASTRUCTURE ONLY.

The gradient checker for the hand-written adjoint.  Only ``tests/`` may import it.

Reverse-mode AD through the reference (``jax.value_and_grad`` in optimize/fit_params.py:731)
differentiates the continuous quantities and holds every integer decision fixed (floor /
astype / searchsorted / argmax / jnp.where masks carry no gradient).  This module takes those
decisions from the float64 NumPy oracle (``oracle.oracle``) and recomputes every continuous
quantity from the 7 parameters in torch float64, following the same reference lines:
quench (quenching_jax.py:23-25, 74), drift (drifting_jax.py:107-124), transverse weights
(detsim_jax.py:332-354), time and template interpolation (sim_jax.py:157-168, 419-423), window
values and boundary corrections (sim_jax.py:170-261), accumulation (sim_jax.py:279-286), the
self-trigger replay (fee_jax.py:189-268), digitize (fee_jax.py:68-69) and hit z
(detsim_jax.py:308-319).

Pinning: ``tests/test_ad_oracle.py`` checks this module's forward values against the NumPy oracle
(bitwise-close), and its gradients against central finite differences of the same
frozen-decision function.
"""
from __future__ import annotations

import numpy as np
import torch

from . import oracle as O

GRAD_PARAMS = ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff")


def vdrift_t(E, p):
    a0, a1, a2, a3, a4, a5 = p.ELECTRON_MOBILITY_PARAMS
    num = a0 + a1 * E + a2 * E ** 1.5 + a3 * E ** 2.5
    den = 1 + (a1 / a0) * E + a4 * E ** 2 + a5 * E ** 3
    return num / den * (p.temperature / 89) ** -1.5 / 1000 * E


def fee_replay(p, dec, ticks, U, wfs_out, v):
    """The self-trigger with the oracle's decisions (fee_jax.py:189-268), digitize (fee_jax.py:68-69)
    and hit z (detsim_jax.py:308-319) as torch float64 functions of the waveforms wfs_out (P, T)
    and the drift velocity v: (adc grid (P, H), z grid (P, H))."""
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
    P = wfs_out.shape[0]
    H = int(p.MAX_ADC_VALUES)
    Q = torch.cumsum(wfs_out * p.t_sampling, dim=1)
    rows = torch.arange(P)
    adcs = []
    for k in range(H):
        e = torch.as_tensor(dec["e"][:, k])
        g = torch.as_tensor(dec["g"][:, k])
        valid = torch.as_tensor(dec["valid"][:, k])
        a = torch.where(valid, Q[rows, e], torch.zeros((), dtype=torch.float64))
        adcs.append(a)
        Qn = Q - Q[rows, g][:, None]
        Q = torch.where(Qn < 0, torch.zeros((), dtype=torch.float64), Qn)
    integral = torch.stack(adcs, dim=1)
    x = torch.clamp(integral * p.GAIN + p.V_PEDESTAL - p.V_CM, min=0)
    adc = torch.clamp(x * p.ADC_COUNTS / (p.V_REF - p.V_CM), max=p.ADC_COUNTS)
    bb = np.asarray(p.tpc_borders, np.float64)
    _, _, plane, _ = O.id2pixel(p, U)
    za, zc = t(bb[plane, 2, 0]), t(bb[plane, 2, 1])
    z = za[:, None] + t(ticks) * p.t_sampling * v * torch.sign(zc - za)[:, None]
    return adc, z


class Frozen:
    """Integer decisions and parameter-independent data of one event, from the NumPy oracle."""

    def __init__(self, p, raw_lut, tracks, fields, tables=None, rows_subset_pids=None):
        """With ``rows_subset_pids`` (pixel ids >= 0 that are rows of the event) only those rows
        are built, replayed and returned, in ascending id order: the stripes landing elsewhere are
        dropped, so large events stay cheap.  Their values equal the full chain's on those rows."""
        self.p = p
        self.tables = tables if tables is not None else O.build_templates(p, raw_lut)
        tables = self.tables
        subset = rows_subset_pids is not None
        if not subset:
            sim = O.simulate(p, raw_lut, tracks, fields, tables=tables)
        else:
            pids = np.unique(np.asarray(rows_subset_pids, np.int64))
            wfs, U, q = O.simulate_wfs(p, raw_lut, tracks, fields, rows_subset_pids=pids, tables=tables)
            r, found = O._rows(U, pids)
            if not found.all() or (pids < 0).any():
                raise ValueError("rows_subset_pids must be pixel ids >= 0 of rows of the event")
            sim = O.simulate_stochastic(p, wfs[r], U[r])
            sim.update(wfs=wfs[r], unique_pixels=U[r], q=q)
        self.out = sim
        q = sim["q"]
        self.q = q
        self.U = sim["unique_pixels"]
        f = lambda n: np.asarray(tracks[:, fields.index(n)], np.float64)  # noqa: E731
        self.z = f("z") - p.shift_z
        self.dE, self.dEdx = f("dE"), f("dEdx")
        b = np.asarray(p.tpc_borders, np.float64)
        self.z_an = b[q["plane"], 2, 0]
        self.z_ca = b[q["plane"], 2, 1]
        self.mask = q["mask"].astype(np.float64)
        N = len(self.z)
        nt = self.tables["R"].shape[-1]
        self.k0 = np.clip(np.floor(q["tau"]).astype(np.int64), 0, nt - 1)
        # stripes (sim_jax.py:152-154, 722-725)
        rb, fb = O._rows(self.U, q["pid"].reshape(N, -1))
        rn, fn = O._rows(self.U, q["pid_nb"])
        self.bins = dict(seg=np.nonzero(fb)[0], k=np.nonzero(fb)[1], row=rb[fb],
                         cx=q["cx"].reshape(N, -1)[fb], cy=q["cy"].reshape(N, -1)[fb])
        if subset:
            # neighbour stripes off the subset are dropped
            segn, kn = np.nonzero(fn)
            self.nbs = dict(seg=segn, row=rn[fn], cx=q["cx_nb"][fn], cy=q["cy_nb"][fn])
        else:
            # neighbours without a row land on row 0, the -1 row (sim_jax.py:723-725)
            rn = np.where(fn, rn, 0)
            segn, kn = np.nonzero(np.ones_like(fn))
            self.nbs = dict(seg=segn, row=rn.ravel(), cx=q["cx_nb"].ravel(), cy=q["cy_nb"].ravel())
        self.dec = sim["grid"]["dec"]
        self.ticks = sim["grid"]["ticks"]
        self.valid_rows = self.U >= 0

    def forward(self, theta):
        """theta: float64 tensor (7,) of (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff).
        Returns (adc grid (P, H), z grid (P, H), wfs_out (P, Nticks-1)) on the oracle's rows."""
        p = self.p
        Ab, kb, E, tau_l, _vd, DL, DT = (theta[i] for i in range(7))
        t = lambda a: torch.as_tensor(a, dtype=torch.float64)  # noqa: E731
        q = self.q
        v = vdrift_t(E, p)
        R = Ab / (1 + kb * t(self.dEdx) / (E * p.lArDensity))
        n0 = R * t(self.dE) * p.MeVToElectrons
        td = (torch.abs(t(self.z) - t(self.z_an)) + 1e-6) / v
        ne = n0 * torch.exp(-td / tau_l) * t(self.mask)
        sl = torch.sqrt(td * 2 * DL)
        st = torch.sqrt(td * 2 * DT)
        edges = t(O.bin_edges(p))

        def w1d(x0):
            e = edges[None, :] - t(x0)[:, None]
            inner = torch.erf(e[:, 1:-1] / (np.sqrt(2) * st[:, None]))
            ones = torch.ones((len(x0), 1), dtype=torch.float64)
            Ek = torch.cat([-ones, inner, ones], dim=1)
            return 0.5 * (Ek[:, 1:] - Ek[:, :-1])

        W = w1d(q["x0"])[:, :, None] * w1d(q["y0"])[:, None, :]
        nbin = (W * ne[:, None, None]).reshape(len(ne), -1)
        t0 = torch.abs(t(self.z) - t(self.z_ca)) / v
        tau = t0 / p.t_sampling
        k0 = t(self.k0)
        f = tau - k0
        lam = sl / v / p.t_sampling
        s = torch.as_tensor(O._sigmas(p))
        idx = torch.as_tensor(q["idx"])
        s0, s1, s2 = s[idx - 1], s[idx], s[idx + 1]
        la = (lam - s1) * (lam - s2) / ((s0 - s1) * (s0 - s2))
        lb = (lam - s0) * (lam - s2) / ((s1 - s0) * (s1 - s2))
        lc = (lam - s0) * (lam - s1) / ((s2 - s0) * (s2 - s1))

        T_main, C_main = torch.as_tensor(self.tables["T_main"]), torch.as_tensor(self.tables["C_main"])
        Rr, CR = torch.as_tensor(self.tables["R"]), torch.as_tensor(self.tables["C_R"])
        nticks = int(p.time_interval[1] / p.t_sampling) + 1
        L = int(p.signal_length)
        nt = Rr.shape[-1]
        P = len(self.U)
        wfs = torch.zeros(P * nticks, dtype=torch.float64)
        j = np.arange(L)

        def acc(seg, row, win, delta):
            kk = self.k0[seg]
            start0 = nt - L - kk
            fs = f[seg]
            for so, wf in ((0, 1.0 - fs), (1, fs)):
                tt = (start0 - so)[:, None] + j[None]
                tt = np.where((tt <= 0) | (tt >= nticks - 1), 0, tt + 1)
                ix = torch.as_tensor((row[:, None] * nticks + tt).ravel())
                wfs.index_add_(0, ix, (win * wf[:, None]).reshape(-1))
                st_ = start0 - so
                cix = torch.as_tensor(np.where((st_ <= 0) | (st_ >= nticks - 1), 0, st_) + row * nticks)
                wfs.index_add_(0, cix, delta * wf)

        b = self.bins
        sg = b["seg"]
        ii = idx[sg]
        cx, cy = torch.as_tensor(b["cx"]), torch.as_tensor(b["cy"])
        win = (T_main[ii, cx, cy][:, nt - L:] * lb[sg][:, None] + T_main[ii - 1, cx, cy][:, nt - L:] * la[sg][:, None]
               + T_main[ii + 1, cx, cy][:, nt - L:] * lc[sg][:, None]) * nbin[sg, b["k"]][:, None]
        C = C_main[ii, cx, cy]
        kk = torch.as_tensor(self.k0[sg])
        ar = torch.arange(len(sg))
        interp = C[ar, kk] * (1.0 - f[sg]) + C[ar, torch.clamp(kk + 1, max=nt - 1)] * f[sg]
        delta = (C[:, nt - L] - interp) * nbin[sg, b["k"]]
        acc(sg, b["row"], win, delta)

        nb = self.nbs
        sg = nb["seg"]
        cx, cy = torch.as_tensor(nb["cx"]), torch.as_tensor(nb["cy"])
        win = Rr[cx, cy][:, nt - L:] * ne[sg][:, None]
        C = CR[cx, cy]
        kk = torch.as_tensor(self.k0[sg])
        ar = torch.arange(len(sg))
        interp = C[ar, kk] * (1.0 - f[sg]) + C[ar, torch.clamp(kk + 1, max=nt - 1)] * f[sg]
        delta = (C[:, nt - L] - interp) * ne[sg]
        acc(sg, nb["row"], win, delta)

        wfs_out = wfs.reshape(P, nticks)[:, 1:]
        adc, z = fee_replay(p, self.dec, self.ticks, self.U, wfs_out, v)
        return adc, z, wfs_out

    def theta0(self):
        return torch.tensor([float(getattr(self.p, n)) for n in GRAD_PARAMS], dtype=torch.float64)

    def vjp(self, w_adc, w_z):
        """Gradient of sum(w_adc * adc) + sum(w_z * z) over the valid rows, w.r.t. the 7 params."""
        th = self.theta0().requires_grad_(True)
        adc, z, _ = self.forward(th)
        vr = torch.as_tensor(self.valid_rows)
        L = (torch.as_tensor(w_adc) * adc[vr]).sum() + (torch.as_tensor(w_z) * z[vr]).sum()
        L.backward()
        return th.grad.detach().numpy(), float(L.detach())

    def value(self, theta_np, w_adc, w_z):
        with torch.no_grad():
            adc, z, _ = self.forward(torch.as_tensor(theta_np, dtype=torch.float64))
            vr = torch.as_tensor(self.valid_rows)
            return float((torch.as_tensor(w_adc) * adc[vr]).sum() + (torch.as_tensor(w_z) * z[vr]).sum())
