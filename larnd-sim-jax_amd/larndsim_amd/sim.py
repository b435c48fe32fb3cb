"""Host side of the LUT simulation: contexts, workspaces and the reference-compatible entry points.

Mirrors ``src/larndsim/sim_jax.py`` of larnd-sim-jax:

* ``simulate_wfs(params, response_template, tracks, fields)`` -> ``(wfs, unique_pixels)``
  (``sim_jax.py:689-736``);
* ``simulate_stochastic(params, wfs, unique_pixels, rngseed)`` -> the 8 hit arrays
  (``sim_jax.py:738-769``);
* ``pad_size`` (``sim_jax.py:61-102``) with the same history semantics.

All compute runs in ``liblsim.so`` (HIP, gfx950); torch tensors are only the device buffers.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import logging
from typing import Any

import numpy as np
import torch

from . import _lib
from .consts import RecombinationMode, template_sigmas, vdrift_of

logger = logging.getLogger(__name__)

size_history_dict: dict = {}


def pad_size(cur_size, tag, pad_threshold=0.05):
    """Padded size with a per-tag history (``sim_jax.py:61-102``)."""
    if isinstance(cur_size, (int, np.integer)):
        cur = (int(cur_size),)
        ret = lambda x: x[0]  # noqa: E731
    else:
        cur = tuple(cur_size)
        ret = lambda x: x  # noqa: E731
    hist = size_history_dict.setdefault(tag, [])
    if cur in hist:
        return ret(cur)
    for s in hist:
        if all(cs <= v <= cs * (1 + pad_threshold) for cs, v in zip(cur, s)):
            return ret(s)
    new = tuple(int(cs * (1 + pad_threshold / 2) + 0.5) for cs in cur)
    hist.append(new)
    hist.sort()
    return ret(new)


def get_size_history():
    return size_history_dict


# ------------------------------------------------------------------------------------------
# parameter conversion
# ------------------------------------------------------------------------------------------

def _f(v) -> float:
    return float(v.detach().cpu().item()) if isinstance(v, torch.Tensor) else float(v)


def dvdrift_of(eField, temperature, mob):
    """Analytic d vdrift / d eField of the mobility parameterisation (consts_jax.py:208-216)."""
    a0, a1, a2, a3, a4, a5 = mob
    E = eField
    num = a0 + a1 * E + a2 * E ** 1.5 + a3 * E ** 2.5
    den = 1 + (a1 / a0) * E + a4 * E ** 2 + a5 * E ** 3
    dnum = a1 + 1.5 * a2 * E ** 0.5 + 2.5 * a3 * E ** 1.5
    dden = (a1 / a0) + 2 * a4 * E + 3 * a5 * E ** 2
    tc = (temperature / 89) ** -1.5 / 1000
    return tc * ((dnum * den - num * dden) / den ** 2 * E + num / den)


def params_struct(p) -> _lib.ParamsC:
    mode = p.recombination_mode
    mode = mode.value if isinstance(mode, RecombinationMode) else int(mode)
    E = _f(p.eField)
    s = _lib.ParamsC()
    for name in ("eField", "Ab", "kb", "lifetime", "vdrift", "long_diff", "tran_diff", "lArDensity", "alpha", "beta",
                 "R_param", "MeVToElectrons", "shift_x", "shift_y", "shift_z", "size_margin", "t_sampling"):
        setattr(s, name, _f(getattr(p, name)))
    s.vdrift_eff = vdrift_of(E, float(p.temperature), p.ELECTRON_MOBILITY_PARAMS)
    s.dvdrift_deField = dvdrift_of(E, float(p.temperature), p.ELECTRON_MOBILITY_PARAMS)
    s.threshold = float(p.DISCRIMINATION_THRESHOLD)
    s.gain = float(p.GAIN)
    s.v_pedestal = float(p.V_PEDESTAL)
    s.v_cm = float(p.V_CM)
    s.v_ref = float(p.V_REF)
    s.adc_counts = float(p.ADC_COUNTS)
    s.recombination_mode = mode
    return s


def _check_noise(p):
    if float(p.RESET_NOISE_CHARGE) != 0.0 or float(p.UNCORRELATED_NOISE_CHARGE) != 0.0:
        raise NotImplementedError(
            "readout noise is not implemented: set RESET_NOISE_CHARGE = UNCORRELATED_NOISE_CHARGE = 0 "
            "(the reference's fits remove noise, fit_params.py:118-120)")


def _static_key(p):
    b = np.asarray(p.tpc_borders, dtype=np.float32)
    return (b.tobytes(), float(p.pixel_pitch), int(p.n_pixels_x), int(p.n_pixels_y), int(p.number_pix_neighbors),
            int(p.nb_tran_diff_bins), int(p.nb_sampling_bins_per_pixel), int(p.signal_length), p.n_ticks,
            int(p.MAX_ADC_VALUES), p.adc_interval, template_sigmas(p).tobytes(), int(p.long_diff_extent))


# ------------------------------------------------------------------------------------------
# simulator: one context per (geometry, configuration, LUT)
# ------------------------------------------------------------------------------------------

@dataclasses.dataclass
class ForwardState:
    """Device buffers of one forward call (kept alive for the backward)."""
    n: int
    n_events: int
    tracks: torch.Tensor
    seg: Any
    ws: torch.Tensor
    max_rows: int
    out: dict
    out_c: Any
    params: Any


class Simulator:
    """Owns a ``lsim_ctx`` with the LUT tables, and runs the forward / backward chain.  With
    ``response=None`` the context has no LUT: only the readout (``fee``) and the parametrized
    current model (``larndsim_amd.param``) run on it."""

    def __init__(self, params, response, device="cuda"):
        if not torch.cuda.is_available():
            raise _lib.LsimError("a ROCm GPU is required (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.H = int(params.MAX_ADC_VALUES)
        self.nout = params.n_ticks - 1
        geo = _lib.Geometry()
        b = np.asarray(params.tpc_borders, dtype=np.float32).astype(np.float64)
        if b.shape[0] > _lib.MAX_TPC:
            raise ValueError("too many TPCs")
        geo.n_tpc = b.shape[0]
        geo.n_pixels_x = int(params.n_pixels_x)
        geo.n_pixels_y = int(params.n_pixels_y)
        geo.pixel_pitch = float(params.pixel_pitch)
        for t in range(b.shape[0]):
            for a in range(3):
                for e in range(2):
                    geo.tpc_borders[t][a][e] = float(b[t, a, e])
        cfg = _lib.Config()
        cfg.number_pix_neighbors = int(params.number_pix_neighbors)
        cfg.nb_tran_diff_bins = int(params.nb_tran_diff_bins)
        cfg.nb_sampling_bins_per_pixel = int(params.nb_sampling_bins_per_pixel)
        cfg.signal_length = int(params.signal_length)
        cfg.n_ticks = params.n_ticks
        cfg.max_adc_values = self.H
        cfg.adc_interval = params.adc_interval
        sig = template_sigmas(params).astype(np.float64)
        cfg.n_templates = len(sig)
        cfg.long_diff_extent = int(params.long_diff_extent)
        for i, s in enumerate(sig):
            cfg.long_diff_template[i] = float(s)
        self._ctx = C.c_void_p()
        _lib.check(self.lib.lsim_ctx_create(C.byref(self._ctx), C.byref(geo), C.byref(cfg)))
        self.tape_bytes = int(self.lib.lsim_tape_bytes())
        if response is None:  # no LUT: the readout and the parametrized current model only
            self.lut = None
            return
        lut = torch.as_tensor(np.asarray(response) if not isinstance(response, torch.Tensor) else response)
        lut = lut.to(device=self.device, dtype=torch.float32).contiguous()
        if lut.ndim != 3:
            raise ValueError(f"response LUT must be (Nx, Ny, Nt), got {tuple(lut.shape)}")
        self.lut = lut
        _lib.check(self.lib.lsim_lut_set(self._ctx, C.c_void_p(lut.data_ptr()), lut.shape[0], lut.shape[1], lut.shape[2],
                                         self._stream()))

    def __del__(self):
        try:
            if getattr(self, "_ctx", None) and self._ctx.value:
                torch.cuda.synchronize(self.device)
                self.lib.lsim_ctx_destroy(self._ctx)
                self._ctx = C.c_void_p()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # -------------------------------------------------------------------------------------
    def segments(self, tracks, fields):
        """Upload (N, F) float32 tracks (numpy or torch) and describe its columns in place."""
        t = tracks if isinstance(tracks, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(tracks, dtype=np.float32))
        t = t.to(device=self.device, dtype=torch.float32).contiguous()
        if t.ndim != 2:
            raise ValueError("tracks must be a 2-D (N, F) array")
        n, F = t.shape
        base = t.data_ptr()
        col = lambda name: C.c_void_p(base + 4 * fields.index(name)) if name in fields else None  # noqa: E731
        for req in ("x", "y", "z", "dE", "dEdx", "eventID"):
            if req not in fields:
                raise ValueError(f"tracks are missing field '{req}'")
        s = _lib.Segments(col("x"), col("y"), col("z"), col("dE"), col("dEdx"), col("eventID"), col("z_start"),
                          col("z_end"), col("dx"), F, n)
        return t, s

    def workspace(self, n, n_events):
        sz = C.c_size_t()
        bound = C.c_int64()
        _lib.check(self.lib.lsim_forward_query(self._ctx, n, n_events, C.byref(sz), C.byref(bound)))
        return torch.empty(int(sz.value), dtype=torch.uint8, device=self.device), int(bound.value)

    def alloc_out(self, max_rows, want_wfs=True, want_tape=False, want_hits=True):
        d, H = self.device, self.H
        R = max(int(max_rows), 1)
        o = dict(
            n_rows=torch.zeros(1, dtype=torch.int64, device=d),
            pixel_ids=torch.empty(R, dtype=torch.int64, device=d),
            wfs=torch.empty((R, self.nout), dtype=torch.float32, device=d) if want_wfs else None,
            adc=torch.empty((R, H), dtype=torch.float32, device=d),
            ticks=torch.empty((R, H), dtype=torch.int32, device=d),
            tape=torch.empty(R * self.tape_bytes, dtype=torch.uint8, device=d) if want_tape else None,
            n_hits=torch.zeros(1, dtype=torch.int64, device=d),
            status=torch.zeros(1, dtype=torch.int32, device=d),
        )
        if want_hits:
            for k in ("hit_adc", "hit_x", "hit_y", "hit_z", "hit_tick", "hit_prob"):
                o[k] = torch.empty(R * H, dtype=torch.float32, device=d)
            o["hit_event"] = torch.empty(R * H, dtype=torch.int64, device=d)
            o["hit_pixel"] = torch.empty(R * H, dtype=torch.int64, device=d)
            o["hit_index"] = torch.empty(R * H, dtype=torch.int32, device=d)
        return o

    @staticmethod
    def out_struct(o) -> _lib.FwdOut:
        s = _lib.FwdOut()
        for name, _ in _lib.FwdOut._fields_:
            t = o.get(name)
            setattr(s, name, C.c_void_p(t.data_ptr()) if t is not None else None)
        return s

    def forward(self, params, tracks, fields, n_events=None, want_wfs=True, want_tape=False, max_rows=None):
        """Run the chain.  Without ``max_rows`` the row count is read back after the pixel
        enumeration (one host sync, like ``jnp.unique`` in the reference) and the outputs are
        allocated exactly; with ``max_rows`` everything is stream ordered."""
        _check_noise(params)
        t, seg = self.segments(tracks, fields)
        if n_events is None:
            n_events = infer_n_events(tracks, fields)
        pc = params_struct(params)
        ws, bound = self.workspace(seg.n, n_events)
        st = self._stream()
        if max_rows is None:
            n_rows = torch.zeros(1, dtype=torch.int64, device=self.device)
            status = torch.zeros(1, dtype=torch.int32, device=self.device)
            _lib.check(self.lib.lsim_forward_prepare(self._ctx, C.byref(pc), C.byref(seg), n_events,
                                                     C.c_void_p(ws.data_ptr()), ws.numel(), C.c_void_p(n_rows.data_ptr()),
                                                     C.c_void_p(status.data_ptr()), st))
            P = int(n_rows.item())
            out = self.alloc_out(P, want_wfs, want_tape)
            out["status"] = status
            oc = self.out_struct(out)
            _lib.check(self.lib.lsim_forward_rows(self._ctx, C.byref(pc), seg.n, n_events, C.c_void_p(ws.data_ptr()),
                                                  ws.numel(), P, C.byref(oc), st))
            max_rows = P
        else:
            out = self.alloc_out(max_rows, want_wfs, want_tape)
            oc = self.out_struct(out)
            _lib.check(self.lib.lsim_forward(self._ctx, C.byref(pc), C.byref(seg), n_events, C.c_void_p(ws.data_ptr()),
                                             ws.numel(), max_rows, C.byref(oc), st))
        return ForwardState(seg.n, n_events, t, seg, ws, int(max_rows), out, oc, params)

    def backward(self, state: ForwardState, d_adc, d_z=None, d_wfs=None):
        """d loss / d (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff) as float64 (7,), from
        cotangents on the ADC grid, the hit z grid and (optionally) the waveforms (max_rows, nout)."""
        if state.out.get("tape") is None:
            raise _lib.LsimError("forward was run without want_tape=True")
        sz = C.c_size_t()
        _lib.check(self.lib.lsim_backward_query(self._ctx, state.n, state.max_rows, C.byref(sz)))
        bws = torch.empty(int(sz.value), dtype=torch.uint8, device=self.device)
        d_adc = d_adc.to(device=self.device, dtype=torch.float32).contiguous()
        dz_p = None
        if d_z is not None:
            d_z = d_z.to(device=self.device, dtype=torch.float32).contiguous()
            dz_p = C.c_void_p(d_z.data_ptr())
        grad = torch.zeros(7, dtype=torch.float64, device=self.device)
        pc = params_struct(state.params)
        if d_wfs is not None:
            dw = torch.zeros((state.max_rows, self.nout), dtype=torch.float32, device=self.device)
            dw[: d_wfs.shape[0]] = d_wfs.to(device=self.device, dtype=torch.float32)
            _lib.check(self.lib.lsim_backward_wfs(self._ctx, C.byref(pc), C.byref(state.seg), state.n_events,
                                                  C.c_void_p(state.ws.data_ptr()), state.ws.numel(), state.max_rows,
                                                  C.byref(state.out_c), C.c_void_p(d_adc.data_ptr()), dz_p,
                                                  C.c_void_p(dw.data_ptr()), C.c_void_p(bws.data_ptr()), bws.numel(),
                                                  C.c_void_p(grad.data_ptr()), self._stream()))
            return grad
        _lib.check(self.lib.lsim_backward(self._ctx, C.byref(pc), C.byref(state.seg), state.n_events,
                                          C.c_void_p(state.ws.data_ptr()), state.ws.numel(), state.max_rows,
                                          C.byref(state.out_c), C.c_void_p(d_adc.data_ptr()), dz_p,
                                          C.c_void_p(bws.data_ptr()), bws.numel(), C.c_void_p(grad.data_ptr()),
                                          self._stream()))
        return grad

    def negative_rows(self, params, seg):
        """Negative ids of the reference's pixel set, on the device: ``(ids, count)`` with the
        distinct negative main-pixel ids ascending in ``ids[:count]`` (padding segments' packed
        event -1 ids, and -1 when a main pixel is off the plane).  Stream ordered, no host read."""
        sz, mx = C.c_size_t(), C.c_int64()
        _lib.check(self.lib.lsim_negative_rows_query(self._ctx, C.byref(sz), C.byref(mx)))
        ws = torch.empty(int(sz.value), dtype=torch.uint8, device=self.device)
        ids = torch.empty(max(int(mx.value), 1), dtype=torch.int64, device=self.device)
        n = torch.zeros(1, dtype=torch.int64, device=self.device)
        _lib.check(self.lib.lsim_negative_rows(self._ctx, C.byref(params_struct(params)), C.byref(seg),
                                               C.c_void_p(ws.data_ptr()), ws.numel(), C.c_void_p(ids.data_ptr()),
                                               ids.numel(), C.c_void_p(n.data_ptr()), self._stream()))
        return ids, n

    def fee(self, params, wfs, pixel_ids, want_tape=False):
        """Self-trigger + digitisation + compaction of caller waveforms (simulate_stochastic); with
        ``want_tape`` the decisions the adjoint replays are kept in ``out["tape"]``."""
        _check_noise(params)
        wfs = wfs.to(device=self.device, dtype=torch.float32).contiguous()
        pids = pixel_ids.to(device=self.device, dtype=torch.int64).contiguous()
        P = wfs.shape[0]
        if wfs.shape[1] != self.nout:
            raise ValueError(f"wfs must have {self.nout} ticks, got {wfs.shape[1]}")
        sz = C.c_size_t()
        _lib.check(self.lib.lsim_fee_query(self._ctx, P, C.byref(sz)))
        ws = torch.empty(int(sz.value), dtype=torch.uint8, device=self.device)
        out = self.alloc_out(P, want_wfs=False, want_tape=want_tape)
        oc = self.out_struct(out)
        _lib.check(self.lib.lsim_fee_forward(self._ctx, C.byref(params_struct(params)), C.c_void_p(wfs.data_ptr()),
                                             C.c_void_p(pids.data_ptr()), P, C.c_void_p(ws.data_ptr()), ws.numel(),
                                             C.byref(oc), self._stream()))
        return out


def infer_n_events(tracks, fields) -> int:
    ev = tracks[:, fields.index("eventID")]
    if isinstance(ev, torch.Tensor):
        m = int(ev.max().item()) if ev.numel() else -1
    else:
        m = int(np.max(ev)) if len(ev) else -1
    return max(m + 1, 1)


_SIMS: dict = {}


def get_simulator(params, response) -> Simulator:
    """Cached simulator per (static configuration, LUT object)."""
    key = (_static_key(params), id(response))
    sim = _SIMS.get(key)
    if sim is None or sim.lut_src is not response:
        sim = Simulator(params, response)
        sim.lut_src = response
        _SIMS[key] = sim
    return sim


def hits_dict(out, n_hits=None):
    """Trim the compacted hit arrays to the hit count (one host read)."""
    nh = int(out["n_hits"].item()) if n_hits is None else n_hits
    return dict(adcs=out["hit_adc"][:nh], pixel_x=out["hit_x"][:nh], pixel_y=out["hit_y"][:nh],
                pixel_z=out["hit_z"][:nh], ticks=out["hit_tick"][:nh], hit_prob=out["hit_prob"][:nh],
                event=out["hit_event"][:nh], hit_pixels=out["hit_pixel"][:nh], hit_index=out["hit_index"][:nh])


# ------------------------------------------------------------------------------------------
# reference-compatible entry points
# ------------------------------------------------------------------------------------------

def reference_rows(wfs, pids, neg_ids, n_neg):
    """The reference's row layout of ``simulate_wfs`` (``sim_jax.py:716-721``): the sorted unique
    main-pixel ids plus an appended -1, padded with -1 to ``pad_size(count, "unique_pixels", 0.2)``.

    ``pids`` / ``wfs`` are the P rows with ids >= 0; ``neg_ids[:n_neg]`` the distinct negative main
    pixel ids (``Simulator.negative_rows``).  The negative rows lead, in ascending order, then the
    -1 rows; all of them hold zero waveforms (they carry no charge).  One host read (``n_neg``)."""
    P = pids.shape[0]
    k = int(n_neg.item()) if isinstance(n_neg, torch.Tensor) else int(n_neg)
    tgt = pad_size(P + k + 1, "unique_pixels", 0.2)
    front = tgt - P
    ids = torch.cat([neg_ids[:k], torch.full((front - k,), -1, dtype=torch.int64, device=pids.device), pids])
    wfs = torch.cat([torch.zeros((front, wfs.shape[1]), dtype=wfs.dtype, device=wfs.device), wfs])
    return wfs, ids


def simulate_wfs(params, response_template, tracks, fields, pad=True):
    """Waveforms of the pixels that are the main pixel of some segment (``sim_jax.py:689-736``).

    Returns ``(wfs, unique_pixels)``: ``wfs`` (P_pad, Nticks-1) float32 and ``unique_pixels``
    (P_pad,) int64 sorted ascending.  With ``pad`` the rows are laid out as the reference's are
    (``reference_rows``: negative ids first, then ``pad_size`` -1 rows, zero waveforms); without it
    only the rows with ids >= 0 are returned.
    """
    sim = get_simulator(params, response_template)
    st = sim.forward(params, tracks, fields, want_wfs=True)
    wfs, pids = st.out["wfs"][: st.max_rows], st.out["pixel_ids"][: st.max_rows]
    if pad:
        neg, n_neg = sim.negative_rows(params, st.seg)
        wfs, pids = reference_rows(wfs, pids, neg, n_neg)
    return wfs, pids


def simulate_stochastic(params, wfs, unique_pixels, rngseed=0):
    """Self-trigger readout of waveforms (``sim_jax.py:738-769``, noise off).

    Returns ``(adcs, pixel_x, pixel_y, pixel_z, ticks, hit_prob, event, hit_pixels)`` compacted in
    (row, hit) order like ``parse_output`` (``sim_jax.py:619-647``).
    """
    sim = None
    for s in _SIMS.values():
        if s.nout == wfs.shape[1]:
            sim = s
            break
    if sim is None:
        raise _lib.LsimError("simulate_stochastic needs a simulator; call simulate_wfs first")
    out = sim.fee(params, torch.as_tensor(wfs), torch.as_tensor(unique_pixels))
    h = hits_dict(out)
    return (h["adcs"], h["pixel_x"], h["pixel_y"], h["pixel_z"], h["ticks"], h["hit_prob"], h["event"], h["hit_pixels"])


# ------------------------------------------------------------------------------------------
# preallocated runner: fixed buffers, one stream-ordered library call per step (capturable)
# ------------------------------------------------------------------------------------------

class Runner:
    """Forward (and optionally backward) over fixed device buffers.

    Construction does one two-phase forward to learn the row count (times ``row_margin``) and
    allocates every buffer; ``forward()`` / ``backward()`` then make exactly one library call each
    on the current stream with no allocation and no host synchronisation, so a step can be
    captured into a hipGraph (``torch.cuda.CUDAGraph``) and replayed.
    """

    def __init__(self, sim: Simulator, params, tracks, fields, n_events=None, want_wfs=True, want_tape=False,
                 row_margin=1.0):
        _check_noise(params)
        self.sim, self.lib = sim, sim.lib
        self.fields = fields
        self.tracks, self.seg = sim.segments(tracks, fields)
        self.n_events = infer_n_events(self.tracks, fields) if n_events is None else n_events
        self.set_params(params)
        st = sim.forward(params, self.tracks, fields, n_events=self.n_events, want_wfs=False)
        self.rows = st.max_rows
        self.max_rows = max(int(self.rows * row_margin), 1)
        self.ws = st.ws
        self.out = sim.alloc_out(self.max_rows, want_wfs, want_tape)
        self.oc = sim.out_struct(self.out)
        self.bws = None
        self.grad = torch.zeros(7, dtype=torch.float64, device=sim.device)
        self.d_adc = None
        self.d_z = None

    def set_params(self, params):
        self.params = params
        self.pc = params_struct(params)

    def forward(self):
        _lib.check(self.lib.lsim_forward(self.sim._ctx, C.byref(self.pc), C.byref(self.seg), self.n_events,
                                         C.c_void_p(self.ws.data_ptr()), self.ws.numel(), self.max_rows,
                                         C.byref(self.oc), self.sim._stream()))

    def set_cotangents(self, d_adc, d_z=None):
        n = self.max_rows * self.sim.H
        self.d_adc = torch.zeros(n, dtype=torch.float32, device=self.sim.device)
        self.d_adc[: d_adc.numel()].copy_(d_adc.reshape(-1)[:n])
        if d_z is not None:
            self.d_z = torch.zeros(n, dtype=torch.float32, device=self.sim.device)
            self.d_z[: d_z.numel()].copy_(d_z.reshape(-1)[:n])
        sz = C.c_size_t()
        _lib.check(self.lib.lsim_backward_query(self.sim._ctx, self.seg.n, self.max_rows, C.byref(sz)))
        self.bws = torch.empty(int(sz.value), dtype=torch.uint8, device=self.sim.device)

    def backward(self):
        dz = C.c_void_p(self.d_z.data_ptr()) if self.d_z is not None else None
        _lib.check(self.lib.lsim_backward(self.sim._ctx, C.byref(self.pc), C.byref(self.seg), self.n_events,
                                          C.c_void_p(self.ws.data_ptr()), self.ws.numel(), self.max_rows,
                                          C.byref(self.oc), C.c_void_p(self.d_adc.data_ptr()), dz,
                                          C.c_void_p(self.bws.data_ptr()), self.bws.numel(),
                                          C.c_void_p(self.grad.data_ptr()), self.sim._stream()))


def timing(enable: bool):
    lib = _lib.load()
    _lib.check(lib.lsim_timing(1 if enable else 0))


def timing_read():
    lib = _lib.load()
    ms = C.c_double()
    n = C.c_int64()
    _lib.check(lib.lsim_timing_read(C.byref(ms), C.byref(n)))
    return float(ms.value), int(n.value)


def set_row_waves(w: int):
    _lib.check(_lib.load().lsim_set_row_waves(int(w)))


def set_row_cost(pairs: int):
    _lib.check(_lib.load().lsim_set_row_cost(int(pairs)))


def set_row_tile(cells: int):
    """Side of the square pixel-cell tiles the waveform kernel processes rows in (1: row order)."""
    _lib.check(_lib.load().lsim_set_row_tile(int(cells)))


def _kernel_ms(runner: Runner, reps: int) -> float:
    runner.forward()
    torch.cuda.synchronize()
    timing(True)
    for _ in range(reps):
        runner.forward()
    ms, n = timing_read()
    timing(False)
    return ms / max(n, 1)


def autotune_row_waves(runner: Runner, candidates=(1, 2, 4, 8), reps=3):
    """Pick the waves-per-row-workgroup setting with the shortest waveform-kernel time."""
    best, best_ms = None, None
    res = {}
    for w in candidates:
        set_row_waves(w)
        res[w] = _kernel_ms(runner, reps)
        if best_ms is None or res[w] < best_ms:
            best, best_ms = w, res[w]
    set_row_waves(best)
    return best, res


def autotune_rows(runner: Runner, waves=(1, 2, 4, 8), costs=(64, 256, 1024), tiles=(8, 16, 32), reps=3):
    """Tune the waveform kernel on the runner's workload, one knob at a time: waves per row
    workgroup (at the default row cost and tile), then the per-row cost the XCD bands are balanced
    with, then the side of the cell tiles the rows are processed in (tiled orders only: they time
    within noise of the plain row order and fetch a third less from HBM).  All three only change
    the schedule and the per-wave partial sums; the settings stay in effect.  Returns (waves, row
    cost, tile, {setting: kernel ms})."""
    set_row_cost(256)
    set_row_tile(16)
    w, res_w = autotune_row_waves(runner, waves, reps)

    def pick(setter, values):
        best, best_ms, res = None, None, {}
        for v in values:
            setter(v)
            res[v] = _kernel_ms(runner, reps)
            if best_ms is None or res[v] < best_ms:
                best, best_ms = v, res[v]
        setter(best)
        return best, res

    c, res_c = pick(set_row_cost, costs)
    t, res_t = pick(set_row_tile, tiles)
    return w, c, t, {**{f"waves={k}": v for k, v in res_w.items()}, **{f"row_cost={k}": v for k, v in res_c.items()},
                     **{f"row_tile={k}": v for k, v in res_t.items()}}

