"""Track chopping and batch assembly on the GPU (optimize/dataio.py).

* ``chop_tracks(tracks, fields, precision)``: optimize/dataio.py:63-106, on the GPU through
  ``lsim_chop_count`` / ``lsim_chop_fill``, bit-identical to the reference's NumPy result.
* ``chop_batch``: one batch of ``TracksDataset.__getitem__`` (:378-400): the event-id remap of
  ``remap_event_ids_to_local`` (:47-61), the chop, and the padding / invalidation of
  ``pad_batch`` (:357-373) fused into the same two launches.
* ``group_events_into_batches``: the event-grouped batch split of ``TracksDataset.__init__``
  (:245-289), host bookkeeping over per-trajectory lengths.

The returned arrays are float32 device tensors with the reference's row layout (one row per step,
columns in ``fields`` order), so they feed ``lsim_segments`` directly with stride ``len(fields)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

ZERO_PAD_FIELDS = ("n_electrons", "dE", "dEdx", "dx", "long_diff", "tran_diff")  # dataio.py:340
NEG_ONE_PAD_FIELDS = ("trackID", "pixel_plane")  # dataio.py:341


def chop_layout(fields) -> _lib.ChopLayout:
    """Column indices of ``fields`` (the reference's ``fields.index``) for the chop kernels."""
    f = list(fields)
    lay = _lib.ChopLayout()
    lay.n_fields = len(f)
    for a, c in enumerate("xyz"):
        lay.start[a] = f.index(c + "_start")
        lay.end[a] = f.index(c + "_end")
        lay.mid[a] = f.index(c)
    lay.dE, lay.dx = f.index("dE"), f.index("dx")
    lay.event = f.index("eventID") if "eventID" in f else -1
    zero = [f.index(n) for n in ZERO_PAD_FIELDS if n in f]
    neg = [f.index(n) for n in NEG_ONE_PAD_FIELDS if n in f]
    lay.n_zero, lay.n_neg_one = len(zero), len(neg)
    for i, c in enumerate(zero):
        lay.zero[i] = c
    for i, c in enumerate(neg):
        lay.neg_one[i] = c
    return lay


def _device_tracks(tracks, fields, device):
    if isinstance(tracks, torch.Tensor):
        t = tracks.to(device=device, dtype=torch.float32)
    else:
        t = torch.from_numpy(np.ascontiguousarray(tracks, dtype=np.float32)).to(device)
    if t.dim() != 2 or t.shape[1] != len(fields):
        raise ValueError(f"tracks must be (N, {len(fields)}), got {tuple(t.shape)}")
    return t.contiguous()


class _Chopper:
    """Count, scan and fill for one track array (one host read of the step total)."""

    def __init__(self, tracks, fields, precision, device):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else (
            tracks.device if isinstance(tracks, torch.Tensor) and tracks.is_cuda else torch.device("cuda"))
        if self.device.type != "cuda":
            raise _lib.LsimError("the chop kernels run on a GPU device")
        self.fields = tuple(fields)
        self.t = _device_tracks(tracks, self.fields, self.device)
        self.n = self.t.shape[0]
        self.precision = float(precision)
        self.lay = chop_layout(self.fields)
        ws = C.c_size_t()
        _lib.check(self.lib.lsim_chop_query(self.n, C.byref(ws)))
        self.ws = torch.empty(max(ws.value, 8), dtype=torch.uint8, device=self.device)
        self.info = torch.empty(3, dtype=torch.int64, device=self.device)
        _lib.check(self.lib.lsim_chop_count(C.byref(self.lay), C.c_void_p(self.t.data_ptr()), self.n,
                                            C.c_double(self.precision), C.c_void_p(self.ws.data_ptr()),
                                            self.ws.numel(), C.c_void_p(self.info.data_ptr()), self._stream()))
        total, bad, _ = self.info.tolist()
        if bad:
            raise ValueError(f"{bad} track(s) with a non-finite length or more than 2^31 - 1 steps")
        self.n_steps = int(total)

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def fill(self, n_rows, src_events=None, batch_events=None, n_local=-1):
        out = torch.empty((n_rows, len(self.fields)), dtype=torch.float32, device=self.device)
        sp = bp = None
        nb = 0
        if src_events is not None:
            sp = torch.as_tensor(np.asarray(src_events, dtype=np.int64)).to(self.device)
            bp = torch.as_tensor(np.asarray(batch_events, dtype=np.int64)).to(self.device)
            if sp.numel() != self.n:
                raise ValueError("source_event_ids needs one id per track row")
            nb = bp.numel()
        _lib.check(self.lib.lsim_chop_fill(
            C.byref(self.lay), C.c_void_p(self.t.data_ptr()), self.n, C.c_double(self.precision),
            C.c_void_p(self.ws.data_ptr()), self.ws.numel(), self.n_steps, n_rows,
            C.c_void_p(sp.data_ptr() if sp is not None and sp.numel() else None),
            C.c_void_p(bp.data_ptr() if bp is not None and nb else None), nb, n_local,
            C.c_void_p(out.data_ptr()), C.c_void_p(self.info.data_ptr()), self._stream()))
        if src_events is not None and self.n:
            missing = int(self.info[2].item())
            if missing:
                raise ValueError(f"Failed to remap eventID to local indices: {missing} track row(s) have an eventID "
                                 "missing from the batch event ids")
        return out


def chop_tracks(tracks, fields, precision=0.001, device=None) -> torch.Tensor:
    """optimize/dataio.py:63-106: every track row becomes max(ceil(length / precision), 1) step rows.

    ``tracks`` is a float32 (N, F) array or tensor, ``fields`` its column names.  Returns the
    (sum of steps, F) float32 tensor on the GPU, equal bit for bit to the reference's NumPy output.
    """
    ch = _Chopper(tracks, fields, precision, device)
    return ch.fill(ch.n_steps)


def chop_batch(tracks, fields, precision=0.001, target_len=None, source_event_ids=None, batch_event_ids=None,
               device=None) -> torch.Tensor:
    """One chopped, padded batch as ``TracksDataset.__getitem__`` builds it (dataio.py:378-400).

    * ``source_event_ids`` / ``batch_event_ids`` given: each row's eventID becomes the index of its
      source id among the sorted batch ids (``remap_event_ids_to_local``, :47-61); an id that is not
      among them raises ``ValueError``.
    * ``target_len`` above the chopped length: zero rows are appended up to it and invalidated
      (eventID -1, n_electrons / dE / dEdx / dx / long_diff / tran_diff 0, trackID / pixel_plane -1),
      and, when ``batch_event_ids`` is given, so is every row whose eventID is not a local id
      0 .. len(batch_event_ids) - 1 (``pad_batch``, :357-373).
    """
    ch = _Chopper(tracks, fields, precision, device)
    if source_event_ids is not None and batch_event_ids is None:
        raise ValueError("source_event_ids needs batch_event_ids")
    if source_event_ids is not None and ch.lay.event < 0:
        raise ValueError("fields has no eventID column")
    n_rows = ch.n_steps
    n_local = -1
    if target_len is not None and int(target_len) > ch.n_steps:
        n_rows = int(target_len)
        if batch_event_ids is not None:
            n_local = len(batch_event_ids)
    return ch.fill(n_rows, source_event_ids, batch_event_ids, n_local)


def group_events_into_batches(traj_event_ids, traj_lengths, max_batch_len, max_nbatch=None):
    """Whole events packed into batches by cumulative track length (dataio.py:245-289).

    ``traj_event_ids[i]`` and ``traj_lengths[i]`` are trajectory i's event and summed dx.
    Trajectories longer than ``max_batch_len`` are dropped; the events' lengths are accumulated in
    float64 in trajectory order, and a new batch starts wherever floor(cumulative length /
    max_batch_len) increases.  Returns (list of trajectory-index lists, total length kept).
    """
    ev = np.asarray(traj_event_ids)
    ln = np.asarray(traj_lengths)
    keep = np.flatnonzero(ln <= max_batch_len)
    if keep.size == 0:
        raise ValueError("All tracks are longer than the batch size! Please check.")
    uniq, inv = np.unique(ev[keep], return_inverse=True)
    evt_len = np.bincount(inv, weights=ln[keep].astype(np.float64), minlength=uniq.size)
    order = np.argsort(inv, kind="stable")
    grp_start = np.searchsorted(inv[order], np.arange(uniq.size + 1))
    cum = np.cumsum(evt_len)
    cut = np.flatnonzero(np.diff(np.floor_divide(cum, max_batch_len)) > 0) + 1
    bounds = np.concatenate([[0], cut, [uniq.size]])
    if max_nbatch and max_nbatch > 0:
        bounds = bounds[:max_nbatch + 1]
    members = keep[order]
    batches = [members[grp_start[a]:grp_start[b]].tolist() for a, b in zip(bounds[:-1], bounds[1:])]
    return batches, float(cum[bounds[-1] - 1])
