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


# ------------------------------------------------------------------------------------------
# the on-disk segment files and the dataset of batches (optimize/dataio.py:44-45, 108-424)
# ------------------------------------------------------------------------------------------

def load_segments(filename, swap_xz=True):
    """The ``segments`` compound dataset of a larnd-sim input file as a numpy structured array,
    prepared as ``TracksDataset.__init__`` does (dataio.py:113-139): x and z columns swapped
    (x, x_start, x_end <-> z, z_start, z_end), a zero ``t0`` column appended when absent, and
    event_id / traj_id renamed to eventID / trackID.  Read with the package's own HDF5 reader."""
    from numpy.lib import recfunctions as rfn

    from .h5lite import read_dataset

    tracks = read_dataset(filename, "segments")
    if swap_xz:
        for a, b in (("x_start", "z_start"), ("x_end", "z_end"), ("x", "z")):
            ta, tb = np.copy(tracks[a]), np.copy(tracks[b])
            tracks[a], tracks[b] = tb, ta
    if "t0" not in tracks.dtype.names:
        tracks = rfn.append_fields(tracks, "t0", np.zeros(tracks.shape[0]), usemask=False)
    replace = {"event_id": "eventID", "traj_id": "trackID"}
    tracks.dtype.names = tuple(replace.get(n, n) for n in tracks.dtype.names)
    return tracks


def np_from_structured(tracks) -> np.ndarray:
    """(N, F) float32 of a structured array, columns in field order (dataio.py:44-45)."""
    from numpy.lib import recfunctions as rfn

    return np.asarray(rfn.structured_to_unstructured(tracks, copy=True, dtype=np.float32), dtype=np.float32)


class TracksDataset:
    """Batches of whole events from a segment file (dataio.py:108-424): the same selection, event
    choice, event-grouped batch split and per-batch step count as the reference; a batch is
    chopped, remapped and padded on the GPU (``chop_batch``).  ``__getitem__`` returns a float32
    device tensor (N, F) in ``get_track_fields()`` order."""

    def __init__(self, filename, nevents, max_nbatch=None, swap_xz=True, random_nevents=False, data_seed=42,
                 track_len_sel=2., max_abs_costheta_sel=0.966, min_abs_segz_sel=15., track_z_bound=28.,
                 max_batch_len=50, print_input=False, chopped=True, pad=True, electron_sampling_resolution=0.1,
                 live_selection=False, device=None):
        tracks = load_segments(filename, swap_xz)
        self.track_fields = tracks.dtype.names
        # load only what the batches can use, trimmed at an event boundary (dataio.py:145-157)
        if max_batch_len is not None and max_nbatch is not None and max_nbatch > 0:
            threshold = max_batch_len * (max_nbatch + 2)
            cutoff = int(np.searchsorted(np.cumsum(tracks["dx"]), threshold, side="left"))
            if cutoff < len(tracks):
                other = tracks["eventID"][:cutoff + 1] != tracks["eventID"][cutoff]
                if other.any():
                    tracks = tracks[:int(np.where(other)[0][-1]) + 1]
        # trajectories (dataio.py:162-224); the row lists index the loaded table, also with the
        # live selection, whose keys come from the selected rows (as in the reference)
        if live_selection:
            sel = tracks[abs(tracks["z"]) < min_abs_segz_sel]
            _, first = np.unique(sel[["eventID", "trackID"]], return_index=True)
            first = np.sort(first)
            last = np.r_[first[1:] - 1, len(sel) - 1]
            s, e = sel[first], sel[last]
            d = np.column_stack((s["x_start"] - e["x_end"], s["y_start"] - e["y_end"], s["z_start"] - e["z_end"]))
            cos = np.abs(d @ np.array([0, 0, 1])) / (np.linalg.norm(d, axis=1) + 1e-10)
            m = (np.sqrt((d ** 2).sum(axis=1)) > track_len_sel) & (cos < max_abs_costheta_sel)
            m &= np.maximum(abs(s["z"]), abs(e["z"])) < track_z_bound
            keys = np.ascontiguousarray(sel[np.repeat(m, last - first + 1)][["eventID", "trackID"]])
        else:
            keys = np.ascontiguousarray(tracks[["eventID", "trackID"]])
        index, inverse = np.unique(keys, return_inverse=True)
        self.traj_keys = np.ascontiguousarray(index)
        self.tracks_struct = tracks
        sorted_idx = np.argsort(inverse)
        vals, starts = np.unique(inverse[sorted_idx], return_index=True)
        ends = np.append(starts[1:], len(sorted_idx))
        traj_rows = [np.empty(0, dtype=int) for _ in range(len(index))]
        for v, a, b in zip(vals, starts, ends):
            traj_rows[int(v)] = sorted_idx[a:b]
        event_ids = np.array([tracks[traj_rows[i][0]]["eventID"] for i in range(len(traj_rows))])
        uniq, first_idx = np.unique(event_ids, return_index=True)
        ordered = uniq[np.argsort(first_idx)]
        if nevents is not None and nevents > 0:
            if random_nevents:
                rng = np.random.default_rng(seed=data_seed)
                chosen = rng.choice(ordered, size=nevents, replace=False) if nevents < len(ordered) else ordered
            else:
                chosen = ordered[:nevents]
            traj_rows = [traj_rows[i] for i in np.where(np.isin(event_ids, chosen))[0]]
        self.trajectory_row_indices = traj_rows
        if max_batch_len is not None:
            lengths = np.array([tracks[r]["dx"].sum() for r in traj_rows])
            evs = np.array([tracks[r[0]]["eventID"] for r in traj_rows])
            self.batch_traj_indices, self.tot_data_length = group_events_into_batches(evs, lengths, max_batch_len,
                                                                                      max_nbatch)
        else:
            self.batch_traj_indices = [[i] for i in range(len(traj_rows))]
            self.tot_data_length = float(np.sum([tracks[r]["dx"].sum() for r in traj_rows]))
        if min(len(b) for b in self.batch_traj_indices) == 0:
            raise ValueError("There exist some empty batch in the simulation input!")
        self.batch_nsteps, self.batch_row_keys, self.batch_event_global_ids, self.batch_row_indices = [], [], [], []
        for b in self.batch_traj_indices:
            rows = np.concatenate([traj_rows[t] for t in b])
            steps = np.maximum(np.ceil(tracks[rows]["dx"] / electron_sampling_resolution), 1).astype(int)
            self.batch_nsteps.append(int(steps.sum()))
            bk = np.ascontiguousarray(np.unique(tracks[rows][["eventID", "trackID"]]))
            self.batch_row_keys.append(bk)
            self.batch_event_global_ids.append(np.asarray(np.unique(bk["eventID"]), dtype=np.int64))
            self.batch_row_indices.append(np.asarray(rows, dtype=np.int64))
        self.max_batch_nsteps = int(max(self.batch_nsteps)) if self.batch_nsteps else 0
        self.chopped, self.pad = chopped, pad
        self.electron_sampling_resolution = electron_sampling_resolution
        self.print_input = print_input
        self.device = device

    def __len__(self):
        return len(self.batch_traj_indices)

    def __getitem__(self, idx):
        if idx < 0 or idx >= len(self):
            raise IndexError("Batch index out of range")
        rows = np.concatenate([self.trajectory_row_indices[t] for t in self.batch_traj_indices[idx]])
        sel = self.tracks_struct[rows]
        arr = np_from_structured(sel)
        target = self.max_batch_nsteps if self.pad and self.max_batch_nsteps > 0 else None
        if not self.chopped:  # remap and pad on the host (dataio.py:47-61, 357-373)
            ids = self.batch_event_global_ids[idx]
            src = np.asarray(sel["eventID"], dtype=np.int64)
            loc = np.searchsorted(ids, src)
            if np.any(loc >= ids.size) or not np.array_equal(ids[np.minimum(loc, ids.size - 1)], src):
                raise ValueError("Failed to remap eventID to local indices")
            f = list(self.track_fields)
            ev = f.index("eventID")
            arr[:, ev] = loc.astype(np.float32)
            n = arr.shape[0]
            if target is not None and target > n:
                arr = np.concatenate([arr, np.zeros((target - n, arr.shape[1]), np.float32)])
                bad = np.zeros(target, bool)
                bad[n:] = True
                bad |= (arr[:, ev] >= 0) & ~np.isin(arr[:, ev].astype(np.int64), np.arange(ids.size))
                arr[bad, ev] = -1
                for name in ZERO_PAD_FIELDS:
                    if name in f:
                        arr[bad, f.index(name)] = 0
                for name in NEG_ONE_PAD_FIELDS:
                    if name in f:
                        arr[bad, f.index(name)] = -1
            dev = self.device if self.device is not None else torch.device("cuda")
            return torch.from_numpy(arr).to(dev)
        return chop_batch(arr, self.track_fields, precision=self.electron_sampling_resolution, target_len=target,
                          source_event_ids=sel["eventID"], batch_event_ids=self.batch_event_global_ids[idx],
                          device=self.device)

    def get_track_fields(self):
        return self.track_fields

    def get_batch_global_event_ids(self, idx=None):
        return self.batch_event_global_ids if idx is None else self.batch_event_global_ids[idx]

    def get_batch_row_keys(self):
        return self.batch_row_keys

    def get_batch_row_indices(self, idx=None):
        return self.batch_row_indices if idx is None else self.batch_row_indices[idx]
