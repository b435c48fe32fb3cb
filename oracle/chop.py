"""CPU restatement of the track chopping step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module (the checker of the HIP ``lsim_chop_*`` kernels); the
product path never does.

Follows ``chop_tracks`` (optimize/dataio.py:63-106): every track row is cut into
n = max(ceil(length / precision), 1) steps along its straight line.  The arithmetic follows
NumPy 2's type rules on the reference's float32 (N, F) array, which fix every rounding:

* length, direction, the step count and the per-step dE are float32 (``length / precision``
  and ``dE * precision / (length + 1e-10)`` keep float32: the Python float is weak);
* step positions are float64 (``steps * precision`` is int64 x float), rounded to float32 when
  stored: start_k = s + (k precision) d, end_k = s + (precision (k + 1)) d;
* the last step ends at the track's end, carries dE (1 - precision (n - 1) / (length + 1e-10))
  and dx = length - precision (n - 1), both evaluated in float64;
* the midpoint x = 0.5 (start + end) is float32; every other column is copied.

Pinned by the reference's own outputs on seeded tracks (tests/golden/chop_tracks_vectors.npz,
written by tests/golden/make_chop_vectors.py) and by invariants (step count, lengths summing to
the track length, dE summing to the track dE, contiguous steps, the exact end point).
"""
from __future__ import annotations

import numpy as np


def step_counts(tracks: np.ndarray, fields, precision: float) -> np.ndarray:
    """Steps per track row (float32 length, as the reference computes it)."""
    f = list(fields)
    s = np.stack([tracks[:, f.index(c + "_start")] for c in "xyz"], axis=1).astype(np.float32)
    e = np.stack([tracks[:, f.index(c + "_end")] for c in "xyz"], axis=1).astype(np.float32)
    seg = e - s
    sq = seg * seg
    length = np.sqrt((sq[:, 0] + sq[:, 1]) + sq[:, 2])  # float32, summed left to right
    return np.maximum(np.ceil(length / np.float32(precision)), 1).astype(np.int64)


def chop_tracks(tracks: np.ndarray, fields, precision: float = 0.001) -> np.ndarray:
    """optimize/dataio.py:63-106 on a float32 (N, F) array -> float32 (sum n_i, F)."""
    f = list(fields)
    t = np.asarray(tracks, dtype=np.float32)
    iS = [f.index(c + "_start") for c in "xyz"]
    iE = [f.index(c + "_end") for c in "xyz"]
    iM = [f.index(c) for c in "xyz"]
    iDE, iDX = f.index("dE"), f.index("dx")
    s = t[:, iS]
    seg = t[:, iE] - s
    sq = seg * seg
    length = np.sqrt((sq[:, 0] + sq[:, 1]) + sq[:, 2])
    den = length + np.float32(1e-10)
    direc = seg / den[:, None]
    n = np.maximum(np.ceil(length / np.float32(precision)), 1).astype(np.int64)
    row = np.repeat(np.arange(len(t)), n)
    off = np.concatenate([[0], np.cumsum(n)])
    k = np.arange(len(row)) - off[row]
    out = t[row].copy()
    out[:, iDE] = (t[row, iDE] * np.float32(precision)) / den[row]
    kp = k.astype(np.float64) * precision
    k1p = precision * (k + 1).astype(np.float64)
    for c in range(3):
        out[:, iS[c]] = (s[row, c].astype(np.float64) + kp * direc[row, c].astype(np.float64)).astype(np.float32)
        out[:, iE[c]] = (s[row, c].astype(np.float64) + k1p * direc[row, c].astype(np.float64)).astype(np.float32)
    out[:, iDX] = np.float32(precision)
    last = off[1:] - 1
    for c in range(3):
        out[last, iE[c]] = t[:, iE[c]]
    nm1 = (n - 1).astype(np.float64)
    out[last, iDE] = (t[:, iDE].astype(np.float64) * (1.0 - precision * nm1 / den.astype(np.float64))).astype(np.float32)
    out[last, iDX] = (length.astype(np.float64) - precision * nm1).astype(np.float32)
    for c in range(3):
        out[:, iM[c]] = np.float32(0.5) * (out[:, iS[c]] + out[:, iE[c]])
    return out


ZERO_PAD = ("n_electrons", "dE", "dEdx", "dx", "long_diff", "tran_diff")
NEG_ONE_PAD = ("trackID", "pixel_plane")


def remap_event_ids(arr, fields, source_ids, batch_ids):
    """optimize/dataio.py:47-61: eventID <- searchsorted(batch_ids, source_ids), ids must exist."""
    f = list(fields)
    src = np.asarray(source_ids, dtype=np.int64)
    bat = np.asarray(batch_ids, dtype=np.int64)
    out = np.array(arr, dtype=np.float32, copy=True)
    if src.size == 0:
        return out
    loc = np.searchsorted(bat, src)
    if np.any(loc >= bat.size) or not np.array_equal(bat[np.minimum(loc, bat.size - 1)], src):
        raise ValueError("missing event id")
    out[:, f.index("eventID")] = loc.astype(np.float32)
    return out


def invalidate(arr, fields, mask):
    """optimize/dataio.py:347-355 on the rows of mask."""
    f = list(fields)
    arr[mask, f.index("eventID")] = -1
    for name in ZERO_PAD:
        if name in f:
            arr[mask, f.index(name)] = 0
    for name in NEG_ONE_PAD:
        if name in f:
            arr[mask, f.index(name)] = -1
    return arr


def pad_batch(arr, fields, target_len, n_local=None):
    """optimize/dataio.py:357-373: zero rows up to target_len, invalidated; with n_local, rows whose
    eventID e >= 0 is not in arange(n_local) after int64 truncation are invalidated too."""
    f = list(fields)
    out = np.array(arr, dtype=np.float32, copy=True)
    cur = out.shape[0]
    if target_len > cur:
        out = np.concatenate([out, np.zeros((target_len - cur, out.shape[1]), np.float32)])
        mask = np.zeros(target_len, bool)
        mask[cur:] = True
        out = invalidate(out, fields, mask)
    if n_local is not None:
        ev = out[:, f.index("eventID")]
        bad = (ev >= 0) & ~np.isin(ev.astype(np.int64), np.arange(n_local, dtype=np.int64))
        out = invalidate(out, fields, bad)
    return out


def group_events(traj_event_ids, traj_lengths, max_batch_len, max_nbatch=None):
    """optimize/dataio.py:245-289 as explicit loops over trajectories and events."""
    ev = list(traj_event_ids)
    ln = list(traj_lengths)
    kept = [i for i in range(len(ln)) if ln[i] <= max_batch_len]
    if not kept:
        raise ValueError("all trajectories longer than max_batch_len")
    events = sorted(set(ev[i] for i in kept))
    pos = {e: j for j, e in enumerate(events)}
    groups = [[] for _ in events]
    elen = [0.0] * len(events)
    for i in kept:
        groups[pos[ev[i]]].append(i)
        elen[pos[ev[i]]] += float(ln[i])
    batches, cur, run = [], [], 0.0
    prev_bin = None
    cum_at_end = []
    for j in range(len(events)):
        run += elen[j]
        b = run // max_batch_len  # floor division as np.floor_divide computes it
        if prev_bin is not None and b > prev_bin:
            batches.append(cur)
            cur = []
        cur = cur + groups[j]
        cum_at_end.append(run)
        prev_bin = b
    batches.append(cur)
    if max_nbatch and max_nbatch > 0:
        batches = batches[:max_nbatch]
    n_ev = sum(len(set(ev[i] for i in b)) for b in batches)
    return batches, cum_at_end[n_ev - 1]
