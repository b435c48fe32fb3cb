"""Hit output files (optimize/simulate.py:125-179): the selection of stored hits, the per-event
HDF5 group layout, and the single npz file.

Host-side bookkeeping after the GPU readout: the hit arrays are copied to the host once and
written.  The HDF5 file is written by the package's own writer (``h5lite``), in the layout
h5py gives the reference's files; no h5py is needed.
"""
from __future__ import annotations

import numpy as np
import torch

from .fee import adc2charge, digitize

NPZ_KEYS = ("adcs", "Q", "x", "y", "z", "ticks", "hit_prob", "event_id")  # simulate.py:178
EVENT_KEYS = ("adc_clean", "adc", "Q", "pixels", "ticks", "eventID", "pix_x", "pix_y", "pix_z")  # :147-155


def _np(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def select_hits(params, adcs, pixel_x, pixel_y, pixel_z, ticks, hit_prob, event, hit_pixels):
    """The stored hits of one batch (simulate.py:130-133): slots whose ADC differs from the ADC of
    the discrimination threshold and whose event is not padding; Q = adc2charge(adc).  Inputs are
    the readout arrays of ``simulate_stochastic`` (any shape, flattened in order)."""
    adc_lowest = float(digitize(params, torch.tensor(float(params.DISCRIMINATION_THRESHOLD), dtype=torch.float32)))
    adc = _np(adcs).reshape(-1).astype(np.float32)
    ev = _np(event).reshape(-1)
    clean = adc - np.float32(adc_lowest)
    mask = (clean != 0) & (ev != -1)
    adc_m = adc[mask]
    q = _np(adc2charge(torch.from_numpy(adc_m.astype(np.float64)), params)).astype(np.float32)
    return dict(adc=adc_m, adc_clean=clean[mask], Q=q, pixels=_np(hit_pixels).reshape(-1)[mask],
                ticks=_np(ticks).reshape(-1)[mask].astype(np.float32), event=ev[mask].astype(np.int64),
                pix_x=_np(pixel_x).reshape(-1)[mask].astype(np.float32),
                pix_y=_np(pixel_y).reshape(-1)[mask].astype(np.float32),
                pix_z=_np(pixel_z).reshape(-1)[mask].astype(np.float32),
                hit_prob=_np(hit_prob).reshape(-1)[mask].astype(np.float32))


def event_groups(hits, local_to_global=None):
    """Per-event datasets of one batch as simulate.py:139-155 writes them: group name
    ``event_<global id>``, datasets ``EVENT_KEYS`` (eventID holds the global id)."""
    out = {}
    for loc in np.unique(hits["event"]).astype(int):
        if loc < 0:
            continue
        g = int(local_to_global.get(loc, loc)) if local_to_global else loc
        m = hits["event"] == loc
        out[f"event_{g}"] = dict(adc_clean=hits["adc_clean"][m], adc=hits["adc"][m], Q=hits["Q"][m],
                                 pixels=hits["pixels"][m], ticks=hits["ticks"][m],
                                 eventID=np.full(int(m.sum()), g, dtype=np.int64), pix_x=hits["pix_x"][m],
                                 pix_y=hits["pix_y"][m], pix_z=hits["pix_z"][m])
    return out


def write_hdf5(path, batch_index, groups):
    """Append ``batch_<i>/event_<g>/<dataset>`` groups to an HDF5 file (simulate.py:135-155, which
    opens the file in append mode and creates one batch group per call).  The file's groups are
    read back and the whole file is rewritten with the new batch; like h5py's create_group, an
    existing batch group of the same name raises ValueError."""
    import os

    from . import h5lite

    tree = h5lite.read_tree(path) if os.path.exists(path) else {}
    name = f"batch_{batch_index}"
    if name in tree:
        raise ValueError(f"{path}: group {name} already exists")
    tree[name] = {ev: {k: np.asarray(v) for k, v in ds.items()} for ev, ds in groups.items()}
    h5lite.write_tree(path, tree)


def read_hdf5(path):
    """The groups of a hit file as nested dicts {batch: {event: {dataset: array}}}."""
    from . import h5lite

    return h5lite.read_tree(path)


def save_npz(path, batches):
    """One npz of all batches' stored hits with the keys of simulate.py:177-178."""
    cat = lambda k: np.concatenate([h[k] for h in batches]) if batches else np.zeros(0)  # noqa: E731
    np.savez(path, adcs=cat("adc"), Q=cat("Q"), x=cat("pix_x"), y=cat("pix_y"), z=cat("pix_z"), ticks=cat("ticks"),
             hit_prob=cat("hit_prob"), event_id=cat("event"))


def load_npz(path):
    """The arrays of a hit npz (either this writer's or the reference's); no pickles are loaded."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
