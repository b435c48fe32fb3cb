"""Pixel-id packing and coordinates (detsim_jax.py:230-319) on torch / numpy integers."""
import numpy as np
import torch

from .consts import get_vdrift


def _ntpc(params):
    return int(np.asarray(params.tpc_borders).shape[0])


def pixel2id(params, pixel_x, pixel_y, pixel_plane, eventID):
    """Packed pixel id, -1 outside the plane (detsim_jax.py:230-244); true int64."""
    nx, ny, ntpc = int(params.n_pixels_x), int(params.n_pixels_y), _ntpc(params)
    px, py = torch.as_tensor(pixel_x).long(), torch.as_tensor(pixel_y).long()
    outside = (px >= nx) | (py >= ny) | (px < 0) | (py < 0)
    pid = ((torch.as_tensor(eventID).long() * ntpc + torch.as_tensor(pixel_plane).long()) * ny + py) * nx + px
    return torch.where(outside, torch.full_like(pid, -1), pid)


def id2pixel(params, pid):
    """(x, y, plane, event) of packed ids with floor semantics (detsim_jax.py:265-278)."""
    nx, ny, ntpc = int(params.n_pixels_x), int(params.n_pixels_y), _ntpc(params)
    pid = torch.as_tensor(pid).long()
    return (torch.remainder(pid, nx), torch.remainder(torch.div(pid, nx, rounding_mode="floor"), ny),
            torch.remainder(torch.div(pid, nx * ny, rounding_mode="floor"), ntpc),
            torch.div(pid, nx * ny * ntpc, rounding_mode="floor"))


def get_pixel_coordinates(params, xpitch, ypitch, plane):
    """Pixel-centre coordinates (detsim_jax.py:296-306)."""
    b = torch.as_tensor(np.asarray(params.tpc_borders, dtype=np.float64))[torch.as_tensor(plane).long().cpu()]
    b = b.to(torch.as_tensor(xpitch).device)
    px = xpitch * params.pixel_pitch + b[..., 0, 0] + params.pixel_pitch / 2
    py = ypitch * params.pixel_pitch + b[..., 1, 0] + params.pixel_pitch / 2
    return torch.stack([px, py], dim=-1)


def get_hit_z(params, ticks, plane):
    """Hit z from the tick (detsim_jax.py:308-319)."""
    b = torch.as_tensor(np.asarray(params.tpc_borders, dtype=np.float64))[torch.as_tensor(plane).long().cpu()]
    b = b.to(torch.as_tensor(ticks).device)
    return b[..., 2, 0] + ticks * params.t_sampling * get_vdrift(params) * torch.sign(b[..., 2, 1] - b[..., 2, 0])


# --- event-id validation (detsim_jax.py:24-152; called by optimize/simulate.py:111-113 and
# optimize/fit_params.py:599-601 before a batch is simulated) --------------------------------

INT64_MAX = int(np.iinfo(np.int64).max)


def _pixel_stride(params):
    return int(params.n_pixels_x) * int(params.n_pixels_y) * _ntpc(params)


def _bin_stride(params):
    nb = int(params.nb_sampling_bins_per_pixel)
    return int(params.n_pixels_x) * nb * int(params.n_pixels_y) * nb * _ntpc(params)


def _stride(params, kind):
    if kind == "pixel":
        return _pixel_stride(params)
    if kind == "bin":
        return _bin_stride(params)
    raise ValueError(f"Unknown packing kind '{kind}'")


def max_safe_event_id(params, kind="pixel"):
    """Largest event id whose packed ids (pixel or sub-pixel bin) stay within int64."""
    s = _stride(params, kind)
    return (INT64_MAX - (s - 1)) // s


def max_possible_id(params, max_event_id, kind="pixel"):
    """The largest packed id an event id can produce: the last cell of its last plane."""
    return (int(max_event_id) + 1) * _stride(params, kind) - 1


def _negatives(ids, what, context):
    bad = ids < -1
    if np.any(bad):
        raise ValueError(f"{context} found {what} below -1: {np.unique(ids[bad])[:16].tolist()}")


def validate_event_ids_for_packing(params, event_ids, kind="pixel", context=""):
    """Event ids must be >= -1 and small enough that their packed ids fit int64."""
    ids = np.asarray(event_ids, dtype=np.int64).ravel()
    if ids.size == 0:
        return
    _negatives(ids, "eventID values", context)
    valid = ids[ids >= 0]
    if valid.size == 0:
        return
    top = int(valid.max())
    if top > max_safe_event_id(params, kind):
        raise OverflowError(f"{context} eventID {top} exceeds the int64-safe limit {max_safe_event_id(params, kind)} "
                            f"for {kind} packing")
    if max_possible_id(params, top, kind) > INT64_MAX:
        raise OverflowError(f"{context} maximum packed {kind} id exceeds int64 max {INT64_MAX}")


def validate_packed_ids_for_decoding(params, packed_ids, kind="pixel", context=""):
    """Packed ids must be >= -1 and decode to event ids within the int64-safe limit."""
    ids = np.asarray(packed_ids, dtype=np.int64).ravel()
    if ids.size == 0:
        return
    _negatives(ids, "packed ids", context)
    valid = ids[ids >= 0]
    if valid.size == 0:
        return
    top = int((valid // _stride(params, kind)).max())
    if top > max_safe_event_id(params, kind):
        raise OverflowError(f"{context} decoded eventID {top} exceeds the int64-safe limit "
                            f"{max_safe_event_id(params, kind)} for {kind} packing")


def validate_local_event_ids(event_ids, context=""):
    """A batch's event ids must be exactly {-1} plus 0 .. n - 1 with no gaps."""
    ids = np.asarray(event_ids, dtype=np.int64).ravel()
    if ids.size == 0:
        return
    _negatives(ids, "eventID values (non-local namespace)", context)
    uniq = np.unique(ids[ids >= 0])
    if uniq.size and not np.array_equal(uniq, np.arange(uniq.size)):
        raise ValueError(f"{context} eventID values do not form contiguous local namespace [0, {uniq.size - 1}]. "
                         f"Found unique IDs: {uniq.tolist()}")
