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
