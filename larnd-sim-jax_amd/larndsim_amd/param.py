"""Parametrized current model on the GPU (``simulate_parametrized``, sim_jax.py:339-372).

The reference's second simulation mode: no response LUT, but an analytic induced current per
(electron, pixel) pair (detsim_jax.py:441-639) scattered into the pixel waveforms, then the same
self-trigger readout as the LUT path.  Two HIP entry points carry it (``lsim_param_electrons``,
``lsim_param_rows``, float64); the pixel set between them is a sort on the device (the reference
takes ``jnp.unique`` there too), and the readout is ``lsim_fee_forward``.

The electrons come either from the transverse bins of ``apply_tran_diff`` (``mc_diff`` off; the
reference needs ``tran_diff_bin_edges`` set then) or, with ``mc_diff``, from one normal draw per
segment and axis: the draws are taken with torch's generator from ``rngseed`` (or passed in as
``normals``), not from jax's stream, so a run matches the reference in distribution, and exactly
for the same draws.  Forward only: parameters that require gradients are refused.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .sim import get_simulator, hits_dict, pad_size, params_struct


def _check(params):
    from .autograd import needs_grad

    if needs_grad(params):
        raise NotImplementedError("gradients of the parametrized current model are not available")


def electrons(params, tracks, fields, rngseed=0, normals=None, bin_edges=None):
    """The electrons of the drifted segments and their pixel pairs (``lsim_param_electrons``):
    dict of device tensors x, y, t0, ne, sig (float64, (E,)), plane (int32), pair_pids (int64, (E, K)),
    and the normal draws used (mc_diff)."""
    sim = get_simulator(params, None)
    t, seg = sim.segments(tracks, fields)
    n = int(seg.n)
    dev = sim.device
    mc = bool(params.mc_diff)
    diffusion = 1 if params.diffusion_in_current_sim else 0
    edges_t = None
    if mc:
        if normals is None:
            g = torch.Generator(device=dev).manual_seed(int(rngseed))
            normals = torch.randn((n, 3), generator=g, device=dev, dtype=torch.float32)
        normals = torch.as_tensor(normals).to(device=dev, dtype=torch.float32).contiguous()
        if tuple(normals.shape) != (n, 3):
            raise ValueError(f"normals must be ({n}, 3)")
        per = 1
    else:
        edges = bin_edges if bin_edges is not None else params.tran_diff_bin_edges
        if edges is None:
            raise ValueError("apply_tran_diff needs tran_diff_bin_edges (or mc_diff)")
        edges_t = torch.as_tensor(np.asarray(edges, np.float64)).to(dev).contiguous()
        per = (edges_t.numel() - 1) ** 2
    nn = int(params.number_pix_neighbors)
    K = (2 * nn + 1) ** 2
    E = n * per
    f64 = dict(dtype=torch.float64, device=dev)
    out = dict(x=torch.empty(E, **f64), y=torch.empty(E, **f64), t0=torch.empty(E, **f64), ne=torch.empty(E, **f64),
               sig=torch.empty(E, **f64), plane=torch.empty(E, dtype=torch.int32, device=dev),
               pair_pids=torch.empty((E, K), dtype=torch.int64, device=dev), normals=normals)
    vp = lambda a: C.c_void_p(a.data_ptr()) if a is not None else None  # noqa: E731
    _lib.check(sim.lib.lsim_param_electrons(
        sim._ctx, C.byref(params_struct(params)), C.byref(seg), 0 if not mc else 1, diffusion, vp(normals),
        vp(edges_t), 0 if edges_t is None else edges_t.numel(), vp(out["x"]), vp(out["y"]), vp(out["t0"]),
        vp(out["ne"]), vp(out["sig"]), vp(out["plane"]), vp(out["pair_pids"]), sim._stream()))
    del t  # the uploaded tracks are no longer read
    return out


def waveforms(params, e):
    """(wfs (P, n_ticks - 1) float32, unique_pixels (P,) int64 sorted) of the electrons' pairs
    (``lsim_param_rows``); rows with a negative id (pixels outside the plane) are zero."""
    sim = get_simulator(params, None)
    pids = e["pair_pids"].reshape(-1)
    U, inv = torch.unique(pids, sorted=True, return_inverse=True)
    P = U.numel()
    order = torch.argsort(inv, stable=True).to(torch.int32)
    offs = torch.zeros(P + 1, dtype=torch.int32, device=pids.device)
    offs[1:] = torch.cumsum(torch.bincount(inv, minlength=P), 0).to(torch.int32)
    wfs = torch.empty((P, sim.nout), dtype=torch.float32, device=pids.device)
    vp = lambda a: C.c_void_p(a.data_ptr())  # noqa: E731
    _lib.check(sim.lib.lsim_param_rows(sim._ctx, C.byref(params_struct(params)), vp(e["x"]), vp(e["y"]), vp(e["t0"]),
                                       vp(e["ne"]), vp(e["sig"]), vp(U), P, vp(order), vp(offs),
                                       1 if params.diffusion_in_current_sim else 0, vp(wfs), sim._stream()))
    return wfs, U


def simulate_parametrized(params, tracks, fields, rngseed=0, normals=None, bin_edges=None):
    """sim_jax.py:339-372: (adcs, pixel_x, pixel_y, pixel_z, ticks, hit_prob, event, hit_pixels), the
    compacted hits of the parametrized current model (noise off).  The pixel set is padded by
    pad_size(..., "unique_pixels") with -1 rows in front, as the reference's sort leaves it."""
    _check(params)
    sim = get_simulator(params, None)
    e = electrons(params, tracks, fields, rngseed, normals, bin_edges)
    wfs, U = waveforms(params, e)
    P = U.numel()
    pad = pad_size(P, "unique_pixels") - P
    if pad > 0:
        U = torch.cat([torch.full((pad,), -1, dtype=torch.int64, device=U.device), U])
        wfs = torch.cat([torch.zeros((pad, wfs.shape[1]), dtype=wfs.dtype, device=wfs.device), wfs])
    h = hits_dict(sim.fee(params, wfs, U))
    return (h["adcs"], h["pixel_x"], h["pixel_y"], h["pixel_z"], h["ticks"], h["hit_prob"], h["event"],
            h["hit_pixels"])
