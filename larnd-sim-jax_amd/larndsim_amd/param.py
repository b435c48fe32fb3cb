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
for the same draws.

If any of the 7 physics parameters is a tensor that requires a gradient, ``adcs`` and
``pixel_z`` (and the waveforms) carry autograd history through the hand-written adjoint
(``lsim_param_backward``): the derivative jax.grad takes through simulate_parametrized, with the
pixel ids, the whole-tick split of each arrival time and the readout's decisions held fixed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .sim import get_simulator, hits_dict, pad_size, params_struct


def electrons(params, tracks, fields, rngseed=0, normals=None, bin_edges=None):
    """The electrons of the drifted segments and their pixel pairs (``lsim_param_electrons``):
    dict of device tensors x, y, t0, ne, sig (float64, (E,)), plane (int32), pair_pids (int64, (E, K)),
    the normal draws used (mc_diff) and the bin edges (transverse bins)."""
    sim = get_simulator(params, None)
    t, seg = sim.segments(tracks, fields)
    out = _electrons(params, sim, seg, rngseed, normals, bin_edges)
    del t  # the uploaded tracks are no longer read
    return out


def _electrons(params, sim, seg, rngseed, normals, bin_edges):
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
               pair_pids=torch.empty((E, K), dtype=torch.int64, device=dev), normals=normals, edges=edges_t)
    vp = lambda a: C.c_void_p(a.data_ptr()) if a is not None else None  # noqa: E731
    _lib.check(sim.lib.lsim_param_electrons(
        sim._ctx, C.byref(params_struct(params)), C.byref(seg), 0 if not mc else 1, diffusion, vp(normals),
        vp(edges_t), 0 if edges_t is None else edges_t.numel(), vp(out["x"]), vp(out["y"]), vp(out["t0"]),
        vp(out["ne"]), vp(out["sig"]), vp(out["plane"]), vp(out["pair_pids"]), sim._stream()))
    return out


def waveforms(params, e):
    """(wfs (P, n_ticks - 1) float32, unique_pixels (P,) int64 sorted) of the electrons' pairs
    (``lsim_param_rows``); rows with a negative id (pixels outside the plane) are zero."""
    wfs, U, _, _ = _waveforms(params, get_simulator(params, None), e)
    return wfs, U


def _waveforms(params, sim, e):
    pids = e["pair_pids"].reshape(-1)
    # one stable sort of the pair pixel ids gives the rows (the distinct ids, ascending), each row's
    # pairs in pair order and the row offsets
    spids, order = torch.sort(pids, stable=True)
    U, counts = torch.unique_consecutive(spids, return_counts=True)
    P = U.numel()
    order = order.to(torch.int32)
    offs = torch.zeros(P + 1, dtype=torch.int32, device=pids.device)
    offs[1:] = torch.cumsum(counts, 0).to(torch.int32)
    wfs = torch.empty((P, sim.nout), dtype=torch.float32, device=pids.device)
    vp = lambda a: C.c_void_p(a.data_ptr())  # noqa: E731
    _lib.check(sim.lib.lsim_param_rows(sim._ctx, C.byref(params_struct(params)), vp(e["x"]), vp(e["y"]), vp(e["t0"]),
                                       vp(e["ne"]), vp(e["sig"]), vp(U), P, vp(order), vp(offs),
                                       1 if params.diffusion_in_current_sim else 0, vp(wfs), sim._stream()))
    return wfs, U, order, offs


def backward(params, sim, seg, e, U, order, offs, tape, d_adc, d_z=None, d_wfs=None):
    """d loss / d (Ab, kb, eField, lifetime, vdrift, long_diff, tran_diff) float64 (7,) on the device
    (``lsim_param_backward``), from cotangents on the readout grid d_adc, d_z (P * H,) and on the
    waveforms d_wfs (P, n_ticks - 1), for the forward that produced ``e``, the rows (U, order, offs)
    and the readout ``tape``."""
    dev = U.device
    P = U.numel()
    E = e["x"].numel()
    sz = C.c_size_t()
    _lib.check(sim.lib.lsim_param_backward_query(sim._ctx, E, P, C.byref(sz)))
    ws = torch.empty(int(sz.value), dtype=torch.uint8, device=dev)
    f32 = lambda a: None if a is None else a.to(device=dev, dtype=torch.float32).contiguous()  # noqa: E731
    d_adc, d_z, d_wfs = f32(d_adc), f32(d_z), f32(d_wfs)
    if d_adc is None or d_adc.numel() != P * sim.H:
        raise ValueError(f"d_adc must hold {P} x {sim.H} values")
    if d_z is not None and d_z.numel() != P * sim.H:
        raise ValueError(f"d_z must hold {P} x {sim.H} values")
    if d_wfs is not None and tuple(d_wfs.shape) != (P, sim.nout):
        raise ValueError(f"d_wfs must be ({P}, {sim.nout})")
    grad = torch.zeros(7, dtype=torch.float64, device=dev)
    vp = lambda a: C.c_void_p(a.data_ptr()) if a is not None else None  # noqa: E731
    edges = e.get("edges")
    _lib.check(sim.lib.lsim_param_backward(
        sim._ctx, C.byref(params_struct(params)), C.byref(seg), 1 if params.mc_diff else 0,
        1 if params.diffusion_in_current_sim else 0, vp(e["normals"]), vp(edges),
        0 if edges is None else edges.numel(), vp(e["x"]), vp(e["y"]), vp(e["t0"]), vp(e["ne"]), vp(e["sig"]), vp(U),
        P, vp(order), vp(offs), vp(tape), vp(d_adc), vp(d_z), vp(d_wfs), vp(ws), ws.numel(), vp(grad),
        sim._stream()))
    return grad


class ParamSimFunction(torch.autograd.Function):
    """The parametrized model as a function of the 7 parameters: outputs the hit ADC values, the hit
    z positions and the waveforms (P, n_ticks - 1) of the unpadded pixel set."""

    @staticmethod
    def forward(ctx, theta, base_params, tracks, fields, rngseed, normals, bin_edges):
        from .consts import GRAD_PARAMS

        p = base_params.replace(**dict(zip(GRAD_PARAMS, [float(v) for v in theta.detach().cpu().tolist()])))
        sim = get_simulator(p, None)
        t, seg = sim.segments(tracks, fields)
        e = _electrons(p, sim, seg, rngseed, normals, bin_edges)
        wfs, U, order, offs = _waveforms(p, sim, e)
        out = sim.fee(p, wfs, U, want_tape=True)
        h = hits_dict(out)
        ctx.state = (p, sim, t, seg, e, U, order, offs, out["tape"])
        ctx.hit_index = h["hit_index"].long()
        ctx.theta_device = theta.device
        ctx.set_materialize_grads(False)
        ctx.others = {k: v for k, v in h.items() if k not in ("adcs", "pixel_z")}
        ctx.others["unique_pixels"] = U
        return h["adcs"], h["pixel_z"], wfs.view(wfs.shape)

    @staticmethod
    def backward(ctx, g_adc, g_z, g_wfs):
        p, sim, t, seg, e, U, order, offs, tape = ctx.state
        n = max(U.numel(), 1) * sim.H
        dev = U.device
        d_adc = torch.zeros(n, dtype=torch.float32, device=dev)
        if g_adc is not None:
            d_adc.index_copy_(0, ctx.hit_index, g_adc.to(torch.float32))
        d_z = None
        if g_z is not None:
            d_z = torch.zeros(n, dtype=torch.float32, device=dev)
            d_z.index_copy_(0, ctx.hit_index, g_z.to(torch.float32))
        d_wfs = g_wfs if (g_wfs is not None and bool(torch.any(g_wfs != 0))) else None
        grad = backward(p, sim, seg, e, U, order, offs, tape, d_adc[: U.numel() * sim.H], d_z if d_z is None else
                        d_z[: U.numel() * sim.H], d_wfs).to(ctx.theta_device)
        ctx.state = None
        return grad, None, None, None, None, None, None


def simulate_parametrized(params, tracks, fields, rngseed=0, normals=None, bin_edges=None):
    """sim_jax.py:339-372: (adcs, pixel_x, pixel_y, pixel_z, ticks, hit_prob, event, hit_pixels), the
    compacted hits of the parametrized current model (noise off).  The readout's rows are the sorted
    pixel set; the reference pads it by pad_size(..., "unique_pixels") with -1 rows in front, which
    carry no charge and no hits, so the hits are the same.  With parameters that require gradients
    ``adcs`` and ``pixel_z`` carry autograd history (``ParamSimFunction``)."""
    from .autograd import needs_grad, theta_of

    if needs_grad(params):
        theta = theta_of(params)
        adcs, z, _ = ParamSimFunction.apply(theta, params, tracks, fields, rngseed, normals, bin_edges)
        o = adcs.grad_fn.others
        return adcs, o["pixel_x"], o["pixel_y"], z, o["ticks"], o["hit_prob"], o["event"], o["hit_pixels"]
    sim = get_simulator(params, None)
    e = electrons(params, tracks, fields, rngseed, normals, bin_edges)
    wfs, U = waveforms(params, e)
    pad_size(U.numel(), "unique_pixels")  # the reference's size history
    h = hits_dict(sim.fee(params, wfs, U))
    return (h["adcs"], h["pixel_x"], h["pixel_y"], h["pixel_z"], h["ticks"], h["hit_prob"], h["event"],
            h["hit_pixels"])
