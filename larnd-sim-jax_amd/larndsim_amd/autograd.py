"""torch.autograd.Function over the HIP forward and hand-written adjoint.

The differentiable inputs are the 7 physics parameters (Ab, kb, eField, lifetime, vdrift,
long_diff, tran_diff), the differentiable outputs the compacted hit ADC values and hit z
positions.  This is the derivative ``jax.value_and_grad`` takes through
``LUTSimulation.predict`` in the reference (optimize/fit_params.py:713-731).
"""
from __future__ import annotations

import torch

from .consts import GRAD_PARAMS
from .sim import hits_dict


class LUTSimFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, sim, base_params, tracks, fields, n_events):
        vals = [float(v) for v in theta.detach().cpu().tolist()]
        p = base_params.replace(**dict(zip(GRAD_PARAMS, vals)))
        st = sim.forward(p, tracks, fields, n_events=n_events, want_wfs=True, want_tape=True)
        h = hits_dict(st.out)
        ctx.state = st
        ctx.sim = sim
        ctx.hit_index = h["hit_index"].long()
        ctx.theta_device = theta.device
        ctx.set_materialize_grads(False)
        # the non-differentiable outputs, kept for the caller; the differentiable ones (adcs,
        # pixel_z) are not stored here: a reference from the context back to its own outputs
        # would keep every step's workspace alive until the cycle collector runs
        others = {k: v for k, v in h.items() if k not in ("adcs", "pixel_z")}
        others["unique_pixels"] = st.out["pixel_ids"][: st.max_rows]
        others["wfs"] = st.out["wfs"][: st.max_rows]
        others["negative_rows"] = sim.negative_rows(p, st.seg)
        ctx.others = others
        wfs = st.out["wfs"][: st.max_rows]  # a view object of its own (not the one kept above)
        return h["adcs"], h["pixel_z"], wfs

    @staticmethod
    def backward(ctx, g_adc, g_z, g_wfs):
        st, H = ctx.state, ctx.sim.H
        d_wfs = g_wfs if (g_wfs is not None and bool(torch.any(g_wfs != 0))) else None
        n = max(st.max_rows, 1) * H
        dev = ctx.sim.device
        d_adc = torch.zeros(n, dtype=torch.float32, device=dev)
        if g_adc is not None:
            d_adc.index_copy_(0, ctx.hit_index, g_adc.to(torch.float32))
        d_z = None
        if g_z is not None:
            d_z = torch.zeros(n, dtype=torch.float32, device=dev)
            d_z.index_copy_(0, ctx.hit_index, g_z.to(torch.float32))
        grad = ctx.sim.backward(st, d_adc, d_z, d_wfs).to(ctx.theta_device)
        ctx.state = None  # the workspace and tape are not needed again (no double backward)
        return grad, None, None, None, None, None


def theta_of(params):
    """Stack the 7 differentiable parameter values (tensors keep their autograd history)."""
    vals = []
    for n in GRAD_PARAMS:
        v = getattr(params, n)
        vals.append(v.to(torch.float64).reshape(()) if isinstance(v, torch.Tensor) else torch.tensor(float(v), dtype=torch.float64))
    devs = {v.device for v in vals}
    dev = next(iter(devs)) if len(devs) == 1 else torch.device("cpu")
    return torch.stack([v.to(dev) for v in vals])


def needs_grad(params) -> bool:
    """True when autograd would record the simulation: grad mode is on and one of the 7 parameters
    is a tensor that requires a gradient (under ``torch.no_grad()`` nothing is recorded)."""
    if not torch.is_grad_enabled():
        return False
    return any(isinstance(getattr(params, n), torch.Tensor) and getattr(params, n).requires_grad for n in GRAD_PARAMS)
