"""Simulation / loss strategies with the reference's plugin interface (optimize/strategies.py).

``HipLUTSimulation`` is the drop-in for ``LUTSimulation`` (strategies.py:87-105): the same
``predict(params, tracks, fields, rngkey) -> dict`` with keys adcs, pixel_x, pixel_y, pixel_z,
ticks, hit_prob, event, hit_pixels, unique_pixels, wfs.  If any of the 7 physics parameters in
``params`` is a torch tensor with ``requires_grad``, ``adcs`` and ``pixel_z`` carry autograd
history through the hand-written adjoint.
"""
from __future__ import annotations

import torch

from .autograd import LUTSimFunction, needs_grad, theta_of
from .fee import adc2charge
from .sim import get_simulator, hits_dict, infer_n_events, reference_rows


class SimulationStrategy:
    def predict(self, params, tracks, fields, rngkey):
        """Runs the simulation and returns a dictionary of outputs (strategies.py:80-85)."""
        raise NotImplementedError


class HipLUTSimulation(SimulationStrategy):
    """LUT simulation on the GPU (strategies.py:87-105).  ``response`` is the raw (Nx, Ny, Nt) LUT
    returned by ``consts.load_lut``.  Noise must be off (rngkey is accepted and unused)."""

    def __init__(self, response):
        self.response = response

    def predict(self, params, tracks, fields, rngkey=None):
        """``unique_pixels`` and ``wfs`` have the reference's row layout (``sim.reference_rows``:
        negative ids, then ``pad_size`` -1 rows with zero waveforms, then the pixel rows); the hit
        arrays are compacted as ``parse_output`` does."""
        sim = get_simulator(params, self.response)
        n_events = infer_n_events(tracks, fields)
        if needs_grad(params):
            theta = theta_of(params)
            base = params.replace(**{n: float(v) for n, v in zip(
                ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff"), theta.detach().cpu().tolist())})
            adcs, z, wfs = LUTSimFunction.apply(theta, sim, base, tracks, fields, n_events)
            # the non-differentiable outputs were kept on the function's context
            o = _last_others(adcs)
            wfs, ids = reference_rows(wfs, o["unique_pixels"], *o["negative_rows"])
            return dict(adcs=adcs, pixel_x=o["pixel_x"], pixel_y=o["pixel_y"], pixel_z=z, ticks=o["ticks"],
                        hit_prob=o["hit_prob"], event=o["event"], hit_pixels=o["hit_pixels"],
                        unique_pixels=ids, wfs=wfs)
        st = sim.forward(params, tracks, fields, n_events=n_events, want_wfs=True)
        neg = sim.negative_rows(params, st.seg)
        h = hits_dict(st.out)
        wfs, ids = reference_rows(st.out["wfs"][: st.max_rows], st.out["pixel_ids"][: st.max_rows], *neg)
        return dict(adcs=h["adcs"], pixel_x=h["pixel_x"], pixel_y=h["pixel_y"], pixel_z=h["pixel_z"], ticks=h["ticks"],
                    hit_prob=h["hit_prob"], event=h["event"], hit_pixels=h["hit_pixels"], unique_pixels=ids, wfs=wfs)


def pad_to_closest_multiple(x, dims_to_pad=None, multiple=128, pad_value=0, pad_front=False):
    """Pad the given dimensions of x up to the next multiple (strategies.py:41-77), at the front or
    the back."""
    x = torch.as_tensor(x)
    dims = range(x.dim()) if dims_to_pad is None else dims_to_pad
    shape = list(x.shape)
    for d in dims:
        shape[d] = ((x.shape[d] + multiple - 1) // multiple) * multiple
    out = torch.full(shape, pad_value, dtype=x.dtype, device=x.device)
    sl = tuple(slice(shape[i] - n, None) if pad_front else slice(0, n) for i, n in enumerate(x.shape))
    out[sl] = x
    return out


class HipLUTProbabilisticSimulation(SimulationStrategy):
    """LUT waveforms and the probabilistic readout on the GPU (strategies.py:107-134): the raw
    (pixel, hit index, tick) distributions for the probabilistic losses.  With parameters that
    require gradients, hit_prob and adcs_distrib carry autograd history through the readout's
    adjoint (selections held fixed) and the waveform adjoint of the LUT simulation."""

    def __init__(self, response):
        self.response = response

    def predict(self, params, tracks, fields, rngkey=None):
        from .detsim import id2pixel
        from .fee import simulate_probabilistic
        from .sim import simulate_wfs

        # the waveforms do not depend on the readout noise, which only the readout uses
        quiet = params.replace(RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)
        if needs_grad(params):
            sim = get_simulator(quiet, self.response)
            theta = theta_of(quiet)
            base = quiet.replace(**{n: float(v) for n, v in zip(
                ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff"), theta.detach().cpu().tolist())})
            adcs, _, wfs = LUTSimFunction.apply(theta, sim, base, tracks, fields, infer_n_events(tracks, fields))
            o = _last_others(adcs)
            wfs, unique_pixels = reference_rows(wfs, o["unique_pixels"], *o["negative_rows"])
        else:
            wfs, unique_pixels = simulate_wfs(quiet, self.response, tracks, fields)
        unique_pixels = pad_to_closest_multiple(unique_pixels, multiple=128, pad_value=-1, pad_front=True)
        wfs = pad_to_closest_multiple(wfs, dims_to_pad=(0,), multiple=128, pad_value=0.0, pad_front=True)
        adcs_distrib, pixel_x, pixel_y, ticks_prob, event = simulate_probabilistic(params, wfs, unique_pixels)
        _, _, pixel_plane, _ = id2pixel(params, unique_pixels)
        return dict(adcs_distrib=adcs_distrib, hit_prob=ticks_prob, pixel_x=pixel_x, pixel_y=pixel_y,
                    pixel_plane=pixel_plane, event=event, unique_pixels=unique_pixels, hit_pixels=unique_pixels,
                    wfs=wfs)


class HipParametrizedSimulation(SimulationStrategy):
    """The parametrized current model on the GPU (strategies.py:136-149): the compacted hit arrays
    of ``param.simulate_parametrized`` under the reference's keys.  ``rngkey`` seeds the normal
    draws of mc_diff.  With parameters that require gradients, ``adcs`` and ``pixel_z`` carry
    autograd history through the model's adjoint."""

    def predict(self, params, tracks, fields, rngkey=0):
        from .param import simulate_parametrized

        adcs, x, y, z, ticks, hit_prob, event, hit_pixels = simulate_parametrized(
            params, tracks, fields, rngseed=0 if rngkey is None else int(rngkey))
        return dict(adcs=adcs, pixel_x=x, pixel_y=y, pixel_z=z, ticks=ticks, hit_prob=hit_prob, event=event,
                    unique_pixels=hit_pixels)


def _last_others(t):
    return t.grad_fn.others if t.grad_fn is not None else {}


class LossStrategy:
    def __init__(self, **kwargs):
        self.kwargs = kwargs

    def compute(self, params, prediction, target):
        raise NotImplementedError


class GenericLossStrategy(LossStrategy):
    """adc -> charge, then ``loss_fn(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ...)``
    (strategies.py:162-180)."""

    def __init__(self, loss_fn, **kwargs):
        super().__init__(**kwargs)
        self.loss_fn = loss_fn

    def compute(self, params, prediction, target):
        Q = adc2charge(prediction["adcs"], params)
        ref_Q = adc2charge(target["adcs"], params)
        return self.loss_fn(params, Q, prediction["pixel_x"], prediction["pixel_y"], prediction["pixel_z"],
                            prediction["ticks"], prediction["hit_prob"], prediction["event"], ref_Q, target["pixel_x"],
                            target["pixel_y"], target["pixel_z"], target["ticks"], target["hit_prob"], target["event"],
                            **self.kwargs)
