"""Gradient-descent fit step of the detector parameters (optimize/fit_params.py:168-199, 667-896).

Reproduces the reference fitter's update for the differentiable LUT path:

* parameters live in an unconstrained space and map to physical values with the sigmoid
  (or exp-log) normalisation of ``map_norm_to_phys`` (fit_params.py:168-179), over the
  ``min``/``max`` bounds of optimize/ranges.py;
* the loss is ``loss_strategy.compute(params, sim_strategy.predict(params, ...), target)``
  differentiated with respect to the normalised parameters (fit_params.py:706-731); here torch
  autograd runs the chain rule through the normalisation and the hand-written adjoint of the
  simulation (``larndsim_amd.autograd``);
* the update is optax's ``chain(clip_by_global_norm(c), multi_transform({p: adam(lr_p)}))``
  (fit_params.py:837-848) written out: Adam with bias correction (b1 0.9, b2 0.999, eps 1e-8),
  one learning rate or schedule per parameter, constant or staircase exponential decay;
* NaN gradients skip the update (fit_params.py:876-879).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

from .fee import adc2charge
from .losses import mse_loss
from .strategies import LossStrategy

# Bounds and nominal values of the fit parameters (optimize/ranges.py:7-21).
RANGES = {
    "lArDensity": dict(nom=1.38, down=1.37, up=1.41, min=1.35, max=1.43),
    "eField": dict(nom=0.5, down=0.45, up=0.55, min=0.4, max=0.6),
    "vdrift": dict(nom=0.1596, down=0.1400, up=0.1800, min=0.13, max=0.19),
    "MeVToElectrons": dict(nom=4.237e4, down=3.48e4, up=5.13e4, min=3e4, max=6e4),
    "alpha": dict(nom=0.93, down=0.85, up=1.1, min=0.8, max=1.2),
    "beta": dict(nom=0.212, down=0.18, up=0.23, min=0.15, max=0.25),
    "R_param": dict(nom=1.25, down=1, up=10, min=1, max=10),
    "Ab": dict(nom=0.8, down=0.78, up=0.88, min=0.75, max=0.9),
    "kb": dict(nom=0.0486, down=0.04, up=0.07, min=0.03, max=0.08),
    "lifetime": dict(nom=2.2e3, down=500, up=5000, min=400, max=6000),
    "long_diff": dict(nom=4.0e-6, down=2e-6, up=9e-6, min=1e-6, max=1e-5),
    "tran_diff": dict(nom=8.8e-6, down=4e-6, up=14e-6, min=3e-6, max=15e-6),
    "shift_x": dict(nom=0, down=-0.22, up=0.22, min=-0.44, max=0.44),
    "shift_y": dict(nom=0, down=-0.22, up=0.22, min=-0.44, max=0.44),
    "shift_z": dict(nom=0, down=-0.22, up=0.22, min=-0.44, max=0.44),
}

_F32_EPS = 2.0 ** -23
_F32_TINY = 2.0 ** -126


def map_norm_to_phys(val, key, scheme="sigmoid", scale=1.0):
    """Unconstrained -> physical value (fit_params.py:168-179)."""
    if scheme == "sigmoid":
        lo, hi = RANGES[key]["min"], RANGES[key]["max"]
        return lo + (hi - lo) * torch.sigmoid(scale * val)
    if scheme == "exp_log":
        nom = RANGES[key]["nom"]
        if nom <= 0:
            raise ValueError(f"exp_log normalization requires positive nominal value for '{key}', got {nom}")
        return nom * torch.exp(scale * val)
    raise ValueError(f"Unsupported normalization scheme in map_norm_to_phys: {scheme}")


def map_phys_to_norm(val, key, scheme="sigmoid", scale=1.0):
    """Physical -> unconstrained value (fit_params.py:182-199)."""
    if scale <= 0:
        raise ValueError(f"Normalization scale must be > 0 for scheme '{scheme}', got {scale}")
    val = torch.as_tensor(val, dtype=torch.float64)
    if scheme == "sigmoid":
        lo, hi = RANGES[key]["min"], RANGES[key]["max"]
        frac = torch.clamp((val - lo) / (hi - lo), _F32_EPS, 1.0 - _F32_EPS)
        return torch.log(frac / (1.0 - frac)) / scale
    if scheme == "exp_log":
        nom = RANGES[key]["nom"]
        if nom <= 0:
            raise ValueError(f"exp_log normalization requires positive nominal value for '{key}', got {nom}")
        return torch.log(torch.clamp(val, min=_F32_TINY) / nom) / scale
    raise ValueError(f"Unsupported normalization scheme in map_phys_to_norm: {scheme}")


def constant_schedule(value):
    return lambda count: value


def exponential_decay(init_value, transition_steps, decay_rate, transition_begin=0, staircase=False, end_value=None):
    """optax.exponential_decay: init * rate ** ((count - begin) / steps), floored with staircase."""
    def f(count):
        c = max(count - transition_begin, 0)
        p = c / transition_steps
        if staircase:
            p = math.floor(p)
        v = init_value * decay_rate ** p
        if end_value is not None:
            v = max(v, end_value) if decay_rate < 1 else min(v, end_value)
        return v
    return f


def clip_by_global_norm(grads: dict, max_norm: float) -> dict:
    """optax.clip_by_global_norm: scale all gradients by max_norm / ||g|| when ||g|| > max_norm."""
    norm = math.sqrt(sum(float(g) ** 2 for g in grads.values()))
    if norm < max_norm:
        return dict(grads)
    return {k: g / norm * max_norm for k, g in grads.items()}


@dataclass
class AdamState:
    mu: float = 0.0
    nu: float = 0.0
    count: int = 0


@dataclass
class Adam:
    """optax.adam (b1 0.9, b2 0.999, eps 1e-8, eps_root 0) for one scalar parameter."""

    lr: object
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1e-8
    state: AdamState = field(default_factory=AdamState)

    def update(self, g: float) -> float:
        s = self.state
        lr = self.lr(s.count) if callable(self.lr) else self.lr
        s.mu = self.b1 * s.mu + (1 - self.b1) * g
        s.nu = self.b2 * s.nu + (1 - self.b2) * g * g
        s.count += 1
        mu_hat = s.mu / (1 - self.b1 ** s.count)
        nu_hat = s.nu / (1 - self.b2 ** s.count)
        return -lr * mu_hat / (math.sqrt(nu_hat) + self.eps)


class PixelChargeLossStrategy(LossStrategy):
    """Per-pixel squared difference of the summed hit charge (``mse_loss``, losses_jax.py:41-56,
    on the hits' charge in ke- from ``adc2charge``, losses_jax.py:380-383)."""

    def compute(self, params, prediction, target):
        Q = adc2charge(prediction["adcs"], params)
        ref_Q = adc2charge(target["adcs"], params)
        pix = prediction["hit_pixels"].to(Q.device)
        ref_pix = target["hit_pixels"].to(Q.device)
        return mse_loss(Q.reshape(-1, 1).double(), pix, ref_Q.reshape(-1, 1).double(), ref_pix)


class GradientDescentFitter:
    """Adam fit of ``relevant_params`` (fit_params.py:781-896) over normalised parameters.

    ``lr`` is one value for every parameter or a dict per parameter; ``lr_scheduler`` is None
    (constant) or ``"exponential_decay"`` with ``lr_kw`` (e.g. ``decay_rate``), staircase over
    ``epoch_size`` steps, as in the reference.
    """

    def __init__(self, sim_strategy, loss_strategy, ref_params, relevant_params, lr, max_clip_norm_val=None,
                 lr_scheduler=None, lr_kw=None, epoch_size=1, normalization_scheme="sigmoid", norm_scale=1.0,
                 init_params=None):
        self.sim_strategy = sim_strategy
        self.loss_strategy = loss_strategy
        self.ref_params = ref_params
        self.relevant = list(relevant_params)
        self.scheme = normalization_scheme
        self.scale = norm_scale
        self.max_clip_norm_val = max_clip_norm_val
        lrs = lr if isinstance(lr, dict) else {k: lr for k in self.relevant}
        self.opt = {}
        for k in self.relevant:
            if lr_scheduler is None:
                sched = constant_schedule(float(lrs[k]))
            elif lr_scheduler == "exponential_decay":
                sched = exponential_decay(float(lrs[k]), transition_steps=epoch_size, staircase=True, **(lr_kw or {}))
            else:
                raise ValueError("The specified optimizer schedules are not yet supported")
            self.opt[k] = Adam(sched)
        start = init_params if init_params is not None else ref_params
        self.norm = {k: float(map_phys_to_norm(float(getattr(start, k)), k, self.scheme, self.scale)) for k in self.relevant}
        self.history = {k: [] for k in self.relevant}
        self.history["loss"] = []

    def physical(self) -> dict:
        return {k: float(map_norm_to_phys(torch.tensor(v, dtype=torch.float64), k, self.scheme, self.scale))
                for k, v in self.norm.items()}

    def current_params(self):
        return self.ref_params.replace(**self.physical())

    def compute_loss(self, tracks, fields, target, with_grad=True):
        """(loss, {param: d loss / d normalised param}, aux) -- fit_params.py:667-740."""
        norm_t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=with_grad) for k, v in self.norm.items()}
        phys = {k: map_norm_to_phys(t, k, self.scheme, self.scale) for k, t in norm_t.items()}
        params = self.ref_params.replace(**phys)
        pred = self.sim_strategy.predict(params, tracks, fields, None)
        loss, aux = self.loss_strategy.compute(params, pred, target)
        if not with_grad:
            return float(loss.detach()), None, aux
        loss.backward()
        grads = {k: float(t.grad) if t.grad is not None else 0.0 for k, t in norm_t.items()}
        return float(loss.detach()), grads, aux

    def process_grads(self, grads: dict) -> dict:
        if any(math.isnan(g) for g in grads.values()):
            return grads
        g = clip_by_global_norm(grads, self.max_clip_norm_val) if self.max_clip_norm_val else grads
        for k in self.relevant:
            self.norm[k] += self.opt[k].update(g[k])
        return grads

    def step(self, tracks, fields, target):
        loss, grads, _ = self.compute_loss(tracks, fields, target)
        self.process_grads(grads)
        self.history["loss"].append(loss)
        for k, v in self.physical().items():
            self.history[k].append(v)
        return loss, grads
