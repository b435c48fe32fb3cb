"""Losses of the fit (losses_jax.py:13-80) on torch tensors."""
from __future__ import annotations

import torch

from .fee import adc2charge  # noqa: F401  (re-export, losses_jax.py:380-383)


def rbf_kernel(x, y, sigma):
    """exp(-|x - y|^2 / (2 sigma^2)) (losses_jax.py:13-18)."""
    d = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return torch.exp(-d / (2 * sigma ** 2))


def mmd(x, y, px, py, sigma):
    """Probability-weighted maximum mean discrepancy (losses_jax.py:21-39)."""
    kxx = (rbf_kernel(x, x, sigma) * px[:, None] * px[None, :]).sum()
    kyy = (rbf_kernel(y, y, sigma) * py[:, None] * py[None, :]).sum()
    kxy = (rbf_kernel(x, y, sigma) * px[:, None] * py[None, :]).sum()
    sx, sy = px.sum(), py.sum()
    return kxx / sx ** 2 + kyy / sy ** 2 - 2 * kxy / (sx * sy)


def mse_adc(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_ticks, ref_hit_prob, ref_event,
            sigma=1, lambda_Q=1):
    """MMD between weighted hit clouds plus relative total-charge error (losses_jax.py:58-80)."""
    w_ref = ref_Q * ref_hit_prob
    w = Q * hit_prob
    ref_s = torch.stack((ref_x + ref_event * 1e5, ref_y, ref_z), dim=-1)
    s = torch.stack((x + event * 1e5, y, z), dim=-1)
    m = mmd(s, ref_s, w, w_ref, sigma)
    tot_ref, tot = w_ref.sum(), w.sum()
    charge_loss = ((tot - tot_ref) / (tot_ref + 1e-6)) ** 2
    aux = dict(charge_loss=charge_loss, mmd_loss_term=m, Q=Q, ref_Q=ref_Q, ref_hit_prob=ref_hit_prob, hit_prob=hit_prob)
    return m + lambda_Q * charge_loss, aux


def mse_loss(adcs, pIDs, adcs_ref, pIDs_ref):
    """Per-pixel squared difference of (P, H) hit arrays (losses_jax.py:41-56)."""
    allp = torch.unique(torch.cat([pIDs, pIDs_ref]))
    sig = torch.zeros((allp.numel(), adcs.shape[1]), dtype=adcs.dtype, device=adcs.device)
    sig = sig.index_add(0, torch.searchsorted(allp, pIDs), adcs)
    sig = sig.index_add(0, torch.searchsorted(allp, pIDs_ref), -adcs_ref)
    return (sig ** 2).sum(), dict()


def pixel_charge_loss(params, Q, hit_pixels, ref_Q, ref_hit_pixels):
    """Per-pixel charge loss: mse_loss on the summed hit charge of each pixel."""
    return mse_loss(Q[:, None], hit_pixels, ref_Q[:, None], ref_hit_pixels)
