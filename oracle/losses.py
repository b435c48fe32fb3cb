"""Dense restatements of the chamfer and NLL fit losses -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module (the checker of the event-blocked HIP versions in
``larndsim_amd.losses``); the product path never does.

They follow the reference literally, on torch CPU tensors so that autograd gives the
reverse-mode gradients (argmin and the ``where`` masks carry none, ``take`` routes the cotangent
to the chosen pair), in the reference's float32 unless ``dtype`` says otherwise:

* ``chamfer_3d`` (src/larndsim/losses_jax.py:195-249): positions (x + event 1e9, y, z), padded to
  the given sizes with zero points of zero weight, the full pairwise squared-distance matrix, the
  argmins both ways, the weighted sums of the distances below 1e4;
* ``nll_loss`` (:329-361): 4-D points (x + event 1e9, y, z, Q / adc_norm) padded with -1e9, the full
  Gaussian kernel matrix, lambda_j = sum_i p_i K_ij, loss = sum p - sum_j log(lambda_j + 1e-10)
  [ref_p_j > 0.5].

No event blocking: every cross-event pair is formed, so the blocked GPU versions are checked
against the reference's own way of separating events.  Pinned by closed forms (single pairs, a
tie, a padding point nearer than the partner, identical clouds, events far apart) in
tests/test_losses_oracle.py.
"""
from __future__ import annotations

import torch


def _positions(x, y, z, event, dtype):
    xo = x.to(torch.float32) + event.to(torch.float32) * 1e9  # float32, as the reference computes it
    return torch.stack((xo, y.to(torch.float32), z.to(torch.float32)), dim=-1).to(dtype)


def chamfer_3d(Q, x, y, z, event, ref_Q, ref_x, ref_y, ref_z, ref_event, new_cur, new_ref, adc_norm=10.0,
               dtype=torch.float32):
    """losses_jax.py:231-249 with the padded sizes given (the reference takes them from pad_size)."""
    a = _positions(x, y, z, event, dtype)
    b = _positions(ref_x, ref_y, ref_z, ref_event, dtype)
    a = torch.cat((a, torch.zeros((new_cur - a.shape[0], 3), dtype=dtype)))
    b = torch.cat((b, torch.zeros((new_ref - b.shape[0], 3), dtype=dtype)))
    wa = torch.cat((Q.to(dtype), torch.zeros(new_cur - Q.shape[0], dtype=dtype))) / adc_norm
    wb = torch.cat((ref_Q.to(dtype), torch.zeros(new_ref - ref_Q.shape[0], dtype=dtype))) / adc_norm
    d = ((a[:, None, :] - b[None, :, :]) ** 2).sum(dim=2)
    iab = torch.argmin(d.detach(), dim=1)
    iba = torch.argmin(d.detach(), dim=0)
    dab = d.gather(1, iab[:, None]).squeeze(1)
    dba = d.gather(0, iba[None, :]).squeeze(0)
    la = torch.where(dab < 1e4, dab * wa * wb[iab], torch.zeros_like(dab))
    lb = torch.where(dba < 1e4, dba * wb * wa[iba], torch.zeros_like(dba))
    return la.sum() + lb.sum()


def nll_loss(Q, x, y, z, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_hit_prob, ref_event, adc_norm=10.0,
             sigma=1.0, pad=0, dtype=torch.float32):
    """losses_jax.py:329-361 with ``pad`` extra -1e9 points on each side."""
    c = torch.cat((_positions(x, y, z, event, dtype), (Q.to(dtype) / adc_norm)[:, None]), dim=1)
    r = torch.cat((_positions(ref_x, ref_y, ref_z, ref_event, dtype), (ref_Q.to(dtype) / adc_norm)[:, None]), dim=1)
    c = torch.cat((c, torch.full((pad, 4), -1e9, dtype=dtype)))
    r = torch.cat((r, torch.full((pad, 4), -1e9, dtype=dtype)))
    p = torch.cat((hit_prob.to(dtype), torch.zeros(pad, dtype=dtype)))
    rp = torch.cat((ref_hit_prob.to(dtype), torch.zeros(pad, dtype=dtype)))
    d = ((c[:, None, :] - r[None, :, :]) ** 2).sum(dim=-1)
    K = torch.exp(-d / (2 * sigma ** 2))
    lam = (p[:, None] * K).sum(dim=0)
    term1 = (torch.log(lam + 1e-10) * (rp > 0.5)).sum()
    return p.sum() - term1


def mse_adc(Q, x, y, z, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_hit_prob, ref_event, sigma=1.0, lambda_Q=1.0,
            dtype=torch.float64):
    """losses_jax.py:13-39, 58-80 literally: the hit clouds (x + event 1e5, y, z) weighted by
    Q hit_prob, their RBF-kernel MMD over every pair (no event blocking), plus the squared relative
    error of the total weighted charge."""
    w_ref = (ref_Q * ref_hit_prob).to(dtype)
    w = (Q * hit_prob).to(dtype)
    a = torch.stack((x.to(dtype) + event.to(dtype) * 1e5, y.to(dtype), z.to(dtype)), dim=-1)
    b = torch.stack((ref_x.to(dtype) + ref_event.to(dtype) * 1e5, ref_y.to(dtype), ref_z.to(dtype)), dim=-1)

    def K(u, v):
        return torch.exp(-((u[:, None, :] - v[None, :, :]) ** 2).sum(dim=-1) / (2 * sigma ** 2))

    sw, sr = w.sum(), w_ref.sum()
    mmd = ((K(a, a) * w[:, None] * w[None, :]).sum() / sw ** 2 + (K(b, b) * w_ref[:, None] * w_ref[None, :]).sum() / sr ** 2
           - 2 * (K(a, b) * w[:, None] * w_ref[None, :]).sum() / (sw * sr))
    return mmd + lambda_Q * ((sw - sr) / (sr + 1e-6)) ** 2
