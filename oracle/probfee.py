"""NumPy float64 restatement of the probabilistic readout -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module (the checker of the HIP ``lsim_prob_fee`` kernels); the
product path never does.

Follows ``_find_one_hit_step`` and ``get_adc_values_average_noise_vmap``
(src/larndsim/fee_jax.py:334-461) literally, path by path: cumulative maxima of q_sum minus each
path's charge, ``log_diff_ndtr`` with its smooth clamps (:13-54), the -1000 soft clamp and the
expected-charge soft mask, log-sum-exp over paths, the future-hit term per path, ``top_k`` with
the lower index first on ties, and the global stop.  It does not use the GPU version's
simplifications (one cumulative maximum per row, the path-independent future-hit term), so it
checks them.  scipy.special.log_ndtr supplies log Phi.

Pinned by tests/test_probfee_oracle.py: log_diff_ndtr against log(Phi(a) - Phi(b)) away from the
clamps, a single clean pulse (all probability of hit 0 on its crossing tick, nothing for hit 1),
and small noise reproducing the deterministic self-trigger ticks of the oracle.
"""
from __future__ import annotations

import numpy as np
from scipy.special import expit, log_ndtr


def softplus(x):
    return np.logaddexp(x, 0.0)


def soft_max(x, lo, s):
    return lo + softplus((x - lo) * s) / s


def soft_where(c, t, f, s):
    w = expit(c * s)
    return w * t + (1.0 - w) * f


def log_diff_ndtr(a, b):
    """fee_jax.py:32-54."""
    la, lb = log_ndtr(a), log_ndtr(b)
    sd = np.where(a > b, lb - la, -1.0)
    ne = soft_max(-np.expm1(sd), 1e-30, 1e10)
    lt = soft_max(np.log(ne), -100.0, 10.0)
    return soft_where(a - b, la + lt, -1000.0, 1000.0)


def logsumexp(x, axis):
    m = np.max(x, axis=axis, keepdims=True)
    return np.squeeze(m + np.log(np.sum(np.exp(x - m), axis=axis, keepdims=True)), axis=axis)


def find_one_hit_step(q_sum, prev_c, prev_lp, sigma, theta, interval, V):
    """fee_jax.py:334-384 for one row; returns (new charges, new log probs, hit log prob per tick,
    expected charge per tick, selection score per tick)."""
    T = q_sum.shape[0]
    zs = 1.0 / sigma
    sh = np.clip(np.arange(T - 1) + interval + 1, 0, T - 1)
    qsl = q_sum[None, :] - prev_c[:, None]
    qmf = np.maximum.accumulate(qsl, axis=1)
    lg = log_diff_ndtr((qmf[:, 1:] - theta) * zs, (qmf[:, :-1] - theta) * zs)
    le = log_diff_ndtr((qsl[:, sh] - theta) * zs, (qsl[:, :-1] - theta) * zs)
    le = np.minimum(le, lg)
    le = soft_max(le, -1000.0, 1.0)
    esp = q_sum[sh] + theta - 0.5 * (q_sum[1:] + q_sum[:-1])
    le = soft_where(esp - theta, le, -1000.0, 10.0)
    hit = logsumexp(le + prev_lp[:, None], 0)
    tot = logsumexp(lg + prev_lp[:, None], 0)
    nq = qsl[:, np.clip(sh + 1, 0, T - 1)]
    mfs = np.maximum.accumulate(qsl[:, ::-1], axis=1)[:, ::-1]
    fe = np.clip(sh + interval + 1, 0, T - 1)
    lf = log_ndtr((mfs[:, fe] - nq - theta) * zs)
    sel = logsumexp(lg + lf + prev_lp[:, None], 0)
    top = np.argsort(-sel, kind="stable")[:V]
    return q_sum[np.clip(sh[top] + 1, 0, T - 1)], tot[top], hit, esp, sel


def prob_fee(wfs, t_sampling, V, S, interval, theta, sigma, stop=1e-9):
    """fee_jax.py:386-461: (log_prob, charge) of shape (P, S, T - 1), plus the selection scores
    (P, S, T - 1) of every executed step (for the tests' tie checks)."""
    out_lp, out_q, sels, _, _ = prob_fee_tape(wfs, t_sampling, V, S, interval, theta, sigma, stop)
    return out_lp, out_q, sels


def prob_fee_tape(wfs, t_sampling, V, S, interval, theta, sigma, stop=1e-9):
    """prob_fee, also returning the selected ticks of every executed step, (P, S, V) (the discrete
    decisions reverse-mode AD holds fixed: top_k is under stop_gradient, fee_jax.py:377), and the
    number of executed steps."""
    wfs = np.asarray(wfs, dtype=np.float32)
    P, T = wfs.shape
    q_sum = np.cumsum((wfs * np.float32(t_sampling)).astype(np.float64), axis=1)
    c = np.zeros((P, V))
    lp = np.full((P, V), -1000.0)
    lp[:, 0] = 0.0
    out_lp = np.full((P, S, T - 1), -1000.0)
    out_q = np.zeros((P, S, T - 1))
    sels = np.full((P, S, T - 1), -np.inf)
    tops = np.zeros((P, S, V), np.int64)
    active = True
    n_steps = 0
    for s in range(S):
        if not active:
            continue
        n_steps += 1
        totals = np.empty(P)
        for p in range(P):
            nc, nl, hit, esp, sel = find_one_hit_step(q_sum[p], c[p], lp[p], sigma, theta, interval, V)
            tops[p, s] = np.argsort(-sel, kind="stable")[:V]
            c[p], lp[p] = nc, nl
            out_lp[p, s], out_q[p, s], sels[p, s] = hit, esp, sel
            totals[p] = logsumexp(nl, 0)
        active = bool(np.any(totals > np.log(stop)))
    return out_lp, out_q, sels, tops, n_steps


def prob_fee_torch(wfs, t_sampling, V, S, interval, theta, sigma, tops, n_steps, q_f32=True):
    """The executed steps of the beam search in torch float64 with the path selections ``tops``
    held fixed, differentiable in ``wfs`` (P, T): (log_prob, charge), each (P, S, T - 1), steps past
    ``n_steps`` at -1000 / 0 as the reference's inactive branch writes them.  The same formulas as
    ``find_one_hit_step``, vectorised over rows and paths; torch.cummax sends the gradient of a
    running maximum to the last index that attains it, torch.minimum splits ties evenly.  With
    ``q_f32`` the charge per tick is the float32 product wfs * t_sampling, as in ``prob_fee_tape``
    (the cast passes gradients unchanged); without it the whole function is float64 (for finite
    differences)."""
    import torch

    P, T = wfs.shape
    dev = wfs.device
    zs = 1.0 / sigma
    sh = torch.clamp(torch.arange(T - 1, device=dev) + interval + 1, max=T - 1)

    def softplus(x):
        return torch.logaddexp(x, torch.zeros((), dtype=x.dtype, device=dev))

    def soft_max_t(x, lo, s):
        return lo + softplus((x - lo) * s) / s

    def soft_where_t(c, t, f, s):
        w = torch.sigmoid(c * s)
        return w * t + (1.0 - w) * f

    def log_diff(a, b):
        la, lb = torch.special.log_ndtr(a), torch.special.log_ndtr(b)
        sd = torch.where(a > b, lb - la, torch.full_like(a, -1.0))
        ne = soft_max_t(-torch.expm1(sd), 1e-30, 1e10)
        lt = soft_max_t(torch.log(ne), -100.0, 10.0)
        return soft_where_t(a - b, la + lt, -1000.0, 1000.0)

    if q_f32:
        q = (wfs.to(torch.float32) * np.float32(t_sampling)).to(wfs.dtype)
    else:
        q = wfs * t_sampling
    q_sum = torch.cumsum(q, dim=1)
    c = torch.zeros((P, V), dtype=wfs.dtype, device=dev)
    lp = torch.full((P, V), -1000.0, dtype=wfs.dtype, device=dev)
    lp = torch.cat([torch.zeros((P, 1), dtype=wfs.dtype, device=dev), lp[:, 1:]], dim=1)
    rows = torch.arange(P, device=dev)[:, None]
    tops = torch.as_tensor(tops, device=dev)
    out_lp, out_q = [], []
    esp = q_sum[:, sh] + theta - 0.5 * (q_sum[:, 1:] + q_sum[:, :-1])
    for s in range(S):
        if s >= n_steps:
            out_lp.append(torch.full((P, T - 1), -1000.0, dtype=wfs.dtype, device=dev))
            out_q.append(torch.zeros((P, T - 1), dtype=wfs.dtype, device=dev))
            continue
        qsl = q_sum[:, None, :] - c[:, :, None]
        qmf, _ = torch.cummax(qsl, dim=2)
        lg = log_diff((qmf[..., 1:] - theta) * zs, (qmf[..., :-1] - theta) * zs)
        le = log_diff((qsl[..., sh] - theta) * zs, (qsl[..., :-1] - theta) * zs)
        le = torch.minimum(le, lg)
        le = soft_max_t(le, -1000.0, 1.0)
        le = soft_where_t(esp[:, None, :] - theta, le, -1000.0, 10.0)
        out_lp.append(torch.logsumexp(le + lp[:, :, None], dim=1))
        out_q.append(esp)
        tot = torch.logsumexp(lg + lp[:, :, None], dim=1)
        t_k = tops[:, s]
        lp = tot[rows, t_k]
        c = q_sum[rows, torch.clamp(sh[t_k] + 1, max=T - 1)]
    return torch.stack(out_lp, dim=1), torch.stack(out_q, dim=1)
