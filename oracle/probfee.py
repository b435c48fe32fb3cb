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
    wfs = np.asarray(wfs, dtype=np.float32)
    P, T = wfs.shape
    q_sum = np.cumsum((wfs * np.float32(t_sampling)).astype(np.float64), axis=1)
    c = np.zeros((P, V))
    lp = np.full((P, V), -1000.0)
    lp[:, 0] = 0.0
    out_lp = np.full((P, S, T - 1), -1000.0)
    out_q = np.zeros((P, S, T - 1))
    sels = np.full((P, S, T - 1), -np.inf)
    active = True
    for s in range(S):
        if not active:
            continue
        totals = np.empty(P)
        for p in range(P):
            nc, nl, hit, esp, sel = find_one_hit_step(q_sum[p], c[p], lp[p], sigma, theta, interval, V)
            c[p], lp[p] = nc, nl
            out_lp[p, s], out_q[p, s], sels[p, s] = hit, esp, sel
            totals[p] = logsumexp(nl, 0)
        active = bool(np.any(totals > np.log(stop)))
    return out_lp, out_q, sels
