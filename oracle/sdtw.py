"""Float64 soft-DTW and its gradient -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module (the checker of the HIP ``lsim_sdtw`` kernels); the product
path never does.

Soft dynamic time warping (Cuturi and Blondel, "Soft-DTW: a differentiable loss function for
time-series", ICML 2017, Algorithms 1 and 2) with the squared-Euclidean ground cost in the form
the reference's SoftDTW uses (src/larndsim/softdtw_jax.py:180-275): D[i,j] = (-2 x_i y_j + x_i^2)
+ y_j^2, R[i,j] = D[i,j] + softmin_g(R[i-1,j-1], R[i-1,j], R[i,j-1]), R[-1,-1] = 0 and +inf
elsewhere outside, value R[n-1,m-1]; E = d value / d D by the backward recursion, and
d value / d x_i = sum_j E[i,j] (2 x_i - 2 y_j).  Evaluated by anti-diagonals.

Pinned by tests/test_sdtw_oracle.py: the value equals -g log sum over every monotone alignment path
of exp(-path cost / g), enumerated by brute force on short sequences; g -> 0 gives the hard DTW
minimum; the gradient equals central differences.
"""
from __future__ import annotations

import numpy as np


def cost(x, y):
    x = np.asarray(x, np.float64)[:, None]
    y = np.asarray(y, np.float64)[None, :]
    return (-2.0 * x * y + x * x) + y * y


def softmin(a, b, c, g):
    m = np.minimum(a, np.minimum(b, c))
    with np.errstate(invalid="ignore", over="ignore"):
        s = np.exp(-(a - m) / g) + np.exp(-(b - m) / g) + np.exp(-(c - m) / g)
    return m - g * np.log(s)


def soft_dtw(x, y, g=1.0):
    """(value, R) with R (n + 1, m + 1), R[0, 0] = 0 and R[i + 1, j + 1] the table above."""
    D = cost(x, y)
    n, m = D.shape
    R = np.full((n + 1, m + 1), np.inf)
    R[0, 0] = 0.0
    for d in range(2, n + m + 1):  # i + j = d over the 1-based table
        i = np.arange(max(1, d - m), min(n, d - 1) + 1)
        j = d - i
        R[i, j] = D[i - 1, j - 1] + softmin(R[i - 1, j - 1], R[i - 1, j], R[i, j - 1], g)
    return R[n, m], R


def soft_dtw_grad(x, y, g=1.0):
    """(value, d value / d x)."""
    D = cost(x, y)
    n, m = D.shape
    val, R = soft_dtw(x, y, g)
    Dp = np.zeros((n + 2, m + 2))
    Dp[1:n + 1, 1:m + 1] = D
    Rp = np.full((n + 2, m + 2), -np.inf)
    Rp[:n + 1, :m + 1] = R
    Rp[n + 1, m + 1] = R[n, m]
    E = np.zeros((n + 2, m + 2))
    E[n + 1, m + 1] = 1.0
    for d in range(n + m, 1, -1):
        i = np.arange(max(1, d - m), min(n, d - 1) + 1)
        j = d - i
        a = np.exp((Rp[i + 1, j] - Rp[i, j] - Dp[i + 1, j]) / g)
        b = np.exp((Rp[i, j + 1] - Rp[i, j] - Dp[i, j + 1]) / g)
        c = np.exp((Rp[i + 1, j + 1] - Rp[i, j] - Dp[i + 1, j + 1]) / g)
        E[i, j] = E[i + 1, j] * a + E[i, j + 1] * b + E[i + 1, j + 1] * c
    E = E[1:n + 1, 1:m + 1]
    xv, yv = np.asarray(x, np.float64), np.asarray(y, np.float64)
    return val, (E * (2.0 * xv[:, None] - 2.0 * yv[None, :])).sum(axis=1)
