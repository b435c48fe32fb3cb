"""The float64 soft-DTW restatement (oracle/sdtw.py) against brute force over alignment paths,
the hard DTW limit and central differences; no GPU."""
import itertools
import math

import numpy as np
import pytest

from oracle import sdtw as SD


def _paths(n, m):
    """Every monotone alignment path from (0, 0) to (n-1, m-1) with steps (1,0), (0,1), (1,1)."""
    out = []

    def walk(i, j, acc):
        if (i, j) == (n - 1, m - 1):
            out.append(acc)
            return
        for di, dj in ((1, 0), (0, 1), (1, 1)):
            if i + di < n and j + dj < m:
                walk(i + di, j + dj, acc + [(i + di, j + dj)])

    walk(0, 0, [(0, 0)])
    return out


@pytest.mark.parametrize("n,m,g", [(1, 1, 1.0), (2, 3, 1.0), (4, 4, 0.5), (5, 3, 2.0), (3, 5, 0.1)])
def test_value_is_soft_minimum_over_all_paths(n, m, g):
    rng = np.random.default_rng(n * 10 + m)
    x, y = rng.normal(size=n), rng.normal(size=m)
    D = SD.cost(x, y)
    costs = np.array([sum(D[i, j] for i, j in p) for p in _paths(n, m)])
    mn = costs.min()
    want = mn - g * math.log(np.exp(-(costs - mn) / g).sum())
    got, _ = SD.soft_dtw(x, y, g)
    assert got == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_hard_limit_and_identity():
    rng = np.random.default_rng(1)
    x, y = rng.normal(size=6), rng.normal(size=7)
    D = SD.cost(x, y)
    hard = min(sum(D[i, j] for i, j in p) for p in _paths(6, 7))
    assert SD.soft_dtw(x, y, 1e-4)[0] == pytest.approx(hard, abs=1e-2)
    # x aligned with itself: the diagonal costs nothing, soft-DTW is below 0 (other paths add mass)
    v, _ = SD.soft_dtw(x, x, 1.0)
    assert v <= 1e-12


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(2)
    x, y = rng.normal(size=9), rng.normal(size=12)
    v, gx = SD.soft_dtw_grad(x, y, 0.7)
    h = 1e-6
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h
        fd = (SD.soft_dtw(x + e, y, 0.7)[0] - SD.soft_dtw(x - e, y, 0.7)[0]) / (2 * h)
        assert gx[i] == pytest.approx(fd, rel=1e-6, abs=1e-8)
