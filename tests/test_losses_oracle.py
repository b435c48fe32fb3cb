"""The dense chamfer and NLL restatements (oracle/losses.py) against closed forms and central
differences; no GPU."""
import math

import pytest
import torch

from oracle import losses as OL


def T(*v, dt=torch.float64):
    return torch.tensor(v, dtype=dt)


def test_chamfer_single_pair():
    # d = 0.25, weights 20 / 10 and 30 / 10: both directions count 0.25 * 2 * 3
    L = OL.chamfer_3d(T(20.0), T(1.0), T(2.0), T(3.0), T(0), T(30.0), T(1.5), T(2.0), T(3.0), T(0), 1, 1,
                      dtype=torch.float64)
    assert float(L) == pytest.approx(3.0, rel=1e-12)


def test_chamfer_identical_clouds_and_far_events():
    g = torch.Generator().manual_seed(0)
    x, y, z = (torch.rand(7, generator=g, dtype=torch.float64) * 10 for _ in range(3))
    q = torch.rand(7, generator=g, dtype=torch.float64) + 0.5
    ev = torch.tensor([0, 0, 1, 1, 1, 2, 2])
    assert float(OL.chamfer_3d(q, x, y, z, ev, q, x, y, z, ev, 7, 7, dtype=torch.float64)) == 0.0
    # simulated hits only in event 0, target hits only in event 1: every nearest distance is ~1e18
    L = OL.chamfer_3d(q[:3], x[:3], y[:3], z[:3], torch.zeros(3, dtype=torch.long), q[3:], x[3:], y[3:], z[3:],
                      torch.ones(4, dtype=torch.long), 3, 4, dtype=torch.float64)
    assert float(L) == 0.0


def test_chamfer_padding_point_and_tie():
    # a at x = 10, b at x = 30 (event 0).  Unpadded: 2 * 400 * wa * wb.  With the target padded,
    # a's nearest point is the padding point at the origin (d = 100 < 400), which has weight 0.
    args = (T(10.0), T(10.0), T(0.0), T(0.0), T(0), T(20.0), T(30.0), T(0.0), T(0.0), T(0))
    wa, wb = 1.0, 2.0
    assert float(OL.chamfer_3d(*args, 1, 1, dtype=torch.float64)) == pytest.approx(2 * 400 * wa * wb)
    assert float(OL.chamfer_3d(*args, 1, 2, dtype=torch.float64)) == pytest.approx(400 * wa * wb)
    # a padded simulated side does not change b's nearest (a at d = 400 beats the origin at 900)
    assert float(OL.chamfer_3d(*args, 2, 1, dtype=torch.float64)) == pytest.approx(2 * 400 * wa * wb)
    # a tie: a at 5 between b0 at 6 (w 1) and b1 at 4 (w 2): argmin takes b0 -> wa (1 + 1 + 2)
    L = OL.chamfer_3d(T(10.0), T(5.0), T(0.0), T(0.0), T(0), T(10.0, 20.0), T(6.0, 4.0), T(0.0, 0.0), T(0.0, 0.0),
                      T(0, 0), 1, 2, dtype=torch.float64)
    assert float(L) == pytest.approx(4.0)


def test_chamfer_x_rounding_of_later_events():
    # event 1: x + 1e9 in float32 is a multiple of 64 (1e9 itself is one), so x = 10 and x = 30 both
    # land on 1e9, while x = 40 lands on 1e9 + 64
    L1 = OL.chamfer_3d(T(10.0), T(10.0), T(0.0), T(0.0), T(1), T(10.0), T(30.0), T(0.0), T(0.0), T(1), 1, 1,
                       dtype=torch.float64)
    assert float(L1) == 0.0
    L2 = OL.chamfer_3d(T(10.0), T(10.0), T(0.0), T(0.0), T(1), T(10.0), T(40.0), T(0.0), T(0.0), T(1), 1, 1,
                       dtype=torch.float64)
    assert float(L2) == pytest.approx(2 * 64.0 ** 2)
    L0 = OL.chamfer_3d(T(10.0), T(10.0), T(0.0), T(0.0), T(0), T(10.0), T(30.0), T(0.0), T(0.0), T(0), 1, 1,
                       dtype=torch.float64)
    assert float(L0) == pytest.approx(2 * 400.0)


def test_nll_single_pair_closed_form():
    p, d2, qa, qb = 0.7, 0.25, 30.0, 50.0  # positions are float32 first: use exact values
    L = OL.nll_loss(T(qa), T(0.0), T(0.0), T(0.0), T(p), T(0), T(qb), T(0.5), T(0.0), T(0.0), T(1.0), T(0),
                    sigma=1.0, pad=3, dtype=torch.float64)
    dq = (qa - qb) / 10.0
    lam = p * math.exp(-(d2 + dq * dq) / 2)
    assert float(L) == pytest.approx(p - math.log(lam + 1e-10), rel=1e-12)
    # target hits with hit_prob <= 0.5 do not enter the likelihood
    L2 = OL.nll_loss(T(qa), T(0.0), T(0.0), T(0.0), T(p), T(0), T(qb), T(0.5), T(0.0), T(0.0), T(0.5), T(0),
                     dtype=torch.float64)
    assert float(L2) == pytest.approx(p)


@pytest.mark.parametrize("which", ["chamfer", "nll"])
def test_gradients_match_central_differences(which):
    g = torch.Generator().manual_seed(3)
    n, m = 9, 11
    ev, rev = torch.tensor([0] * 5 + [1] * 4), torch.tensor([0] * 6 + [1] * 5)
    x, y, z = (torch.rand(n, generator=g, dtype=torch.float64) * 4 for _ in range(3))
    rx, ry, rz = (torch.rand(m, generator=g, dtype=torch.float64) * 4 for _ in range(3))
    q, rq = torch.rand(n, generator=g, dtype=torch.float64) * 20 + 5, torch.rand(m, generator=g, dtype=torch.float64) * 20 + 5
    p, rp = torch.rand(n, generator=g, dtype=torch.float64), torch.rand(m, generator=g, dtype=torch.float64)
    # positions pass through float32 as in the reference: keep y, q and y +- h exact there
    y, q = torch.round(y * 1024) / 1024, torch.round(q * 1024) / 1024
    y.requires_grad_(True)
    q.requires_grad_(True)

    def f(yy, qq):
        if which == "chamfer":
            return OL.chamfer_3d(qq, x, yy, z, ev, rq, rx, ry, rz, rev, n, m, dtype=torch.float64)
        return OL.nll_loss(qq, x, yy, z, p, ev, rq, rx, ry, rz, rp, rev, sigma=1.5, dtype=torch.float64)

    L = f(y, q)
    gy, gq = torch.autograd.grad(L, (y, q))
    h = 2.0 ** -12
    for k in range(n):
        e = torch.zeros(n, dtype=torch.float64)
        e[k] = h
        with torch.no_grad():
            fy = (f(y + e, q) - f(y - e, q)) / (2 * h)
            fq = (f(y, q + e) - f(y, q - e)) / (2 * h)
        assert float(gy[k]) == pytest.approx(float(fy), rel=1e-5, abs=1e-8)
        assert float(gq[k]) == pytest.approx(float(fq), rel=1e-5, abs=1e-8)


def test_mse_time_closed_form():
    """losses_jax.py:82-87: ticks of empty ADC slots count as 0, then the per-pixel squared
    difference over the union of pixels."""
    from larndsim_amd.losses import mse_time

    adcs = torch.tensor([[80.0, 0.0], [90.0, 95.0]])
    ticks = torch.tensor([[10.0, 1998.0], [20.0, 30.0]])
    ref = torch.tensor([[85.0, 70.0], [0.0, 0.0]])
    ticks_ref = torch.tensor([[12.0, 40.0], [1998.0, 1998.0]])
    L, _ = mse_time(None, adcs, torch.tensor([5, 9]), ticks, ref, torch.tensor([5, 7]), ticks_ref)
    # pixel 5: (10 - 12)^2 + (0 - 40)^2; pixel 7: 0 (masked); pixel 9: 20^2 + 30^2
    assert float(L) == pytest.approx(4 + 1600 + 400 + 900)


def test_dense_mse_adc_closed_forms():
    """The dense mse_adc restatement: identical clouds give 0; one point each at distance d gives
    MMD 2 - 2 exp(-d^2 / 2 sigma^2) with equal charges, plus the relative charge error squared."""
    import math

    from oracle import losses as OL

    t = lambda *v: torch.tensor(v, dtype=torch.float64)  # noqa: E731
    x, y, z, Q, hp, ev = t(1.0, 2.0), t(0.5, -1.0), t(3.0, 3.5), t(10.0, 20.0), t(1.0, 1.0), t(0.0, 1.0)
    assert float(OL.mse_adc(Q, x, y, z, hp, ev, Q, x, y, z, hp, ev)) == pytest.approx(0.0, abs=1e-15)
    d = 0.7
    v = OL.mse_adc(t(5.0), t(0.0), t(0.0), t(0.0), t(1.0), t(0.0), t(5.0), t(d), t(0.0), t(0.0), t(1.0), t(0.0))
    assert float(v) == pytest.approx(2 - 2 * math.exp(-d * d / 2), rel=1e-12)
    v = OL.mse_adc(t(6.0), t(0.0), t(0.0), t(0.0), t(1.0), t(0.0), t(5.0), t(0.0), t(0.0), t(0.0), t(1.0), t(0.0))
    assert float(v) == pytest.approx((1.0 / (5.0 + 1e-6)) ** 2, rel=1e-9)
