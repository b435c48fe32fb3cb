"""The fit step (fit_params.py:168-199, 781-896) on the CPU: Adam's closed-form behaviour, global-norm
clipping, the sigmoid / exp-log normalisations, the staircase decay, and a full fitter loop on
a smooth stand-in simulation whose optimum is known."""
import math

import pytest
import torch

from larndsim_amd import fit


def test_adam_first_steps_closed_form():
    # a constant gradient g gives m_hat = g and v_hat = g^2 at every step: update = -lr g / (|g| + eps)
    opt = fit.Adam(lr=0.1)
    for _ in range(5):
        u = opt.update(3.0)
        assert u == pytest.approx(-0.1 * 3.0 / (3.0 + 1e-8), rel=1e-12)
    # sign flip after one step: m = 0.9*0.1*g + 0.1*(-g) = -0.01 g, v = ..., bias-corrected ratio is closed form
    opt = fit.Adam(lr=1.0)
    opt.update(2.0)
    u2 = opt.update(-2.0)
    m = 0.9 * (0.1 * 2.0) + 0.1 * (-2.0)
    v = 0.999 * (0.001 * 4.0) + 0.001 * 4.0
    assert u2 == pytest.approx(-(m / (1 - 0.81)) / (math.sqrt(v / (1 - 0.999 ** 2)) + 1e-8), rel=1e-12)


def test_clip_by_global_norm():
    g = {"a": 3.0, "b": 4.0}
    c = fit.clip_by_global_norm(g, 1.0)
    assert math.hypot(c["a"], c["b"]) == pytest.approx(1.0)
    assert c["a"] / c["b"] == pytest.approx(0.75)
    assert fit.clip_by_global_norm(g, 10.0) == g


def test_normalisation_round_trip():
    for key in ("Ab", "kb", "eField", "lifetime", "long_diff", "tran_diff"):
        r = fit.RANGES[key]
        for v in (r["min"] * 1.01, r["nom"], r["max"] * 0.99):
            n = fit.map_phys_to_norm(v, key)
            assert float(fit.map_norm_to_phys(n, key)) == pytest.approx(v, rel=1e-9)
        mid = float(fit.map_norm_to_phys(torch.tensor(0.0, dtype=torch.float64), key))
        assert mid == pytest.approx((r["min"] + r["max"]) / 2)
        n = fit.map_phys_to_norm(r["nom"], key, scheme="exp_log")
        assert float(n) == pytest.approx(0.0, abs=1e-12)
    with pytest.raises(ValueError):
        fit.map_norm_to_phys(torch.tensor(0.0), "Ab", scheme="nope")


def test_exponential_decay_staircase():
    f = fit.exponential_decay(0.1, transition_steps=10, decay_rate=0.5, staircase=True)
    assert [f(c) for c in (0, 9, 10, 25)] == [0.1, 0.1, 0.05, 0.025]
    g = fit.exponential_decay(0.1, transition_steps=10, decay_rate=0.5)
    assert g(5) == pytest.approx(0.1 * 0.5 ** 0.5)


class _LinearSim:
    """Stand-in strategy: 'adcs' linear in the parameters (differentiable through torch)."""

    def predict(self, params, tracks, fields, rngkey):
        w = torch.tensor([100.0, -50.0, 30.0], dtype=torch.float64)
        x = torch.stack([torch.as_tensor(params.Ab, dtype=torch.float64) * 10,
                         torch.as_tensor(params.kb, dtype=torch.float64) * 100,
                         torch.as_tensor(params.eField, dtype=torch.float64) * 10])
        return {"adcs": w * x}


class _SquaredLoss:
    def compute(self, params, pred, target):
        return ((pred["adcs"] - target["adcs"]) ** 2).sum(), {}


def test_fitter_recovers_a_smooth_optimum(params4):
    sim = _LinearSim()
    target_p = params4.replace(Ab=0.82, kb=0.05, eField=0.52)
    with torch.no_grad():
        target = sim.predict(target_p, None, None, None)
    f = fit.GradientDescentFitter(sim, _SquaredLoss(), params4, ["Ab", "kb", "eField"], lr=0.05, max_clip_norm_val=1.0)
    for _ in range(400):
        f.step(None, None, target)
    phys = f.physical()
    for k in ("Ab", "kb", "eField"):
        assert phys[k] == pytest.approx(getattr(target_p, k), rel=2e-3), k
    assert f.history["loss"][-1] < 1e-3 * f.history["loss"][0]


def test_blockwise_mmd_equals_full_matrix():
    """Event-blocked MMD (no kernel matrices kept) against the reference formula on the
    event-offset clouds, value and gradients (x, px); kernels in float32, sums in float64."""
    from larndsim_amd import losses

    g = torch.Generator().manual_seed(0)
    nx, ny = 70, 55
    ex = torch.randint(0, 3, (nx,), generator=g)
    ey = torch.randint(0, 3, (ny,), generator=g)
    x = torch.randn(nx, 3, generator=g, dtype=torch.float64) * 2
    y = torch.randn(ny, 3, generator=g, dtype=torch.float64) * 2
    px = torch.rand(nx, generator=g, dtype=torch.float64)
    py = torch.rand(ny, generator=g, dtype=torch.float64)
    x1, p1 = x.clone().requires_grad_(True), px.clone().requires_grad_(True)
    off = lambda c, e: c + torch.stack([e * 1e5, 0 * e, 0 * e], dim=-1).to(c.dtype)  # noqa: E731
    full = losses.mmd(off(x1, ex), off(y, ey), p1, py, 1.3)
    full.backward()
    x2, p2 = x.clone().requires_grad_(True), px.clone().requires_grad_(True)
    blk = losses.mmd_blockwise(x2, y, p2, py, 1.3, ex, ey, chunk=16)
    blk.backward()
    # kernel values are evaluated in float32 (sums in float64)
    assert float(blk) == pytest.approx(float(full), rel=1e-6, abs=1e-9)
    torch.testing.assert_close(x2.grad, x1.grad, rtol=1e-5, atol=1e-8)
    torch.testing.assert_close(p2.grad, p1.grad, rtol=1e-5, atol=1e-8)
