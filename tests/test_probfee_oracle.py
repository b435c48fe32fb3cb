"""The probabilistic readout restatement (oracle/probfee.py) against closed forms and the
deterministic self-trigger: no GPU."""
import numpy as np
import pytest
from scipy.special import ndtr

from oracle import oracle as O
from oracle import probfee as PF


def test_log_diff_ndtr_away_from_the_clamps():
    rng = np.random.default_rng(0)
    b = rng.uniform(-6, 4, 500)
    a = b + rng.uniform(0.05, 5, 500)
    want = np.log(ndtr(a) - ndtr(b))
    got = PF.log_diff_ndtr(a, b)
    ok = want > -90  # the log_term clamp at -100 is smooth but not negligible below
    assert np.allclose(got[ok], want[ok], rtol=1e-9, atol=1e-9)
    # a <= b is pushed to -1000
    assert np.all(PF.log_diff_ndtr(b - 0.1, b) < -999)


def _pulses(T, at, charge, width=3):
    w = np.zeros(T, np.float32)
    for t, q in zip(at, charge):
        w[t:t + width] += q / width / 0.1  # current per tick; charge = current * t_sampling
    return w


def test_single_clean_pulse():
    T, theta, sigma = 400, 7000.0, 300.0
    w = _pulses(T, [100], [40000.0])[None, :]
    lp, q, _ = PF.prob_fee(w, 0.1, 20, 3, 18, theta, sigma)
    p0 = np.exp(lp[0, 0])
    assert p0.sum() == pytest.approx(1.0, abs=1e-3)
    assert int(np.argmax(p0)) == 99  # the tick before the first cumulative charge above theta
    assert np.exp(lp[0, 1]).sum() < 1e-6  # no second hit


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_small_noise_reproduces_deterministic_ticks(seed):
    from larndsim_amd import consts

    p = consts.default_params(RESET_NOISE_CHARGE=0.0)
    rng = np.random.default_rng(seed)
    T = 2000
    at = np.sort(rng.choice(np.arange(50, 1900, 60), size=4, replace=False))
    w = _pulses(T, at, rng.uniform(12000, 60000, size=4), width=int(rng.integers(1, 6)))[None, :]
    w += rng.uniform(0, 50, size=(1, T)).astype(np.float32)  # a small pedestal current
    _, ticks, dec = O.get_adc_values(p, w.astype(np.float64))
    lp, q, _ = PF.prob_fee(w, p.t_sampling, int(p.fee_paths_scaling), int(p.MAX_ADC_VALUES), p.adc_interval,
                           p.DISCRIMINATION_THRESHOLD, 30.0)
    nvalid = int(dec["valid"][0].sum())
    assert nvalid >= 3
    for s in range(nvalid):
        prob = np.exp(lp[0, s])
        assert prob.sum() == pytest.approx(1.0, abs=1e-2), s
        assert int(np.argmax(prob)) == int(ticks[0, s]), s


def _noisy_trains(P, T, seed):
    """Pulse trains with a pedestal current and noise: no flat stretches of the cumulative charge
    (ties of the running maximum would make the frozen-decision function non-smooth)."""
    rng = np.random.default_rng(seed)
    w = np.zeros((P, T), np.float32)
    for p in range(P):
        at = np.sort(rng.choice(np.arange(20, T - 60, 25), size=3, replace=False))
        w[p] = _pulses(T, at, rng.uniform(6000, 30000, size=3), width=int(rng.integers(2, 6)))
    return w + rng.uniform(5, 60, size=(P, T)).astype(np.float32)


def test_torch_replay_equals_the_numpy_restatement():
    """prob_fee_torch with the numpy run's selections reproduces its outputs (float64)."""
    import torch

    T, V, S, theta, sigma = 260, 5, 4, 7000.0, 900.0
    w = _noisy_trains(3, T, 4)
    lp, q, _, tops, n = PF.prob_fee_tape(w, 0.1, V, S, 18, theta, sigma)
    tl, tq = PF.prob_fee_torch(torch.from_numpy(w.astype(np.float64)), 0.1, V, S, 18, theta, sigma, tops, n)
    np.testing.assert_allclose(tl.numpy(), lp, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(tq.numpy(), q, rtol=1e-12, atol=1e-9)


def test_torch_replay_gradient_is_central_differences():
    """The autograd gradient of a random projection of both outputs against central differences
    of the same frozen-selection function along random directions of the waveforms."""
    import torch

    T, V, S, theta, sigma = 200, 4, 3, 7000.0, 900.0
    w0 = torch.from_numpy(_noisy_trains(2, T, 7).astype(np.float64))
    _, _, _, tops, n = PF.prob_fee_tape(w0.numpy().astype(np.float32), 0.1, V, S, 18, theta, sigma)
    rng = np.random.default_rng(8)
    g_lp = torch.from_numpy(rng.normal(size=(2, S, T - 1)))
    g_q = torch.from_numpy(rng.normal(size=(2, S, T - 1)) * 1e-4)

    def f(w):
        lp, q = PF.prob_fee_torch(w, 0.1, V, S, 18, theta, sigma, tops, n, q_f32=False)
        # only probabilities that matter (log prob above -50) carry a cotangent
        return ((g_lp * torch.exp(lp)).sum() + (g_q * q).sum())

    w = w0.clone().requires_grad_(True)
    f(w).backward()
    g = w.grad.numpy()
    for k in range(3):
        u = torch.from_numpy(rng.normal(size=w0.shape))
        h = 1e-3
        fd = (float(f(w0 + h * u)) - float(f(w0 - h * u))) / (2 * h)
        an = float((torch.from_numpy(g) * u).sum())
        assert an == pytest.approx(fd, rel=1e-5, abs=1e-9), (k, an, fd)
