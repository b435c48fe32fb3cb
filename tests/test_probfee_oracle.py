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
