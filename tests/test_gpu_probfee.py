"""The probabilistic readout on the GPU (lsim_prob_fee, float32) against the float64 restatement
of the reference's beam search (oracle/probfee.py) on hand-built pulse trains and on simulated
waveforms: expected charges to 2e-6 of their scale; hit probabilities within twice the error of
the same function evaluated in float32 (torch on the CPU, the oracle's path selections) plus 2e-5;
each hit's total probability to 1e-4 + 0.1 %.  Later hit indices are compared on rows whose top-V
selections were not near ties."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pulse_rows(seed, n, T=2000):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, T), np.float32)
    for r in range(n):
        k = int(rng.integers(1, 6))
        for t, q in zip(np.sort(rng.choice(np.arange(30, 1950, 45), size=k, replace=False)),
                        rng.uniform(5000, 60000, size=k)):
            wd = int(rng.integers(1, 8))
            rows[r, t:t + wd] += q / wd / 0.1
        rows[r] += rng.uniform(0, 80, size=T).astype(np.float32)
    return rows


def _compare(params, w, stop=1e-9):
    from larndsim_amd.fee import get_adc_values_average_noise_vmap
    from oracle import probfee as PF

    lp, q = get_adc_values_average_noise_vmap(params, torch.from_numpy(w).cuda(), stop)
    lp, q = lp.cpu().numpy().astype(np.float64), q.cpu().numpy().astype(np.float64)
    args = (params.t_sampling, int(params.fee_paths_scaling), int(params.MAX_ADC_VALUES), params.adc_interval,
            params.DISCRIMINATION_THRESHOLD, params.RESET_NOISE_CHARGE)
    rl, rq, sels, tops, n_steps = PF.prob_fee_tape(w, *args, stop)
    # the same function in float32 (torch on the CPU, the oracle's selections): the rounding scale
    l32, _ = PF.prob_fee_torch(torch.from_numpy(w), *args, tops, n_steps)
    p32 = np.exp(l32.double().numpy())
    V = int(params.fee_paths_scaling)
    scale = np.abs(rq).max() + 1.0
    compared = 0
    worst_g, worst_32 = 0.0, 0.0
    for r in range(w.shape[0]):
        for s in range(lp.shape[1]):
            assert np.allclose(q[r, s], rq[r, s], rtol=0, atol=2e-6 * scale), (r, s)
            # probabilities, not log probabilities: the reference's soft mask of log_diff_ndtr
            # (sigmoid of 1000 (a - b)) turns float32 rounding of the z differences of nearly flat
            # ticks into log-probability changes of ~1e-2 where the probability is small
            pg, pr = np.exp(lp[r, s]), np.exp(rl[r, s])
            eg, e32 = np.abs(pg - pr).max(), np.abs(p32[r, s] - pr).max()
            worst_g, worst_32 = max(worst_g, eg), max(worst_32, e32)
            assert eg <= 2 * e32 + 2e-5, (r, s, eg, e32)
            assert abs(pg.sum() - pr.sum()) < 1e-4 + 1e-3 * pr.sum(), (r, s)
            compared += 1
            srt = np.sort(sels[r, s])[::-1]
            if np.isfinite(srt[V]) and srt[V - 1] > -20.0 and srt[V - 1] - srt[V] < 1e-3 * max(1.0, abs(srt[V - 1])):
                break  # near tie among paths that carry probability: later steps may follow others
    print(f"[probfee] worst probability error: GPU {worst_g:.2e}, torch float32 {worst_32:.2e}", flush=True)
    return compared


@pytest.mark.parametrize("sigma", [900.0, 60.0])
def test_pulse_trains(sigma):
    from larndsim_amd import consts

    params = consts.default_params(RESET_NOISE_CHARGE=sigma)
    w = _pulse_rows(int(sigma), 12)
    n = _compare(params, w)
    print("compared steps", n, flush=True)
    assert n >= 12 * 3  # on average at least three hit indices per row


def test_simulated_waveforms(params4, lut, c1):
    from larndsim_amd.sim import Simulator

    sim = Simulator(params4, lut)
    tracks, fields = c1
    st = sim.forward(params4, tracks, fields, n_events=1, want_wfs=True)
    wfs = st.out["wfs"][: st.max_rows].cpu().numpy()
    # rows with the most charge
    pick = np.argsort(-wfs.sum(axis=1))[:10]
    params = params4.replace(RESET_NOISE_CHARGE=900.0)
    n = _compare(params, np.ascontiguousarray(wfs[pick]))
    print("compared steps", n, flush=True)
    assert n >= 10


def test_probabilistic_outputs_and_stop(params4):
    """simulate_probabilistic's shapes and coordinates; a row set with no charge stops after the
    first step (every later step -1000 / 0)."""
    from larndsim_amd import detsim
    from larndsim_amd.fee import get_adc_values_average_noise_vmap, simulate_probabilistic

    params = params4.replace(RESET_NOISE_CHARGE=900.0)
    w = torch.zeros((3, 2000), device="cuda")
    lp, q = get_adc_values_average_noise_vmap(params, w)
    assert tuple(lp.shape) == (3, 10, 1999)
    assert torch.all(lp[:, 1:] == -1000.0) and torch.all(q[:, 1:] == 0.0)
    pids = detsim.pixel2id(params, torch.tensor([3, 10, 20]), torch.tensor([4, 5, 6]), torch.tensor([0, 1, 0]),
                           torch.tensor([0, 0, 2]))
    adcs, px, py, tp, ev = simulate_probabilistic(params, torch.from_numpy(_pulse_rows(3, 3)).cuda(), pids.cuda())
    assert tuple(adcs.shape) == (3, 10, 1999) and tuple(tp.shape) == (3, 10, 1999)
    assert ev.tolist() == [0, 0, 2]
    xy = detsim.get_pixel_coordinates(params, torch.tensor([3, 10, 20]), torch.tensor([4, 5, 6]),
                                      torch.tensor([0, 1, 0]))
    assert torch.allclose(px.cpu().double(), xy[:, 0].double()) and torch.allclose(py.cpu().double(), xy[:, 1].double())


def test_probabilistic_strategy(params4, lut, c1):
    """HipLUTProbabilisticSimulation (strategies.py:107-134): rows padded at the front to a multiple
    of 128 with pixel id -1, the readout of those rows, digitised expected charges."""
    from larndsim_amd.fee import digitize, get_adc_values_average_noise_vmap
    from larndsim_amd.strategies import HipLUTProbabilisticSimulation

    tracks, fields = c1
    p = params4.replace(RESET_NOISE_CHARGE=900.0)
    out = HipLUTProbabilisticSimulation(lut).predict(p, tracks, fields)
    P = out["wfs"].shape[0]
    print("probabilistic strategy rows", P, flush=True)
    assert P % 128 == 0 and tuple(out["hit_prob"].shape) == (P, 10, 1999)
    ids = out["unique_pixels"]
    npad = int((ids == -1).sum())
    assert npad > 0 and bool((ids[:npad] == -1).all()) and bool((out["wfs"][:npad] == 0).all())
    lp, q = get_adc_values_average_noise_vmap(p, out["wfs"])
    assert torch.equal(lp, out["hit_prob"])
    assert torch.equal(digitize(p, q), out["adcs_distrib"])
