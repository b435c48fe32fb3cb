"""The probabilistic readout's adjoint on the GPU (lsim_prob_fee_backward, float32) against
torch-float64 autograd of the oracle's frozen-selection replay of the reference's beam search
(oracle/probfee.py prob_fee_torch, fee_jax.py:334-461 with top_k under stop_gradient), and the
strategy's chain through the waveform adjoint of the LUT simulation.

Loss: sum(G exp(log_prob)) + sum(Gq charge) with random G, Gq, the shape of the reference's
probabilistic losses (they act on probabilities, so a cotangent on a log probability comes with its
probability as a factor).  Rows whose beam search meets a near tie at the top-V boundary among
paths that carry probability are left out, as in the forward test: there float32 and float64 may
keep different paths."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noisy_trains(P, T, seed):
    """Pulse trains over a pedestal current with noise (no flat stretches of the cumulative charge)."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(5, 60, size=(P, T)).astype(np.float32)
    for p in range(P):
        for t in np.sort(rng.choice(np.arange(40, T - 80, 37), size=int(rng.integers(1, 5)), replace=False)):
            wd = int(rng.integers(2, 7))
            w[p, t:t + wd] += np.float32(rng.uniform(6000, 50000) / wd / 0.1)
    return w


def _tie_free_rows(sels, n_steps, V):
    keep = []
    for r in range(sels.shape[0]):
        ok = True
        for s in range(n_steps):
            srt = np.sort(sels[r, s])[::-1]
            if np.isfinite(srt[V]) and srt[V - 1] > -20.0 and srt[V - 1] - srt[V] < 1e-3 * max(1.0, abs(srt[V - 1])):
                ok = False
                break
        keep.append(ok)
    return np.array(keep)


def _check(params, w, seed):
    from larndsim_amd.fee import get_adc_values_average_noise_vmap
    from oracle import probfee as PF

    P, T = w.shape
    V, S = int(params.fee_paths_scaling), int(params.MAX_ADC_VALUES)
    args = (params.t_sampling, V, S, params.adc_interval, params.DISCRIMINATION_THRESHOLD, params.RESET_NOISE_CHARGE)
    _, _, sels, tops, n = PF.prob_fee_tape(w, *args)
    keep = _tie_free_rows(sels, n, V)
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(P, S, T - 1))
    Gq = rng.normal(size=(P, S, T - 1)) * 1e-4

    def torch_grad(dt):
        wt = torch.from_numpy(w).to(dt).requires_grad_(True)
        tl, tq = PF.prob_fee_torch(wt, *args, tops, n)
        ((torch.from_numpy(G).to(dt) * torch.exp(tl)).sum() + (torch.from_numpy(Gq).to(dt) * tq).sum()).backward()
        return wt.grad.double().numpy()

    g_ref = torch_grad(torch.float64)
    g_32 = torch_grad(torch.float32)  # the same function evaluated in float32 on the CPU: the rounding scale

    wg = torch.from_numpy(w).cuda().requires_grad_(True)
    lp, q = get_adc_values_average_noise_vmap(params, wg)
    assert lp.requires_grad and q.requires_grad
    Gg, Gqg = torch.from_numpy(G).float().cuda(), torch.from_numpy(Gq).float().cuda()
    ((Gg * torch.exp(lp)).sum() + (Gqg * q).sum()).backward()
    g = wg.grad.double().cpu().numpy()

    print(f"[pf-adj] rows {P}, steps {n}, tie-free rows {int(keep.sum())}", flush=True)
    assert keep.sum() >= P // 2
    worst = 0.0
    for r in np.nonzero(keep)[0]:
        scale = np.abs(g_ref[r]).max()
        err = np.abs(g[r] - g_ref[r]).max() / scale
        err32 = np.abs(g_32[r] - g_ref[r]).max() / scale
        u = rng.normal(size=T)
        proj = abs(float(g[r] @ u) - float(g_ref[r] @ u)) / (np.abs(g_ref[r]) @ np.abs(u))
        at = int(np.argmax(np.abs(g[r] - g_ref[r])))
        print(f"[pf-adj] row {r}: |g| max {scale:.3e}, max error / max {err:.2e} at tick {at} "
              f"(torch float32: {err32:.2e}), projection error {proj:.2e}", flush=True)
        worst = max(worst, err)
        assert scale > 0
        # float32 rounding of the steep soft masks (sigmoid of 1000 dz) moves single ticks by up to a
        # few 1e-3 of the largest component; any projection of the whole gradient stays within 1e-4
        assert err < 5e-3, r
        assert proj < 1e-4, r
    return worst


@pytest.mark.parametrize("sigma", [900.0, 300.0])
def test_adjoint_on_pulse_trains(params4, sigma):
    params = params4.replace(RESET_NOISE_CHARGE=sigma)
    w = _noisy_trains(8, 2000, int(sigma))
    print("worst", _check(params, w, 11), flush=True)


def test_adjoint_on_simulated_waveforms(params4, lut, c1):
    from larndsim_amd.sim import Simulator

    sim = Simulator(params4, lut)
    tracks, fields = c1
    st = sim.forward(params4, tracks, fields, n_events=1, want_wfs=True)
    wfs = st.out["wfs"][: st.max_rows].cpu().numpy()
    pick = np.argsort(-wfs.sum(axis=1))[:8]
    params = params4.replace(RESET_NOISE_CHARGE=900.0)
    print("worst", _check(params, np.ascontiguousarray(wfs[pick]), 12), flush=True)


def test_adjoint_of_unused_steps_is_zero(params4):
    """Rows without charge stop after the first step; cotangents on the later (constant) steps and
    on an all-zero row give a zero gradient."""
    from larndsim_amd.fee import get_adc_values_average_noise_vmap

    params = params4.replace(RESET_NOISE_CHARGE=900.0)
    w = torch.zeros((3, 2000), device="cuda", requires_grad=True)
    lp, q = get_adc_values_average_noise_vmap(params, w)
    (lp[:, 1:].sum() + q[:, 1:].sum()).backward()
    assert torch.all(w.grad == 0)


def test_strategy_gradient_chains_through_the_waveforms(params4, lut, c1):
    """HipLUTProbabilisticSimulation with parameters that require gradients: the parameter gradient
    of a loss on hit_prob and adcs_distrib equals the waveform adjoint of the LUT simulation applied
    to the readout adjoint's waveform gradient."""
    from larndsim_amd.fee import get_adc_values_average_noise_vmap, digitize
    from larndsim_amd.sim import Simulator
    from larndsim_amd.strategies import HipLUTProbabilisticSimulation, pad_to_closest_multiple

    tr, f = c1
    names = ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff")
    p = params4.replace(RESET_NOISE_CHARGE=900.0)
    th = {n: torch.tensor(float(getattr(p, n)), dtype=torch.float64, requires_grad=True) for n in names}
    out = HipLUTProbabilisticSimulation(lut).predict(p.replace(**th), torch.from_numpy(tr).cuda(), f)
    Pp = out["wfs"].shape[0]
    gen = torch.Generator(device="cuda").manual_seed(3)
    G = torch.randn(out["hit_prob"].shape, device="cuda", generator=gen)
    Ga = torch.randn(out["adcs_distrib"].shape, device="cuda", generator=gen) * 1e-2
    ((G * torch.exp(out["hit_prob"])).sum() + (Ga * out["adcs_distrib"]).sum()).backward()
    g = np.array([float(th[n].grad) for n in names])

    quiet = params4.replace(RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)
    sim = Simulator(quiet, lut)
    st = sim.forward(quiet, tr, f, want_tape=True, want_wfs=True)
    P = st.max_rows
    rows = st.out["wfs"][:P].detach().clone().requires_grad_(True)
    w = pad_to_closest_multiple(torch.cat([torch.zeros((Pp - P, rows.shape[1]), device="cuda"), rows]), dims_to_pad=(0,),
                                multiple=128, pad_front=True)
    assert w.shape[0] == Pp
    lp, q = get_adc_values_average_noise_vmap(p, w)
    ((G * torch.exp(lp)).sum() + (Ga * digitize(p, q)).sum()).backward()
    g2 = sim.backward(st, torch.zeros(P * sim.H), None, rows.grad).cpu().numpy()
    print("strategy", g, "direct", g2, flush=True)
    assert np.abs(g).max() > 0
    assert np.allclose(g, g2, rtol=1e-6, atol=0)
