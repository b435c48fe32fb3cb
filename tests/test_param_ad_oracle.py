"""Pins of the torch-float64 autograd checker of the parametrized model (oracle/param_ad.py, CPU
only): its forward equals the NumPy oracle of the model, and its gradient equals central finite
differences of the NumPy oracle itself, with every decision free (pixel ids, whole ticks and the
readout's decisions are checked unchanged over the difference interval)."""
import numpy as np
import pytest
import torch

from oracle import param as PA
from oracle import param_ad as PD

CASES = [(True, 1, False), (True, 0, True), (False, 0, True)]
IDS = ["mc-n1-nodiff", "mc-n0", "bins-n0"]


def _setup(params4, c1, mc, nn, diffusion, n_seg=150):
    tr, f = c1
    tr = np.ascontiguousarray(tr[:n_seg])
    p = params4.replace(mc_diff=mc, number_pix_neighbors=nn, diffusion_in_current_sim=diffusion)
    normals = np.random.default_rng(5).normal(size=(len(tr), 3)).astype(np.float32).astype(np.float64)
    edges = np.linspace(-0.2, 0.2, 6)
    return p, tr, f, (normals if mc else None), (None if mc else edges)


@pytest.mark.parametrize("mc,nn,diffusion", CASES, ids=IDS)
def test_forward_matches_numpy_oracle(params4, c1, mc, nn, diffusion):
    p, tr, f, normals, edges = _setup(params4, c1, mc, nn, diffusion)
    fr = PD.FrozenParam(p, tr, f, normals=normals, bin_edges=edges)
    with torch.no_grad():
        adc, z, wfs = fr.forward(fr.theta0())
    o = fr.out
    assert np.isfinite(o["wfs"]).all()
    scale = np.abs(o["wfs"]).max()
    np.testing.assert_allclose(wfs.numpy(), o["wfs"][:, 1:], rtol=0, atol=1e-11 * scale)
    vr = fr.valid_rows
    np.testing.assert_allclose(adc.numpy()[vr], o["grid"]["adcs"][vr], rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(z.numpy()[vr], o["grid"]["z"][vr], rtol=1e-12, atol=1e-12)


def _run(p, tr, f, normals, edges, ref):
    with np.errstate(all="ignore"):
        o = PA.simulate_parametrized(p, tr, f, normals=normals, bin_edges=edges)
    assert np.array_equal(o["pair_pids"], ref["pair_pids"]) and np.array_equal(o["unique_pixels"], ref["unique_pixels"])
    return o


def _flipped(o, ref):
    """Rows whose readout decisions differ from the reference run's, and rows reached by an electron
    whose whole-tick arrival time differs (the current then moves by a tick: a step in the waveform)."""
    g, gr = o["grid"], ref["grid"]
    d = (g["ticks"] != gr["ticks"]).any(axis=1)
    for k in ("e", "g", "valid"):
        d |= (g["dec"][k] != gr["dec"][k]).any(axis=1)
    moved = o["t0_tick"] != ref["t0_tick"]
    d[np.searchsorted(ref["unique_pixels"], ref["pair_pids"][moved].ravel())] = True
    return d


def _value(o, w_adc, w_z, w_wfs):
    g = o["grid"]
    vr = o["unique_pixels"] >= 0
    return float((w_adc * g["adcs"][vr]).sum() + (w_z * g["z"][vr]).sum() + (w_wfs * o["wfs"][:, 1:]).sum())


@pytest.mark.parametrize("mc,nn,diffusion", CASES, ids=IDS)
def test_gradient_matches_central_differences(params4, c1, mc, nn, diffusion):
    """A row whose readout decision or an electron's whole tick flips within the difference
    interval gets no cotangent."""
    p, tr, f, normals, edges = _setup(params4, c1, mc, nn, diffusion)
    fr = PD.FrozenParam(p, tr, f, normals=normals, bin_edges=edges)
    names = PD.GRAD_PARAMS
    runs, flipped = {}, np.zeros(len(fr.U), bool)
    for n in names:
        v0 = float(getattr(p, n))
        h = 1e-6 * abs(v0)
        runs[n] = (h, [_run(p.replace(**{n: v0 + s * h}), tr, f, normals, edges, fr.out) for s in (1, -1)])
        for o in runs[n][1]:
            flipped |= _flipped(o, fr.out)
    assert flipped.mean() < 0.05
    vr = fr.valid_rows
    hits = (fr.ticks[vr] < fr.nt - 4) & ~flipped[vr][:, None]
    assert hits.sum() > 5
    rng = np.random.default_rng(11)
    w_adc = rng.normal(size=hits.shape) * hits
    w_z = rng.normal(size=hits.shape) * hits
    # a waveform cotangent of the scale of the hit terms' (1 / t_sampling per ADC count and tick)
    w_wfs = rng.normal(size=(len(fr.U), fr.nt - 1)) * 1e-3 * ~flipped[:, None]
    g, _ = fr.vjp(w_adc, w_z, w_wfs)
    for i, n in enumerate(names):
        h, (op, om) = runs[n]
        fd = (_value(op, w_adc, w_z, w_wfs) - _value(om, w_adc, w_z, w_wfs)) / (2 * h)
        if n == "vdrift":  # never read by the simulation
            assert g[i] == 0.0 and fd == 0.0
            continue
        assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-9 * abs(fd) + 1e-12), n
