"""The parametrized current model on the GPU (lsim_param_electrons, lsim_param_rows, float64)
against the float64 oracle (oracle/param.py, simulate_parametrized, sim_jax.py:339-372): the
electrons to 1e-12, the pair pixel ids and the sorted pixel set exactly, the waveforms to 1e-6 of
each row's peak (float32 output), and the compacted hits through the shared readout.

A decay surface of the model crossing zero sends 1/scale to infinity, and those few
nearest-neighbour pairs overflow to NaN in float64 as in the reference (which overflows more often
in float32): such a row is NaN in both, and yields no hits in both."""
import types

import numpy as np
import pytest
import torch

from helpers import compare_hits

pytestmark = pytest.mark.gpu


CASES = [  # (mc_diff, number_pix_neighbors, diffusion in the current)
    (True, 0, True),
    (True, 2, True),  # 22 of 1243 rows meet an overflowing pair, the same rows on both sides
    (False, 0, True),  # transverse bins: 25 electrons per segment (the central pixel never overflows)
    (True, 1, False),  # no diffusion: one-sided exponentials, finite everywhere
    (True, 4, True),  # the neighbourhood of the benchmark (81 pixels per electron)
]


@pytest.mark.parametrize("mc,nn,diffusion", CASES, ids=["mc-n0", "mc-n2", "bins-n0", "mc-n1-nodiff", "mc-n4"])
def test_parametrized_matches_oracle(params4, c1, mc, nn, diffusion):
    from larndsim_amd import param as GP
    from larndsim_amd.sim import get_simulator
    from oracle import oracle as O
    from oracle import param as PA

    tr, f = c1
    tr = np.ascontiguousarray(tr[:400])
    p = params4.replace(mc_diff=mc, number_pix_neighbors=nn, diffusion_in_current_sim=diffusion)
    edges = np.linspace(-0.2, 0.2, 6)
    normals = np.random.default_rng(5).normal(size=(len(tr), 3)).astype(np.float32)
    print(f"[param] mc={mc} n={nn} diffusion={diffusion}", flush=True)
    e = GP.electrons(p, torch.from_numpy(tr).cuda(), f, normals=torch.from_numpy(normals).cuda(), bin_edges=edges)
    with np.errstate(all="ignore"):
        o = PA.simulate_parametrized(p, tr, f, normals=normals.astype(np.float64), bin_edges=edges)
    oe = o["electrons"]
    v = O.vdrift(p)
    z_an = O._borders(p)[oe["plane"], 2, 0]
    want = dict(x=oe["x"], y=oe["y"], ne=oe["ne"], t0=np.abs(oe["z"] - z_an) / v, sig=oe["sl"] / v)
    for k, w in want.items():
        np.testing.assert_allclose(e[k].cpu().numpy(), w, rtol=1e-12, atol=1e-12, err_msg=k)
    assert np.array_equal(e["plane"].cpu().numpy(), oe["plane"])
    assert np.array_equal(e["pair_pids"].cpu().numpy(), o["pair_pids"])

    wfs, U = GP.waveforms(p, e)
    U = U.cpu().numpy()
    assert np.array_equal(U, o["unique_pixels"])
    g = wfs.cpu().numpy().astype(np.float64)
    ow = o["wfs"][:, 1:]
    valid = U >= 0
    nan_g, nan_o = np.isnan(g[valid]).any(axis=1), np.isnan(ow[valid]).any(axis=1)
    assert np.array_equal(nan_g, nan_o)
    print(f"[param] {valid.sum()} rows, {nan_o.sum()} with an overflowing pair", flush=True)
    assert nan_o.mean() < 0.2
    gv, ov = g[valid][~nan_o], ow[valid][~nan_o]
    peak = np.abs(ov).max(axis=1, keepdims=True) + 1e-30
    err = (np.abs(gv - ov) / peak).max()
    print(f"[param] waveform error / row peak {err:.2e}", flush=True)
    assert err < 1e-6

    # the readout of the valid rows against the oracle's hits
    sim = get_simulator(p, None)
    out = sim.fee(p, wfs[torch.from_numpy(valid).cuda()], torch.from_numpy(U[valid]).cuda())
    n = compare_hits(types.SimpleNamespace(out=out, max_rows=int(valid.sum())), o, f"param-{mc}-{nn}-{diffusion}")
    assert n > 0


def test_strategy_and_reference_outputs(params4, c1):
    """HipParametrizedSimulation.predict returns the reference's keys; the hits equal those of
    simulate_parametrized with the same seed, and a different seed moves the electrons."""
    from larndsim_amd import param as GP
    from larndsim_amd.strategies import HipParametrizedSimulation

    tr, f = c1
    trd = torch.from_numpy(np.ascontiguousarray(tr[:400])).cuda()
    p = params4.replace(mc_diff=True, number_pix_neighbors=0)
    out = HipParametrizedSimulation().predict(p, trd, f, rngkey=7)
    assert set(out) == {"adcs", "pixel_x", "pixel_y", "pixel_z", "ticks", "hit_prob", "event", "unique_pixels"}
    again = GP.simulate_parametrized(p, trd, f, rngseed=7)
    assert torch.equal(out["adcs"], again[0]) and torch.equal(out["unique_pixels"], again[7])
    e7 = GP.electrons(p, trd, f, rngseed=7)
    e8 = GP.electrons(p, trd, f, rngseed=8)
    assert not torch.equal(e7["x"], e8["x"])
    # parameters that require gradients: the same hits, now with autograd history
    g = HipParametrizedSimulation().predict(p.replace(Ab=torch.tensor(0.8, dtype=torch.float64, requires_grad=True)),
                                            trd, f, rngkey=7)
    assert g["adcs"].grad_fn is not None and torch.equal(g["adcs"].detach(), out["adcs"])
