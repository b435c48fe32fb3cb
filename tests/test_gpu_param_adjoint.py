"""Adjoint of the parametrized current model on the GPU (lsim_param_backward) against torch-float64
autograd of the frozen-decision chain (oracle/param_ad.py, itself pinned by central differences of
the NumPy oracle): cotangents on the hit ADC values, the hit z positions and every waveform tick.

Tolerance: relative error <= 1e-3 per parameter (the GPU waveforms and cotangents are float32);
the vdrift gradient is exactly zero (the parameter is never read by the simulation)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3

CASES = [(True, 1, False), (True, 0, True), (False, 0, True), (True, 2, False)]
IDS = ["mc-n1-nodiff", "mc-n0", "bins-n0", "mc-n2-nodiff"]


def _setup(params4, c1, mc, nn, diffusion, n_seg=400):
    tr, f = c1
    tr = np.ascontiguousarray(tr[:n_seg])
    p = params4.replace(mc_diff=mc, number_pix_neighbors=nn, diffusion_in_current_sim=diffusion)
    normals = np.random.default_rng(5).normal(size=(len(tr), 3)).astype(np.float32)
    edges = np.linspace(-0.2, 0.2, 6)
    return p, tr, f, normals, edges


@pytest.mark.parametrize("dense", [False, True], ids=["hits", "waveforms"])
@pytest.mark.parametrize("mc,nn,diffusion", CASES, ids=IDS)
def test_param_backward_matches_autograd(params4, c1, mc, nn, diffusion, dense):
    from larndsim_amd.autograd import theta_of
    from larndsim_amd.param import ParamSimFunction
    from oracle import param_ad as PD

    p, tr, f, normals, edges = _setup(params4, c1, mc, nn, diffusion)
    print(f"[param-adj] mc={mc} n={nn} diffusion={diffusion} dense={dense}", flush=True)
    fr = PD.FrozenParam(p, tr, f, normals=normals.astype(np.float64) if mc else None, bin_edges=None if mc else edges)
    assert np.isfinite(fr.out["wfs"]).all()
    vr = fr.valid_rows
    hits = fr.ticks[vr] < fr.nt - 4
    rng = np.random.default_rng(7)
    w_adc = rng.normal(size=hits.shape) * hits
    w_z = rng.normal(size=hits.shape) * hits
    w_wfs = rng.normal(size=(len(fr.U), fr.nt - 1)).astype(np.float32) * 1e-3 if dense else None
    g_ref, _ = fr.vjp(w_adc, w_z, None if w_wfs is None else w_wfs.astype(np.float64))

    theta = theta_of(p).cuda().requires_grad_(True)
    adcs, z, wfs = ParamSimFunction.apply(theta, p, torch.from_numpy(tr).cuda(), f, 0,
                                          torch.from_numpy(normals).cuda() if mc else None, None if mc else edges)
    o = adcs.grad_fn.others
    U = o["unique_pixels"].cpu().numpy()
    assert np.array_equal(U, fr.U)
    # the cotangents on the oracle's grid, gathered at the GPU's hits (same decisions)
    H = int(p.MAX_ADC_VALUES)
    gidx = o["hit_index"].long().cpu().numpy()
    full_t = np.full((len(U), H), fr.nt - 1.0)
    full_t[vr] = fr.ticks[vr]
    assert np.array_equal(o["ticks"].cpu().numpy(), full_t.reshape(-1)[gidx]), "hit decisions differ"
    wa = np.zeros((len(U), H))
    wa[vr] = w_adc
    wz = np.zeros((len(U), H))
    wz[vr] = w_z
    dev = adcs.device
    L = (torch.from_numpy(wa.reshape(-1)[gidx]).to(dev, torch.float32) * adcs).sum() + (
        torch.from_numpy(wz.reshape(-1)[gidx]).to(dev, torch.float32) * z).sum()
    if dense:
        L = L + (torch.from_numpy(w_wfs).to(dev) * wfs).sum()
    L.backward()
    g = theta.grad.cpu().numpy()
    print("gpu ", g)
    print("ref ", g_ref)
    assert g[4] == 0.0 and g_ref[4] == 0.0
    mask = np.abs(g_ref) > 0
    err = np.abs(g[mask] - g_ref[mask]) / np.abs(g_ref[mask])
    print("rel err", err, flush=True)
    assert err.max() < GRAD_TOL


def test_param_strategy_gradients_and_no_grad(params4, c1):
    """HipParametrizedSimulation.predict with parameters that require gradients: adcs and pixel_z
    carry history, a loss on them back-propagates finite gradients to the parameters; under
    torch.no_grad() the same call returns plain tensors with the same hits."""
    from larndsim_amd.strategies import HipParametrizedSimulation

    p, tr, f, _, _ = _setup(params4, c1, True, 1, False, n_seg=300)
    leaves = {n: torch.tensor(float(getattr(p, n)), dtype=torch.float64, requires_grad=True)
              for n in ("eField", "lifetime", "tran_diff", "long_diff")}
    pg = p.replace(**leaves)
    s = HipParametrizedSimulation()
    out = s.predict(pg, torch.from_numpy(tr).cuda(), f, rngkey=3)
    assert out["adcs"].grad_fn is not None and out["pixel_z"].grad_fn is not None
    (out["adcs"].double().sum() + out["pixel_z"].double().sum()).backward()
    for n, v in leaves.items():
        assert v.grad is not None and torch.isfinite(v.grad), n
    with torch.no_grad():
        plain = s.predict(pg, torch.from_numpy(tr).cuda(), f, rngkey=3)
    assert plain["adcs"].grad_fn is None
    assert torch.equal(plain["adcs"], out["adcs"].detach()) and torch.equal(plain["ticks"], out["ticks"])


@pytest.mark.parametrize("diffusion", [False, True], ids=["nodiff", "diff"])
def test_param_adjoint_on_the_200k_segment_event(params4, diffusion):
    """At the benchmark's size (200k segments, 81 pixels per electron, 16.2 M pairs) against central
    differences of the GPU forward itself: a random cotangent on every tick of the finite rows, and
    the parameters that move only the charge (Ab, kb, lifetime; no pixel or whole-tick decision
    moves with them, the waveform is linear in Ab), to 1e-3."""
    from larndsim_amd import synthetic
    from larndsim_amd.autograd import theta_of
    from larndsim_amd.param import ParamSimFunction

    tr, f = synthetic.workload("c2")
    trk = torch.from_numpy(np.ascontiguousarray(tr)).cuda()
    p = params4.replace(mc_diff=True, diffusion_in_current_sim=diffusion)
    print(f"[param-adj] 200k segments diffusion={diffusion}", flush=True)
    theta = theta_of(p).cuda().requires_grad_(True)
    _, _, wfs = ParamSimFunction.apply(theta, p, trk, f, 1, None, None)
    finite = torch.isfinite(wfs).all(dim=1)
    print(f"[param-adj] {wfs.shape[0]} rows, {int((~finite).sum())} with an overflowing pair", flush=True)
    g = torch.Generator(device="cuda").manual_seed(3)
    w = torch.randn(wfs.shape, generator=g, device="cuda", dtype=torch.float64) * finite[:, None]
    (w * torch.nan_to_num(wfs.double(), nan=0.0)).sum().backward()
    grad = theta.grad.cpu().numpy()
    assert np.isfinite(grad).all()

    def value(th):
        with torch.no_grad():
            _, _, wv = ParamSimFunction.apply(th, p, trk, f, 1, None, None)
            assert torch.equal(torch.isfinite(wv).all(dim=1), finite)
            return float((w * torch.nan_to_num(wv.double(), nan=0.0)).sum())

    th0 = theta.detach()
    for i, n in ((0, "Ab"), (1, "kb"), (3, "lifetime")):
        # the float32 waveforms' rounding sets the differences' noise: a wide interval (the charge
        # is linear in Ab and smooth in kb and lifetime: the truncation error stays below 1e-5)
        h = 1e-2 * abs(float(th0[i]))
        e = torch.zeros_like(th0)
        e[i] = h
        fd = (value(th0 + e) - value(th0 - e)) / (2 * h)
        print(f"[param-adj] {n}: adjoint {grad[i]:.6e} differences {fd:.6e}", flush=True)
        assert grad[i] == pytest.approx(fd, rel=1e-3), n
