"""Adjoint parity: the HIP backward against torch-float64 autograd of the frozen-decision chain.

Tolerance: relative error <= 1e-3 per parameter (the gradient accuracy target); the vdrift
gradient is exactly zero (the parameter is never read by the simulation)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3


@pytest.fixture(scope="module")
def frozen(params4, lut, c1, tables):
    from oracle import ad_oracle as A

    tr, f = c1
    return A.Frozen(params4, lut, tr, f, tables=tables)


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


def _cotangents(frozen, seed):
    vr = frozen.valid_rows
    hits = frozen.ticks[vr] < 1997
    rng = np.random.default_rng(seed)
    return rng.normal(size=hits.shape) * hits, rng.normal(size=hits.shape) * hits


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_autograd(sim, params4, c1, frozen, seed):
    tr, f = c1
    print(f"[gpu] backward c1 seed {seed}", flush=True)
    st = sim.forward(params4, tr, f, want_tape=True)
    P = st.max_rows
    ticks = st.out["ticks"][:P].cpu().numpy()
    assert np.array_equal(ticks, frozen.ticks[frozen.valid_rows].astype(np.int64)), "hit decisions differ"
    w_adc, w_z = _cotangents(frozen, seed)
    g_ref, _ = frozen.vjp(w_adc, w_z)
    g = sim.backward(st, torch.from_numpy(w_adc.astype(np.float32)), torch.from_numpy(w_z.astype(np.float32)))
    g = g.cpu().numpy()
    print("gpu ", g)
    print("ref ", g_ref)
    assert g[4] == 0.0 and g_ref[4] == 0.0
    mask = np.abs(g_ref) > 0
    err = _rel(g[mask], g_ref[mask])
    print("rel err", err)
    assert err.max() < GRAD_TOL


def _edge_event(n_c1=300):
    """C1 segments plus segments next to the anode of each TPC, whose windows start at or before
    output tick 0 (window start nt - L - k0 <= 0 once the drift time to the cathode exceeds
    (nt - L) ticks), so the clamped first samples are exercised in both passes of the adjoint."""
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    near = tr[n_c1:n_c1 + 120].copy()
    z = near[:, f.index("z")]
    near[:, f.index("z")] = 28.6 + np.mod(z, 1.9)
    near[::2, f.index("z")] *= -1  # every other one at the TPC 1 anode
    for k in ("z_start", "z_end"):
        near[:, f.index(k)] = near[:, f.index("z")]
    return np.concatenate([tr[:n_c1], near]), f


VARIANTS = [dict(nb_tran_diff_bins=3), dict(nb_tran_diff_bins=1), dict(number_pix_neighbors=0),
            dict(MAX_ADC_VALUES=4), dict(signal_length=187)]


@pytest.mark.parametrize("dense", [False, True], ids=["hits", "waveforms"])
@pytest.mark.parametrize("kw", VARIANTS, ids=["nb3", "nb1", "n0", "H4", "L187"])
def test_backward_other_settings(params4, lut, kw, dense):
    """The adjoint in every compiled bin-count instance (5 is the default; 3 and 1 merge bins
    differently), with number_pix_neighbors = 0 (rows get no neighbour stripes but still the
    diffusion bins that spill over from segments of the adjacent pixels), fewer self-trigger passes
    than the compiled maximum, and the longest window; with cotangents on the hit grids only and
    with a cotangent on every waveform tick as well (the dense variant), on an event whose
    near-anode windows start at or before tick 0."""
    from larndsim_amd.sim import Simulator
    from oracle import ad_oracle as A
    from oracle import oracle as O

    p0 = params4.replace(**kw)
    tr, f = _edge_event()
    fz = A.Frozen(p0, lut, tr, f, tables=O.build_templates(p0, lut))
    k0 = fz.k0[fz.q["mask"] > 0]
    assert (1950 - int(p0.signal_length) - k0 <= 0).any()
    s0 = Simulator(p0, lut)
    print(f"[gpu] backward edge event {kw} dense={dense}", flush=True)
    st = s0.forward(p0, tr, f, want_tape=True)
    P = st.max_rows
    assert np.array_equal(st.out["ticks"][:P].cpu().numpy(), fz.ticks[fz.valid_rows].astype(np.int64))
    w_adc, w_z = _cotangents(fz, 3)
    vr = torch.as_tensor(fz.valid_rows)
    th = fz.theta0().requires_grad_(True)
    adc, z, wfs = fz.forward(th)
    loss = (torch.as_tensor(w_adc) * adc[vr]).sum() + (torch.as_tensor(w_z) * z[vr]).sum()
    w_wfs = None
    if dense:
        w_wfs = np.random.default_rng(7).normal(size=(P, s0.nout))
        loss = loss + (torch.as_tensor(w_wfs) * wfs[vr]).sum()
    loss.backward()
    g_ref = th.grad.detach().numpy()
    g = s0.backward(st, torch.from_numpy(w_adc.astype(np.float32)), torch.from_numpy(w_z.astype(np.float32)),
                    None if w_wfs is None else torch.from_numpy(w_wfs.astype(np.float32))).cpu().numpy()
    print("gpu ", g)
    print("ref ", g_ref)
    mask = np.abs(g_ref) > 0
    assert mask.sum() >= 5
    assert g[4] == 0.0
    assert _rel(g[mask], g_ref[mask]).max() < GRAD_TOL


def test_autograd_function(params4, lut, c1, frozen):
    """Gradient through HipLUTSimulation.predict (the reference's LUTSimulation.predict path)."""
    from larndsim_amd.consts import GRAD_PARAMS
    from larndsim_amd.strategies import HipLUTSimulation

    tr, f = c1
    theta = torch.tensor([float(getattr(params4, n)) for n in GRAD_PARAMS], dtype=torch.float64, requires_grad=True)
    p = params4.replace(**{n: theta[i] for i, n in enumerate(GRAD_PARAMS)})
    pred = HipLUTSimulation(lut).predict(p, tr, f, 0)
    n = pred["adcs"].numel()
    rng = np.random.default_rng(5)
    wa = torch.from_numpy(rng.normal(size=n)).float().cuda()
    wz = torch.from_numpy(rng.normal(size=n)).float().cuda()
    loss = (wa * pred["adcs"]).sum() + (wz * pred["pixel_z"]).sum()
    loss.backward()
    # the same cotangents placed on the oracle's (row, hit) grid
    vr = frozen.valid_rows
    idx = pred["hit_index"].cpu().numpy() if "hit_index" in pred else None
    grid_a = np.zeros(frozen.ticks[vr].size)
    grid_z = np.zeros(frozen.ticks[vr].size)
    hi = (frozen.ticks[vr].ravel() < 1997).nonzero()[0]
    assert len(hi) == n
    grid_a[hi] = wa.cpu().numpy()
    grid_z[hi] = wz.cpu().numpy()
    g_ref, _ = frozen.vjp(grid_a.reshape(frozen.ticks[vr].shape), grid_z.reshape(frozen.ticks[vr].shape))
    g = theta.grad.numpy()
    mask = np.abs(g_ref) > 0
    assert _rel(g[mask], g_ref[mask]).max() < GRAD_TOL
    assert g[4] == 0.0


def test_autograd_releases_workspace(params4, lut, c1):
    """Repeated predict + backward keeps device memory flat without the cycle collector: the
    autograd context holds no reference to its own outputs and drops the forward state after
    the backward."""
    import gc

    from larndsim_amd.consts import GRAD_PARAMS
    from larndsim_amd.strategies import HipLUTSimulation

    tr, f = c1
    strat = HipLUTSimulation(lut)
    gc.collect()
    gc.disable()
    try:
        used = []
        for _ in range(6):
            theta = {n: torch.tensor(float(getattr(params4, n)), dtype=torch.float64, requires_grad=True)
                     for n in GRAD_PARAMS}
            pred = strat.predict(params4.replace(**theta), tr, f, 0)
            pred["adcs"].sum().backward()
            del pred, theta
            torch.cuda.synchronize()
            used.append(torch.cuda.memory_allocated())
        print("allocated after each step:", used)
        assert used[-1] <= used[1]
    finally:
        gc.enable()


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_with_waveform_cotangent(sim, params4, c1, frozen, seed):
    """lsim_backward_wfs: cotangents on the waveforms as well as on the ADC and z grids (the row
    cotangent becomes one breakpoint per tick) against float64 autograd of the same chain."""
    tr, f = c1
    print(f"[gpu] backward with waveform cotangent c1 seed {seed}", flush=True)
    st = sim.forward(params4, tr, f, want_tape=True)
    vr = frozen.valid_rows
    P = st.max_rows
    assert P == int(vr.sum())
    w_adc, w_z = _cotangents(frozen, seed)
    rng = np.random.default_rng(100 + seed)
    w_wfs = rng.normal(size=(P, sim.nout))
    th = frozen.theta0().requires_grad_(True)
    adc, z, wfs = frozen.forward(th)
    vrt = torch.as_tensor(vr)
    L = (torch.as_tensor(w_adc) * adc[vrt]).sum() + (torch.as_tensor(w_z) * z[vrt]).sum() + \
        (torch.as_tensor(w_wfs) * wfs[vrt]).sum()
    L.backward()
    g_ref = th.grad.detach().numpy()
    g = sim.backward(st, torch.from_numpy(w_adc.astype(np.float32)), torch.from_numpy(w_z.astype(np.float32)),
                     torch.from_numpy(w_wfs.astype(np.float32))).cpu().numpy()
    print("gpu ", g)
    print("ref ", g_ref)
    mask = np.abs(g_ref) > 0
    err = _rel(g[mask], g_ref[mask])
    print("rel err", err)
    assert err.max() < GRAD_TOL


def test_autograd_through_waveforms(params4, lut, c1):
    """A loss on HipLUTSimulation.predict's waveforms back-propagates to the parameters, equal to
    Simulator.backward with the same waveform cotangent."""
    from larndsim_amd.sim import Simulator
    from larndsim_amd.strategies import HipLUTSimulation

    tr, f = c1
    strat = HipLUTSimulation(lut)
    names = ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff")
    th = {n: torch.tensor(float(getattr(params4, n)), dtype=torch.float64, requires_grad=True) for n in names}
    out = strat.predict(params4.replace(**th), torch.from_numpy(tr).cuda(), f)
    w = torch.linspace(-1, 1, out["wfs"].numel(), device="cuda").reshape(out["wfs"].shape)
    (w * out["wfs"]).sum().backward()
    g = np.array([float(th[n].grad) for n in names])
    sim = Simulator(params4, lut)
    st = sim.forward(params4, tr, f, want_tape=True)
    d_adc = torch.zeros(st.max_rows * sim.H)
    # predict's rows: the reference's leading negative-id and -1 rows (no charge), then the P pixel rows
    assert (out["unique_pixels"][: out["wfs"].shape[0] - st.max_rows] < 0).all()
    g2 = sim.backward(st, d_adc, None, w[w.shape[0] - st.max_rows:]).cpu().numpy()
    print("autograd", g, "direct", g2, flush=True)
    assert np.allclose(g, g2, rtol=1e-6, atol=0)
    assert np.abs(g).max() > 0


def _waveform_mse(out, tgt):
    """sum over pixels and ticks of (w - w_target)^2 / sum w_target^2, rows matched by pixel id (a
    pixel missing on one side counts as zeros)."""
    ids, w = out["unique_pixels"], out["wfs"]
    tids, tw = tgt["unique_pixels"], tgt["wfs"]
    u, inv = torch.unique(torch.cat([ids, tids]), return_inverse=True)
    A = torch.zeros(len(u), w.shape[1], dtype=torch.float64, device=w.device).index_add(0, inv[: len(ids)], w.double())
    B = torch.zeros(len(u), w.shape[1], dtype=torch.float64, device=w.device).index_add(0, inv[len(ids):], tw.double())
    return ((A - B) ** 2).sum() / (B ** 2).sum()


def test_waveform_loss_matches_central_differences(params4, lut, c1):
    """A loss on the waveforms has no hit decision in it, so the adjoint must equal central
    differences (h = 1e-3 relative) of the whole loss, through the float32 forward, for every
    parameter that leaves the drift time alone.  eField moves the drift times, and the waveforms
    are linear in them between whole ticks: over +-h many segments cross a tick
    and the differences average the slopes on both sides, so eField is held to 5e-2 here (the
    oracle test of the same loss shows the differences closing on the gradient as h shrinks)."""
    from larndsim_amd.strategies import HipLUTSimulation

    tr, f = c1
    strat = HipLUTSimulation(lut)
    trd = torch.from_numpy(tr).cuda()
    names = ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff")
    v0 = {n: float(getattr(params4, n)) for n in names}
    scale = {"Ab": 1.02, "kb": 0.97, "eField": 1.01, "lifetime": 0.9, "vdrift": 1.0, "long_diff": 1.1, "tran_diff": 0.92}
    with torch.no_grad():
        tgt = strat.predict(params4.replace(**{n: v0[n] * scale[n] for n in names}), trd, f)
    th = {n: torch.tensor(v0[n], dtype=torch.float64, requires_grad=True) for n in names}
    _waveform_mse(strat.predict(params4.replace(**th), trd, f), tgt).backward()
    g = np.array([float(th[n].grad) if th[n].grad is not None else 0.0 for n in names])
    fd = np.zeros(len(names))
    with torch.no_grad():
        for i, n in enumerate(names):
            h = 1e-3 * abs(v0[n])
            lp = float(_waveform_mse(strat.predict(params4.replace(**{n: v0[n] + h}), trd, f), tgt))
            lm = float(_waveform_mse(strat.predict(params4.replace(**{n: v0[n] - h}), trd, f), tgt))
            fd[i] = (lp - lm) / (2 * h)
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)
    print("adjoint", g, "\ncentral", fd, "\nrel", rel, flush=True)
    assert g[4] == 0.0 and fd[4] == 0.0
    smooth = [names.index(n) for n in ("Ab", "kb", "lifetime", "long_diff", "tran_diff")]
    assert all(abs(fd[i]) * abs(v0[names[i]]) > 1e-6 for i in smooth)  # d loss / d log(param) above noise
    assert rel[smooth].max() < GRAD_TOL
    assert rel[names.index("eField")] < 5e-2
