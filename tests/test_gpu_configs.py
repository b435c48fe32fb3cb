"""The three large workloads on the GPU:

* 128 events of log-uniform 2k-400k segments (9.16 M segments) in one graph replay: the full
  pixel-row set bit-exact against the oracle's main-pixel ids, the smallest, the largest and two
  seeded events equal (bitwise) to their runs alone, waveforms and hits of sampled rows against
  the oracle (sim_jax.py:689-736; detsim_jax.py:241-242 packs the event into the pixel id);
* the 50k-segment event: the adjoint with cotangents on a seeded sample of rows, on the hit grids
  and on every waveform tick, against the float64 frozen-decision gradient of the same rows
  (fit_params.py:667-740 differentiates the same chain);
* 32 events of 20k-200k segments (3.3 M segments): one step of the product fitter (mse_adc,
  Adam, the reference's fit-test settings, fit_params.py:781-896), and the batch's adjoint with
  cotangents on one event's rows against the float64 gradient of that event alone.
"""
import numpy as np
import pytest
import torch

from helpers import wfs_error

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


def _npx(p):
    return int(p.n_pixels_x) * int(p.n_pixels_y) * np.asarray(p.tpc_borders).shape[0]


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


# ------------------------------------------------------------------------------------------
# 128 events, one graph replay
# ------------------------------------------------------------------------------------------

def test_c3_batch_in_one_graph_replay(sim, params4, lut, tables):
    from larndsim_amd import synthetic
    from larndsim_amd.sim import Runner
    from oracle import oracle as O

    tr, f = synthetic.workload("c3")
    sizes = synthetic.event_sizes("c3")
    E = len(sizes)
    print(f"[gpu] 128-event batch: {len(tr)} segments, one graph replay", flush=True)
    run = Runner(sim, params4, torch.from_numpy(tr).cuda(), f, n_events=E, want_wfs=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run.forward()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run.forward()
    run.out["wfs"].fill_(-7)
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    P = run.rows
    assert int(run.out["n_rows"].item()) == P
    pids = run.out["pixel_ids"][:P].cpu().numpy()

    # the whole row set against the oracle's main pixels
    main = O.main_pixel_ids(params4, tr, f)
    U = np.unique(main[main >= 0])
    assert np.array_equal(pids, U), "batched pixel rows differ from the oracle's"
    ev_row = pids // _npx(params4)
    assert np.array_equal(np.unique(ev_row), np.arange(E))

    ticks = run.out["ticks"][:P].cpu().numpy()
    assert int(run.out["n_hits"].item()) == int((ticks < 1997).sum())
    ev = tr[:, f.index("eventID")].astype(np.int64)

    # per-event runs, bitwise: smallest, largest, two seeded others
    rng = np.random.default_rng(31)
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), *map(int, rng.choice(E, 2, replace=False))})
    for e in picks:
        sub = tr[ev == e].copy()
        sub[:, f.index("eventID")] = 0
        print(f"[gpu] event {e} alone ({len(sub)} segments)", flush=True)
        st1 = sim.forward(params4, sub, f, n_events=1)
        rows = np.nonzero(ev_row == e)[0]
        r0, r1 = int(rows[0]), int(rows[-1]) + 1
        assert np.array_equal(pids[r0:r1] - e * _npx(params4), st1.out["pixel_ids"][: st1.max_rows].cpu().numpy())
        for key in ("wfs", "adc", "ticks"):
            assert torch.equal(run.out[key][r0:r1], st1.out[key][: st1.max_rows]), (e, key)
        del st1

    # waveforms and hits of sampled rows of two events against the oracle
    for e in (int(np.argmin(sizes)), picks[len(picks) // 2]):
        sub = tr[ev == e].copy()
        sub[:, f.index("eventID")] = 0
        rows = np.nonzero(ev_row == e)[0]
        sample = np.sort(rng.choice(rows, min(10, len(rows)), replace=False))
        local = pids[sample] - e * _npx(params4)
        w_ref, U_ref, _ = O.simulate_wfs(params4, lut, sub, f, rows_subset_pids=local, tables=tables)
        ref = w_ref[np.searchsorted(U_ref, local)]
        got = run.out["wfs"][torch.from_numpy(sample).cuda()].cpu().numpy().astype(np.float64)
        rel, ab = wfs_error(got, ref)
        print(f"[c3] event {e}: {len(sample)} sampled rows, wfs rel err {rel:.3e} abs {ab:.3e}", flush=True)
        assert rel < 1e-5
        a_ref, t_ref, _ = O.get_adc_values(params4, ref)
        same = ticks[sample] == t_ref
        assert same.mean() >= 1 - 1e-3
        ok = same.all(axis=1)
        adc = run.out["adc"][torch.from_numpy(sample).cuda()].cpu().numpy()
        np.testing.assert_allclose(adc[ok], O.digitize(params4, a_ref)[ok], rtol=1e-5, atol=1e-4)


# ------------------------------------------------------------------------------------------
# adjoint on sampled rows of large events
# ------------------------------------------------------------------------------------------

def _sampled_adjoint(sim, p, lut, tables, st, tr_event, f, rows_global, pid_offset, seed, dense):
    """GPU adjoint with cotangents only on rows ``rows_global`` of the forward state ``st`` against
    the float64 frozen-decision gradient of the same cotangents on the event ``tr_event`` alone
    (its pixel ids are the batch's minus ``pid_offset``).  Rows whose hit ticks differ from the
    oracle's (threshold ties) get no cotangent."""
    from oracle import ad_oracle as A

    P, H = st.max_rows, sim.H
    pids = st.out["pixel_ids"][:P].cpu().numpy()
    local = pids[rows_global] - pid_offset
    fz = A.Frozen(p, lut, tr_event, f, tables=tables, rows_subset_pids=local)
    assert np.array_equal(fz.U, local)
    gt = st.out["ticks"][:P].cpu().numpy()[rows_global]
    same = (gt == fz.ticks.astype(np.int64)).all(axis=1)
    print(f"[adjoint] {len(rows_global)} rows, {int(same.sum())} without a tie, "
          f"{int((fz.ticks[same] < 1997).sum())} hits", flush=True)
    assert same.mean() >= 0.9
    rng = np.random.default_rng(seed)
    hits = (fz.ticks < 1997) & same[:, None]
    w_adc = rng.normal(size=hits.shape) * hits
    w_z = rng.normal(size=hits.shape) * hits
    w_wfs = rng.normal(size=(len(local), sim.nout)) * same[:, None] if dense else None
    th = fz.theta0().requires_grad_(True)
    adc, z, wfs = fz.forward(th)
    loss = (torch.as_tensor(w_adc) * adc).sum() + (torch.as_tensor(w_z) * z).sum()
    if dense:
        loss = loss + (torch.as_tensor(w_wfs) * wfs).sum()
    loss.backward()
    g_ref = th.grad.detach().numpy()

    d_adc = torch.zeros((P, H), dtype=torch.float32)
    d_z = torch.zeros((P, H), dtype=torch.float32)
    d_adc[rows_global] = torch.from_numpy(w_adc.astype(np.float32))
    d_z[rows_global] = torch.from_numpy(w_z.astype(np.float32))
    d_wfs = None
    if dense:
        d_wfs = torch.zeros((P, sim.nout), dtype=torch.float32, device=sim.device)
        d_wfs[torch.from_numpy(rows_global).cuda()] = torch.from_numpy(w_wfs.astype(np.float32)).cuda()
    g = sim.backward(st, d_adc, d_z, d_wfs).cpu().numpy()
    print("gpu ", g)
    print("ref ", g_ref)
    mask = np.abs(g_ref) > 0
    assert mask.sum() >= 5
    assert g[4] == 0.0
    err = _rel(g[mask], g_ref[mask])
    print("rel err", err, flush=True)
    assert err.max() < GRAD_TOL


@pytest.mark.parametrize("dense", [False, True], ids=["hits", "waveforms"])
def test_c4_adjoint_on_sampled_rows(sim, params4, lut, tables, dense):
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c4")
    print(f"[gpu] 50k-segment event, adjoint on sampled rows, dense={dense}", flush=True)
    st = sim.forward(params4, tr, f, n_events=1, want_tape=True)
    P = st.max_rows
    ticks = st.out["ticks"][:P].cpu().numpy()
    rng = np.random.default_rng(41)
    with_hits = np.nonzero((ticks < 1997).any(axis=1))[0]
    rows = np.unique(np.concatenate([rng.choice(with_hits, 18, replace=False), rng.choice(P, 6, replace=False)]))
    _sampled_adjoint(sim, params4, lut, tables, st, tr, f, rows, 0, 42, dense)


def test_c5_fit_step_and_event_gradient(sim, params4, lut, tables):
    from larndsim_amd import consts, synthetic
    from larndsim_amd.fit import GradientDescentFitter, clip_by_global_norm, map_phys_to_norm
    from larndsim_amd.losses import mse_adc_blockwise
    from larndsim_amd.strategies import GenericLossStrategy, HipLUTSimulation

    tr, f = synthetic.workload("c5")
    sizes = synthetic.event_sizes("c5")
    ftr = torch.from_numpy(tr).cuda()
    print(f"[gpu] 32-event batch ({len(tr)} segments): one fit step", flush=True)
    target_params = synthetic.perturbed_params(params4, seed=5)
    strat = HipLUTSimulation(lut)
    with torch.no_grad():
        target = strat.predict(target_params, ftr, f)
    names = [n for n in consts.GRAD_PARAMS if n != "vdrift"]
    fitter = GradientDescentFitter(strat, GenericLossStrategy(mse_adc_blockwise), params4, names, lr=2e-2,
                                   max_clip_norm_val=1.0, lr_scheduler="exponential_decay",
                                   lr_kw=dict(decay_rate=0.97))
    before = dict(fitter.norm)
    loss, grads = fitter.step(ftr, f, target)
    torch.cuda.synchronize()
    assert np.isfinite(loss) and loss > 0
    gv = np.array([grads[n] for n in names])
    assert np.isfinite(gv).all() and (gv != 0).all()
    # Adam's first step on the clipped gradient g: mu_hat = g, nu_hat = g^2, so each normalised
    # parameter moves by -lr g / (|g| + eps)
    gc = clip_by_global_norm(grads, 1.0)
    for n in names:
        want = -2e-2 * gc[n] / (abs(gc[n]) + 1e-8)
        assert abs((fitter.norm[n] - before[n]) - want) <= 1e-12 + 1e-9 * abs(want), n
    assert before["Ab"] == pytest.approx(float(map_phys_to_norm(float(params4.Ab), "Ab")), abs=1e-12)
    del target, fitter

    # float64 spot check of one event's gradient inside the batch
    print("[gpu] 32-event batch adjoint, cotangents on the smallest event", flush=True)
    st = sim.forward(params4, ftr, f, n_events=len(sizes), want_tape=True)
    P = st.max_rows
    pids = st.out["pixel_ids"][:P].cpu().numpy()
    e = int(np.argmin(sizes))
    ev = tr[:, f.index("eventID")].astype(np.int64)
    sub = tr[ev == e].copy()
    sub[:, f.index("eventID")] = 0
    ticks = st.out["ticks"][:P].cpu().numpy()
    rows_e = np.nonzero(pids // _npx(params4) == e)[0]
    rng = np.random.default_rng(51)
    with_hits = rows_e[(ticks[rows_e] < 1997).any(axis=1)]
    rows = np.unique(np.concatenate([rng.choice(with_hits, 14, replace=False), rng.choice(rows_e, 4, replace=False)]))
    _sampled_adjoint(sim, params4, lut, tables, st, sub, f, rows, e * _npx(params4), 52, dense=False)
