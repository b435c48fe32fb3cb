"""Batching, determinism and graph capture of the HIP path.

* several uneven events in one call give, per event, exactly (bitwise) the rows, waveforms,
  ADC values and ticks of the same event run alone (the reference's batching contract:
  detsim_jax.py:241-242 packs the event id into the pixel id, so events never mix);
* two runs on the same input are bitwise identical;
* a forward captured into a HIP graph replays to the eager result, replay after replay;
* the gradient of a batch is the sum of the per-event gradients.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


@pytest.fixture(scope="module")
def batch():
    from larndsim_amd import synthetic

    rng = np.random.default_rng(21)
    sizes = [700, 3000, 150, 1800]
    return synthetic.make_batch(rng, sizes), synthetic.SEGMENT_FIELDS, sizes


def _np(st, key):
    return st.out[key][: st.max_rows].cpu().numpy()


def test_batched_equals_per_event(sim, params4, batch):
    tr, f, sizes = batch
    print("[gpu] batched forward of", sizes, flush=True)
    stb = sim.forward(params4, tr, f, n_events=len(sizes))
    ev = tr[:, f.index("eventID")].astype(int)
    pid_b = _np(stb, "pixel_ids")
    npx = params4.n_pixels_x * params4.n_pixels_y * 2
    ev_b = pid_b // npx
    for e in range(len(sizes)):
        sub = tr[ev == e].copy()
        sub[:, f.index("eventID")] = 0
        st1 = sim.forward(params4, sub, f, n_events=1)
        rows = ev_b == e
        assert np.array_equal(pid_b[rows] - e * npx, _np(st1, "pixel_ids")), e
        for key in ("wfs", "adc", "ticks"):
            assert np.array_equal(_np(stb, key)[rows], _np(st1, key)), (e, key)


def test_run_twice_bitwise(sim, params4, batch):
    tr, f, sizes = batch
    a = sim.forward(params4, tr, f, n_events=len(sizes))
    b = sim.forward(params4, tr, f, n_events=len(sizes))
    for key in ("pixel_ids", "wfs", "adc", "ticks"):
        assert np.array_equal(_np(a, key), _np(b, key)), key
    ha, hb = a.out["n_hits"].item(), b.out["n_hits"].item()
    assert ha == hb
    assert torch.equal(a.out["hit_adc"][:ha], b.out["hit_adc"][:hb])


@pytest.mark.parametrize("tile,cost", [(1, 64), (8, 1024), (32, 256)])
def test_schedule_settings_keep_outputs_bitwise(sim, params4, batch, tile, cost):
    """The row order (cell tiles) and the work estimate of the XCD bands change only which
    workgroup computes a row and when: every output is bitwise the default schedule's."""
    from larndsim_amd import sim as S

    tr, f, sizes = batch
    a = sim.forward(params4, tr, f, n_events=len(sizes))
    try:
        S.set_row_tile(tile)
        S.set_row_cost(cost)
        b = sim.forward(params4, tr, f, n_events=len(sizes))
    finally:
        S.set_row_tile(16)
        S.set_row_cost(256)
    for key in ("pixel_ids", "wfs", "adc", "ticks"):
        assert np.array_equal(_np(a, key), _np(b, key)), key


def test_linearity_in_charge_bitwise(sim, params4, batch):
    """Doubling dE doubles n_e exactly, and every waveform value with it."""
    tr, f, sizes = batch
    a = sim.forward(params4, tr, f, n_events=len(sizes))
    t2 = tr.copy()
    t2[:, f.index("dE")] *= 2
    b = sim.forward(params4, t2, f, n_events=len(sizes))
    assert np.array_equal(_np(a, "pixel_ids"), _np(b, "pixel_ids"))
    assert np.array_equal(2 * _np(a, "wfs"), _np(b, "wfs"))


def test_graph_replay(sim, params4, batch):
    from larndsim_amd.sim import Runner

    tr, f, sizes = batch
    run = Runner(sim, params4, torch.from_numpy(tr).cuda(), f, n_events=len(sizes), want_wfs=True)
    run.forward()
    torch.cuda.synchronize()
    ref = {k: run.out[k][: run.rows].clone() for k in ("pixel_ids", "wfs", "adc", "ticks")}
    ref_hits = run.out["n_hits"].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        run.forward()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run.forward()
    torch.cuda.synchronize()
    for rep in range(3):
        for k in ("wfs", "adc", "ticks"):
            run.out[k].fill_(-7)
        run.out["n_hits"].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, v in ref.items():
            assert torch.equal(run.out[k][: run.rows], v), (rep, k)
        assert torch.equal(run.out["n_hits"], ref_hits), rep
        # eager work between replays
        run.forward()
        torch.cuda.synchronize()


def test_batched_gradient_is_sum_of_events(sim, params4, batch):
    tr, f, sizes = batch
    ev = tr[:, f.index("eventID")].astype(int)
    stb = sim.forward(params4, tr, f, n_events=len(sizes), want_tape=True)
    P = stb.max_rows
    H = params4.MAX_ADC_VALUES
    rng = np.random.default_rng(4)
    d_adc = rng.normal(size=(P, H)).astype(np.float32)
    d_z = rng.normal(size=(P, H)).astype(np.float32)
    gb = sim.backward(stb, torch.from_numpy(d_adc), torch.from_numpy(d_z)).cpu().numpy()
    pid_b = _np(stb, "pixel_ids")
    npx = params4.n_pixels_x * params4.n_pixels_y * 2
    gsum = np.zeros(7)
    for e in range(len(sizes)):
        sub = tr[ev == e].copy()
        sub[:, f.index("eventID")] = 0
        st1 = sim.forward(params4, sub, f, n_events=1, want_tape=True)
        rows = pid_b // npx == e
        gsum += sim.backward(st1, torch.from_numpy(d_adc[rows]), torch.from_numpy(d_z[rows])).cpu().numpy()
    assert gb[4] == 0.0
    np.testing.assert_allclose(gb, gsum, rtol=1e-9, atol=0)
