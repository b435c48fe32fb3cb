"""Readout parity on the GPU: the eight compacted hit arrays of ``lsim_forward`` and
``lsim_fee_forward`` (parse_output, sim_jax.py:619-647; simulate_stochastic, sim_jax.py:738-769;
get_pixel_coordinates / get_hit_z, detsim_jax.py:296-319) against the FP64 oracle, segments
moved by non-zero shifts (shift_tracks, sim_jax.py:108-119), and the row layout of
``HipLUTSimulation.predict`` (unique_pixels / wfs padded as sim_jax.py:716-721 pads them).

Bars: ids, ticks, events bit-exact; hit_prob exact; ADC to 1e-5 relative; x, y, z to 1e-5 cm."""
import numpy as np
import pytest
import torch

from helpers import compare_forward, compare_hits, mixed_event

pytestmark = pytest.mark.gpu

SHIFT = dict(shift_x=0.1234, shift_y=-0.0777, shift_z=0.3125)


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


def test_hits_c1(sim, params4, lut, c1, tables):
    tr, f = c1
    st, ref = compare_forward(sim, params4, lut, tr, f, tables=tables, tag="c1 hits")
    assert compare_hits(st, ref, "c1") > 100


def test_hits_mixed_event(sim, params4, lut, tables):
    """Both anode planes (hit z runs the other way in TPC 1), plane borders, out-of-TPC segments
    and padding rows."""
    tr, f = mixed_event(params4)
    st, ref = compare_forward(sim, params4, lut, tr, f, tables=tables, n_events=1, tag="mixed hits")
    planes = (ref["hit_pixels"] // (params4.n_pixels_x * params4.n_pixels_y)) % 2
    assert set(np.unique(planes)) == {0, 1}
    assert compare_hits(st, ref, "mixed") > 50


@pytest.mark.parametrize("event", ["c1", "mixed"])
def test_shifted_segments(sim, params4, lut, c1, tables, event):
    """shift_x/y/z are subtracted from the positions before anything else (sim_jax.py:108-119):
    rows, waveforms and hits follow the shifted positions.  The shifts are not multiples of the
    sub-bin width (0.04434 cm) or the tick length, so bins, pixels and windows all move."""
    p = params4.replace(**SHIFT)
    tr, f = c1 if event == "c1" else mixed_event(params4)
    st, ref = compare_forward(sim, p, lut, tr, f, tables=tables, n_events=1, tag=f"shifted {event}")
    compare_hits(st, ref, f"shifted {event}")
    # the shift changed the outcome (the same segments unshifted give other rows or waveforms)
    st0 = sim.forward(params4, tr, f, n_events=1)
    torch.cuda.synchronize()
    same_rows = st0.max_rows == st.max_rows and torch.equal(st0.out["pixel_ids"][: st.max_rows],
                                                             st.out["pixel_ids"][: st.max_rows])
    assert not same_rows or not torch.equal(st0.out["wfs"][: st.max_rows], st.out["wfs"][: st.max_rows])


def test_stochastic_readout_of_oracle_waveforms(sim, params4, lut, c1, tables):
    """lsim_fee_forward (simulate_stochastic on caller waveforms): the oracle's waveforms, rounded
    to float32, with the oracle's rows (the -1 row included), through the GPU readout and the
    oracle's readout of the same float32 values."""
    from larndsim_amd.sim import hits_dict
    from oracle import oracle as O

    tr, f = c1
    wfs, U, _ = O.simulate_wfs(params4, lut, tr, f, tables=tables)
    w32 = wfs.astype(np.float32)
    ref = O.simulate_stochastic(params4, w32.astype(np.float64), U)
    print("[gpu] fee forward on oracle waveforms", flush=True)
    out = sim.fee(params4, torch.from_numpy(w32), torch.from_numpy(U))
    torch.cuda.synchronize()
    h = {k: v.cpu().numpy() for k, v in hits_dict(out).items()}
    same = out["ticks"].cpu().numpy() == ref["grid"]["ticks"].astype(np.int64)
    print(f"[fee] tick agreement {same.mean():.6f}", flush=True)
    assert same.all()
    assert len(h["adcs"]) == len(ref["adcs"]) > 100
    for k in ("ticks", "event", "hit_pixels"):
        assert np.array_equal(h[k].astype(np.int64), ref[k].astype(np.int64)), k
    assert np.array_equal(h["hit_prob"], ref["hit_prob"].astype(np.float32))
    np.testing.assert_allclose(h["adcs"], ref["adcs"], rtol=1e-5, atol=1e-4)
    for k in ("pixel_x", "pixel_y", "pixel_z"):
        np.testing.assert_allclose(h[k], ref[k], rtol=0, atol=1e-5, err_msg=k)


def _reference_layout(q_main_pid):
    """unique_pixels as simulate_wfs lays it out (sim_jax.py:716-721), for a fresh pad_size
    history: sort(unique(main) ++ [-1] padded with -1 to int(count * 1.1 + 0.5))."""
    u = np.append(np.unique(q_main_pid), -1)
    n = int(len(u) * (1 + 0.2 / 2) + 0.5)
    return np.sort(np.concatenate([u, np.full(n - len(u), -1)])).astype(np.int64)


@pytest.mark.parametrize("grad", [False, True], ids=["values", "autograd"])
def test_predict_row_layout(params4, lut, tables, grad):
    """HipLUTSimulation.predict returns unique_pixels and wfs with the reference's shape and ids on
    an event with padding rows (eventID -1) and off-plane segments: the negative ids first, then
    the -1 rows, then the pixel rows; the extra rows hold zero waveforms."""
    from larndsim_amd import sim as S
    from larndsim_amd.consts import GRAD_PARAMS
    from larndsim_amd.strategies import HipLUTSimulation
    from oracle import oracle as O

    tr, f = mixed_event(params4)
    q = O.segment_quantities(params4, tr, f)
    assert (q["main_pid"] < -1).any() and (q["main_pid"] == -1).any()
    want = _reference_layout(q["main_pid"])
    S.size_history_dict.pop("unique_pixels", None)
    p = params4
    if grad:
        p = params4.replace(**{n: torch.tensor(float(getattr(params4, n)), dtype=torch.float64, requires_grad=True)
                               for n in GRAD_PARAMS})
    print(f"[gpu] predict row layout grad={grad}", flush=True)
    out = HipLUTSimulation(lut).predict(p, tr, f, 0)
    ids = out["unique_pixels"].cpu().numpy()
    assert np.array_equal(ids, want)
    wfs = out["wfs"].detach()
    assert wfs.shape == (len(want), params4.n_ticks - 1)
    P = int((want >= 0).sum())
    assert torch.count_nonzero(wfs[: len(want) - P]) == 0
    if grad:
        (out["wfs"].sum() + out["adcs"].sum()).backward()
        assert p.Ab.grad is not None and float(p.Ab.grad) != 0.0


def test_predict_under_no_grad(params4, lut, c1):
    """Parameters that require gradients, called under torch.no_grad(): no autograd record, the
    plain outputs come back (the reference's predict works the same in both modes)."""
    from larndsim_amd.consts import GRAD_PARAMS
    from larndsim_amd.strategies import HipLUTSimulation

    tr, f = c1
    p = params4.replace(**{n: torch.tensor(float(getattr(params4, n)), dtype=torch.float64, requires_grad=True)
                           for n in GRAD_PARAMS})
    strat = HipLUTSimulation(lut)
    with torch.no_grad():
        a = strat.predict(p, tr, f, 0)
    b = strat.predict(params4, tr, f, 0)
    assert a["adcs"].grad_fn is None
    for k in ("adcs", "pixel_x", "pixel_y", "pixel_z", "ticks", "hit_prob", "event", "hit_pixels", "unique_pixels",
              "wfs"):
        assert torch.equal(a[k], b[k]), k
