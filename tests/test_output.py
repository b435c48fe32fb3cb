"""Hit output files (optimize/simulate.py:125-179) on hand-made readout arrays: the stored-hit
selection, the per-event group layout and the npz round trip."""
import numpy as np
import pytest
import torch

from larndsim_amd import consts, output
from larndsim_amd.fee import adc2charge, digitize


@pytest.fixture(scope="module")
def params():
    return consts.default_params()


def _readout(params):
    low = float(digitize(params, torch.tensor(float(params.DISCRIMINATION_THRESHOLD), dtype=torch.float32)))
    adcs = np.array([low, 120.0, 130.0, 90.0, 200.0, 101.0], np.float32)
    event = np.array([0, 0, -1, 1, 1, 0])
    pid = np.arange(6) * 11
    x, y, z = (np.arange(6, dtype=np.float32) + k for k in (0.5, 10.5, 20.5))
    ticks = np.array([5, 6, 7, 8, 9, 10], np.float32)
    prob = np.ones(6, np.float32)
    return adcs, x, y, z, ticks, prob, event, pid


def test_select_hits(params):
    adcs, x, y, z, ticks, prob, event, pid = _readout(params)
    h = output.select_hits(params, torch.from_numpy(adcs), x, y, z, ticks, prob, torch.from_numpy(event), pid)
    keep = np.array([1, 3, 4, 5])  # slot 0 sits at the threshold ADC, slot 2 is padding
    assert np.array_equal(h["adc"], adcs[keep])
    assert np.array_equal(h["pixels"], pid[keep]) and np.array_equal(h["event"], event[keep])
    assert np.allclose(h["Q"], adc2charge(torch.from_numpy(adcs[keep]).double(), params).numpy(), rtol=1e-6)
    assert np.allclose(h["adc_clean"], adcs[keep] - adcs[0])
    g = output.event_groups(h, {0: 1005, 1: 2000})
    assert sorted(g) == ["event_1005", "event_2000"]
    assert set(g["event_1005"]) == set(output.EVENT_KEYS)
    assert np.array_equal(g["event_1005"]["pixels"], pid[[1, 5]])
    assert np.array_equal(g["event_2000"]["eventID"], [2000, 2000])


def test_npz_round_trip(params, tmp_path):
    adcs, x, y, z, ticks, prob, event, pid = _readout(params)
    h1 = output.select_hits(params, adcs, x, y, z, ticks, prob, event, pid)
    h2 = output.select_hits(params, adcs[::-1].copy(), x, y, z, ticks, prob, event, pid)
    path = tmp_path / "hits.npz"
    output.save_npz(path, [h1, h2])
    z_ = output.load_npz(path)
    assert tuple(sorted(z_)) == tuple(sorted(output.NPZ_KEYS))
    assert np.array_equal(z_["adcs"], np.concatenate([h1["adc"], h2["adc"]]))
    assert np.array_equal(z_["event_id"], np.concatenate([h1["event"], h2["event"]]))
    assert np.array_equal(z_["z"], np.concatenate([h1["pix_z"], h2["pix_z"]]))


def test_hdf5_batches_round_trip(params, tmp_path):
    """write_hdf5 appends batch_<i>/event_<g>/<dataset> groups (simulate.py:135-155); the file reads
    back with every dataset's values and dtype, and a second write of the same batch raises as
    h5py's create_group does."""
    adcs, x, y, z, ticks, prob, event, pid = _readout(params)
    h = output.select_hits(params, adcs, x, y, z, ticks, prob, event, pid)
    g0 = output.event_groups(h, {0: 1005, 1: 2000})
    g1 = output.event_groups(h, {0: 7, 1: 8})
    path = tmp_path / "hits.h5"
    output.write_hdf5(path, 0, g0)
    output.write_hdf5(path, 1, g1)
    back = output.read_hdf5(path)
    assert sorted(back) == ["batch_0", "batch_1"]
    assert sorted(back["batch_1"]) == ["event_7", "event_8"]
    for b, g in (("batch_0", g0), ("batch_1", g1)):
        for ev, ds in g.items():
            assert set(back[b][ev]) == set(output.EVENT_KEYS)
            for k, v in ds.items():
                got = back[b][ev][k]
                assert got.dtype == np.asarray(v).dtype and np.array_equal(got, v), (b, ev, k)
    with pytest.raises(ValueError):
        output.write_hdf5(path, 1, g1)
