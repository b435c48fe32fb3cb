"""The package's HDF5 reader (larndsim_amd/h5lite.py) on two of the reference's own input files
(prepared_data/input_{0,1}.h5, kept as fixtures under tests/golden/), and the host logic of
TracksDataset (optimize/dataio.py:108-424).  No GPU.

The decoding is pinned by the file's own redundancy: the 26 field names and types of the
compound header, each segment's length equal to the distance between its end points, its
position at their midpoint and its energy equal to dE/dx times length -- a wrong member offset,
type or byte order breaks these identities."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = [os.path.join(GOLD, "segments_input_0.h5"), os.path.join(GOLD, "segments_input_1.h5")]
FIELDS = ("eventID", "z_end", "trackID", "tran_diff", "z_start", "x_end", "y_end", "n_electrons", "pdgId", "x_start",
          "y_start", "t_start", "t0_start", "t0_end", "t0", "dx", "long_diff", "pixel_plane", "t_end", "dEdx", "dE", "t",
          "y", "x", "z", "n_photons")


@pytest.mark.parametrize("fn", FILES)
def test_segments_decode(fn):
    from larndsim_amd.h5lite import read_dataset

    a = read_dataset(fn, "segments")
    assert a.dtype.names == FIELDS and a.ndim == 1 and len(a) > 100
    assert a.dtype["t0"] == np.float64 and a.dtype["x"] == np.float32 and a.dtype["eventID"].kind == "u"
    d = np.sqrt((a["x_end"] - a["x_start"]) ** 2 + (a["y_end"] - a["y_start"]) ** 2 + (a["z_end"] - a["z_start"]) ** 2)
    np.testing.assert_allclose(a["dx"], d, rtol=2e-5, atol=1e-5)
    for c in "xyz":
        np.testing.assert_allclose(a[c], (a[c + "_start"] + a[c + "_end"]) / 2, rtol=0, atol=1e-5)
    np.testing.assert_allclose(a["dE"], a["dEdx"] * a["dx"], rtol=1e-5, atol=1e-6)
    assert set(np.unique(a["pdgId"])) <= {11, -11, 13, -13, 22, 211, -211, 2212}  # particle codes
    t0 = a["t0"]
    np.testing.assert_allclose(t0, (a["t0_start"] + a["t0_end"]) / 2, rtol=1e-6)  # float64 members too


def test_reader_refuses_other_files(tmp_path):
    from larndsim_amd.h5lite import read_dataset

    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file at all" * 4)
    with pytest.raises(ValueError):
        read_dataset(str(p), "segments")
    with pytest.raises(KeyError):
        read_dataset(FILES[0], "missing")


def test_load_segments_swaps_x_and_z():
    from larndsim_amd.dataio import load_segments
    from larndsim_amd.h5lite import read_dataset

    raw = read_dataset(FILES[0], "segments")
    t = load_segments(FILES[0])
    for a, b in (("x", "z"), ("x_start", "z_start"), ("x_end", "z_end")):
        assert np.array_equal(t[a], raw[b]) and np.array_equal(t[b], raw[a])
    assert np.array_equal(t["y"], raw["y"]) and t.dtype.names == FIELDS
    assert np.array_equal(load_segments(FILES[0], swap_xz=False)["x"], raw["x"])


def test_dataset_bookkeeping():
    """Event choice, event-grouped batches and step counts as dataio.py:225-330 computes them."""
    from larndsim_amd.dataio import TracksDataset, load_segments

    t = load_segments(FILES[0])
    # trajectories are numbered by their sorted (eventID, trackID) keys, so the events in order of
    # first appearance over trajectories are the ascending event ids
    first_events = np.unique(t["eventID"])
    ds = TracksDataset(FILES[0], nevents=4, max_batch_len=8, electron_sampling_resolution=0.1)
    got = np.concatenate(ds.get_batch_global_event_ids())
    assert np.array_equal(np.sort(got), first_events[:4])
    for b in range(len(ds)):
        rows = ds.get_batch_row_indices(b)
        assert set(np.unique(t[rows]["eventID"])) == set(ds.get_batch_global_event_ids(b))
        assert ds.batch_nsteps[b] == int(np.maximum(np.ceil(t[rows]["dx"] / 0.1), 1).sum())
    assert ds.max_batch_nsteps == max(ds.batch_nsteps)
    # every kept trajectory is at most max_batch_len long, and no event is split across batches
    ev_batches = [set(e) for e in ds.get_batch_global_event_ids()]
    assert sum(len(e) for e in ev_batches) == len(set().union(*ev_batches))
    a = TracksDataset(FILES[0], nevents=3, random_nevents=True, data_seed=7, max_batch_len=None)
    b = TracksDataset(FILES[0], nevents=3, random_nevents=True, data_seed=7, max_batch_len=None)
    assert [list(x) for x in a.get_batch_global_event_ids()] == [list(x) for x in b.get_batch_global_event_ids()]
