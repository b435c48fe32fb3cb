"""Real detector events from the reference's own input files (tests/golden/segments_input_*.h5)
through the GPU path: a TracksDataset batch (read, selected, chopped, remapped and padded on the
GPU) equals the oracle's NumPy chop of the same rows bit for bit, and the forward of that batch
matches the float64 oracle (pixel rows exact, waveforms to 1e-5 of each row's peak, hit ticks and
ADC; the response LUT is the seeded synthetic one, the reference's LUT file is not available)."""
import os

import numpy as np
import pytest
import torch

from helpers import compare_forward

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("name,chopped", [("segments_input_0.h5", True), ("segments_input_1.h5", True),
                                          ("segments_input_0.h5", False)])
def test_real_events_through_the_gpu_path(params4, lut, tables, name, chopped):
    from larndsim_amd.dataio import TracksDataset, np_from_structured
    from larndsim_amd.sim import Simulator
    from oracle import chop as oc

    ds = TracksDataset(os.path.join(GOLD, name), nevents=6, max_batch_len=60, electron_sampling_resolution=0.1,
                       chopped=chopped)
    fields = ds.get_track_fields()
    batch = ds[0]
    rows = ds.get_batch_row_indices(0)
    sel = ds.tracks_struct[rows]
    ids = ds.get_batch_global_event_ids(0)
    want = oc.remap_event_ids(np_from_structured(sel), fields, sel["eventID"], ids)
    if chopped:
        want = oc.chop_tracks(want, fields, 0.1)
    want = oc.pad_batch(want, fields, ds.max_batch_nsteps, n_local=len(ids))
    got = batch.cpu().numpy()
    print(f"[files] {name} chopped={chopped}: {len(ids)} events, batch {got.shape}", flush=True)
    assert np.array_equal(_bits(got), _bits(want))
    sim = Simulator(params4, lut)
    st, ref = compare_forward(sim, params4, lut, got, fields, tables=tables, n_events=len(ids), tag=f"file {name}")
    assert st.max_rows > 0
