"""Track chopping and batch assembly on the GPU (lsim_chop_count / lsim_chop_fill) against the
reference's own chop_tracks outputs (tests/golden/chop_tracks_vectors.npz) and the oracle, bit for
bit, including empty input, single tracks, long tracks, the event remap and padding, and the
chopped rows fed straight into the simulation."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "chop_tracks_vectors.npz")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", ["p004", "p01", "p0001"])
def test_chop_equals_reference_output(gold, case):
    from larndsim_amd.dataio import chop_tracks

    fields = tuple(str(f) for f in gold["fields"])
    out = chop_tracks(gold[f"{case}_in"], fields, float(gold[f"{case}_precision"]), device="cuda:0")
    ref = gold[f"{case}_out"]
    print(case, tuple(out.shape), flush=True)
    assert tuple(out.shape) == ref.shape
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("n,lrange,prec,seed", [
    (20000, (0.0, 1.5), 0.04, 0),
    (3000, (0.0, 30.0), 0.1, 1),
    (1, (80.0, 80.0), 0.001, 2),     # one long track: 80 000 steps
    (5000, (0.0, 0.3), 0.001, 3),
])
def test_chop_equals_oracle(n, lrange, prec, seed):
    from larndsim_amd import synthetic
    from larndsim_amd.dataio import chop_tracks
    from oracle import chop as oc

    rng = np.random.default_rng(seed)
    t = synthetic.make_tracks(rng, n, lrange, n_events=4)
    fields = tuple(synthetic.SEGMENT_FIELDS)
    print("chop", n, lrange, prec, seed, flush=True)
    want = oc.chop_tracks(t, fields, prec)
    got = chop_tracks(torch.from_numpy(t).cuda(), fields, prec).cpu().numpy()
    assert got.shape == want.shape
    bad = np.flatnonzero(np.any(_bits(got) != _bits(want), axis=1))
    assert bad.size == 0, (bad[:5], got[bad[:1]], want[bad[:1]])


def test_chop_empty_and_degenerate():
    from larndsim_amd import synthetic
    from larndsim_amd.dataio import chop_batch, chop_tracks
    from oracle import chop as oc

    fields = tuple(synthetic.SEGMENT_FIELDS)
    f = list(fields)
    empty = np.zeros((0, len(f)), np.float32)
    assert tuple(chop_tracks(empty, fields, 0.1, device="cuda:0").shape) == (0, len(f))
    pad = chop_batch(empty, fields, 0.1, target_len=5, device="cuda:0").cpu().numpy()
    assert np.array_equal(_bits(pad), _bits(oc.pad_batch(empty, fields, 5)))
    # all-zero rows (zero length, zero dE) and a column order other than the generator's
    t = np.zeros((3, len(f)), np.float32)
    t[1, f.index("x_end")] = 0.35
    perm = np.random.default_rng(0).permutation(len(f))
    fp = tuple(fields[i] for i in perm)
    got = chop_tracks(t[:, perm], fp, 0.1, device="cuda:0").cpu().numpy()
    want = oc.chop_tracks(t[:, perm], fp, 0.1)
    assert np.array_equal(_bits(got), _bits(want))
    bad = t.copy()
    bad[2, f.index("x_end")] = np.inf
    with pytest.raises(ValueError):
        chop_tracks(bad, fields, 0.1, device="cuda:0")


def test_chop_batch_remap_and_padding():
    from larndsim_amd import synthetic
    from larndsim_amd.dataio import chop_batch
    from oracle import chop as oc

    rng = np.random.default_rng(5)
    fields = tuple(synthetic.SEGMENT_FIELDS)
    f = list(fields)
    t = synthetic.make_tracks(rng, 700, (0.0, 2.0), n_events=3)
    gids = np.array([1005, 2000, 31337])
    src = gids[t[:, f.index("eventID")].astype(int)]
    for target in (None, 1, 40000):
        want = oc.chop_tracks(oc.remap_event_ids(t, fields, src, gids), fields, 0.1)
        if target is not None and target > want.shape[0]:
            want = oc.pad_batch(want, fields, target, n_local=len(gids))
        got = chop_batch(t, fields, 0.1, target_len=target, source_event_ids=src, batch_event_ids=gids,
                         device="cuda:0").cpu().numpy()
        print("batch", target, got.shape, flush=True)
        assert np.array_equal(_bits(got), _bits(want))
    # rows of an event outside the batch's local ids are invalidated by the padding step
    t2 = t.copy()
    t2[:50, f.index("eventID")] = 7
    want = oc.pad_batch(oc.chop_tracks(t2, fields, 0.1), fields, 40000, n_local=3)
    got = chop_batch(t2, fields, 0.1, target_len=40000, batch_event_ids=gids, device="cuda:0").cpu().numpy()
    assert np.array_equal(_bits(got), _bits(want))
    assert np.all(got[:, f.index("eventID")] != 7)
    with pytest.raises(ValueError):
        chop_batch(t, fields, 0.1, source_event_ids=src, batch_event_ids=gids[:2], device="cuda:0")


def test_chopped_rows_feed_the_simulation(params4, lut):
    """Tracks chopped on the GPU run through the forward and give the same hits as the oracle's
    chopped rows (the chop output is the simulation's segment array, stride len(fields))."""
    from larndsim_amd import synthetic
    from larndsim_amd.dataio import chop_tracks
    from larndsim_amd.sim import Simulator
    from oracle import chop as oc

    rng = np.random.default_rng(9)
    fields = tuple(synthetic.SEGMENT_FIELDS)
    t = synthetic.make_tracks(rng, 40, (5.0, 20.0), n_events=2)
    sim = Simulator(params4, lut)
    dev = chop_tracks(t, fields, 0.04, device="cuda:0")
    host = torch.from_numpy(oc.chop_tracks(t, fields, 0.04)).cuda()
    a = sim.forward(params4, dev, fields, n_events=2)
    b = sim.forward(params4, host, fields, n_events=2)
    torch.cuda.synchronize()
    ha, hb = a.out, b.out
    assert int(ha["n_hits"].item()) == int(hb["n_hits"].item()) > 0
    for k in ("adc", "pixel_ids", "ticks"):
        assert torch.equal(ha[k], hb[k]), k
