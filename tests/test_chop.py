"""Track chopping and batch assembly on the host side (no GPU): the oracle against the reference's
own chop_tracks outputs (tests/golden/chop_tracks_vectors.npz), its invariants, the batching
restatements, and the argument checks of the lsim_chop_* entry points."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from oracle import chop as oc  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden", "chop_tracks_vectors.npz")
CASES = ("p004", "p01", "p0001")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", CASES)
def test_oracle_equals_reference_output(gold, case):
    fields = tuple(str(f) for f in gold["fields"])
    got = oc.chop_tracks(gold[f"{case}_in"], fields, float(gold[f"{case}_precision"]))
    ref = gold[f"{case}_out"]
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_float32_positions_would_not_match(gold):
    """The golden vectors resolve the float64 step positions: rounding them in float32 instead (a
    plausible slip) changes some rows."""
    fields = tuple(str(f) for f in gold["fields"])
    t, prec = gold["p004_in"], float(gold["p004_precision"])
    good = oc.chop_tracks(t, fields, prec)
    f = list(fields)
    n = oc.step_counts(t, fields, prec)
    row = np.repeat(np.arange(len(t)), n)
    k = np.arange(len(row)) - np.concatenate([[0], np.cumsum(n)])[row]
    s = t[row, f.index("x_start")]
    seg = t[:, [f.index(c + "_end") for c in "xyz"]] - t[:, [f.index(c + "_start") for c in "xyz"]]
    ln = np.sqrt(((seg * seg)[:, 0] + (seg * seg)[:, 1]) + (seg * seg)[:, 2])
    d = seg[:, 0] / (ln + np.float32(1e-10))
    bad = s + (k.astype(np.float32) * np.float32(prec)) * d[row]
    assert not np.array_equal(bad, good[:, f.index("x_start")])


@pytest.mark.parametrize("case", CASES)
def test_chop_invariants(gold, case):
    fields = tuple(str(f) for f in gold["fields"])
    f = list(fields)
    t, prec = gold[f"{case}_in"], float(gold[f"{case}_precision"])
    out = oc.chop_tracks(t, fields, prec)
    seg = t[:, [f.index(c + "_end") for c in "xyz"]].astype(np.float64) - \
        t[:, [f.index(c + "_start") for c in "xyz"]].astype(np.float64)
    length = np.linalg.norm(seg, axis=1)
    n = oc.step_counts(t, fields, prec)
    # closed form of the step count, away from exact multiples of the step
    frac = length / prec - np.floor(length / prec)
    safe = (frac > 1e-3) & (frac < 1 - 1e-3)
    assert np.array_equal(n[safe], np.maximum(np.ceil(length[safe] / prec), 1).astype(np.int64))
    # zero length and shorter than a step give one step; a float32 length of 3 steps rounds to 3 or 4
    assert n[0] == 1 and n[1] in (3, 4) and n[2] == 1
    off = np.concatenate([[0], np.cumsum(n)])
    assert out.shape[0] == off[-1]
    for i in range(len(t)):
        r = out[off[i]:off[i + 1]]
        # the steps tile the track: dx sums to its length, dE to its dE, first start = track start,
        # last end = track end, each start equals the previous end
        assert abs(r[:, f.index("dx")].astype(np.float64).sum() - length[i]) <= 1e-5 + 1e-5 * length[i]
        assert abs(r[:, f.index("dE")].astype(np.float64).sum() - t[i, f.index("dE")]) <= 2e-5 * max(1, n[i])
        for c in "xyz":
            assert r[0, f.index(c + "_start")] == t[i, f.index(c + "_start")]
            assert r[-1, f.index(c + "_end")] == t[i, f.index(c + "_end")]
            assert np.array_equal(r[1:, f.index(c + "_start")], r[:-1, f.index(c + "_end")])
        assert np.all(r[:-1, f.index("dx")] == np.float32(prec))
        for name in ("eventID", "trackID", "pdgId", "dEdx"):
            assert np.all(r[:, f.index(name)] == t[i, f.index(name)])


def test_remap_and_pad_restatement():
    from larndsim_amd.synthetic import SEGMENT_FIELDS

    fields = tuple(SEGMENT_FIELDS)
    f = list(fields)
    rng = np.random.default_rng(3)
    a = rng.normal(size=(6, len(f))).astype(np.float32)
    src = np.array([40, 40, 7, 99, 7, 40])
    bat = np.array([7, 40, 99])
    r = oc.remap_event_ids(a, fields, src, bat)
    assert r[:, f.index("eventID")].tolist() == [1, 1, 0, 2, 0, 1]
    with pytest.raises(ValueError):
        oc.remap_event_ids(a, fields, np.array([7, 8, 7, 7, 7, 7]), bat)
    r[3, f.index("eventID")] = 5.0  # not a local id of a 3-event batch
    p = oc.pad_batch(r, fields, 9, n_local=3)
    assert p.shape == (9, len(f))
    for row in (3, 6, 7, 8):
        assert p[row, f.index("eventID")] == -1 and p[row, f.index("trackID")] == -1
        assert p[row, f.index("pixel_plane")] == -1 and p[row, f.index("dE")] == 0 and p[row, f.index("dx")] == 0
    assert np.all(p[6:, f.index("x")] == 0)
    assert p[3, f.index("x")] == r[3, f.index("x")]  # invalidation keeps the other columns
    assert np.array_equal(p[[0, 1, 2, 4, 5]], r[[0, 1, 2, 4, 5]])


@pytest.mark.parametrize("seed", range(6))
def test_group_events_matches_restatement(seed):
    from larndsim_amd.dataio import group_events_into_batches

    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    ev = rng.integers(0, max(1, n // 3), size=n) * 7 + 1000
    ln = rng.uniform(0, 30, size=n).astype(np.float32)
    max_len = float(rng.choice([20.0, 50.0, 100.0]))
    nb = [None, 0, 3][seed % 3]
    try:
        want = oc.group_events(ev, ln, max_len, nb)
    except ValueError:
        with pytest.raises(ValueError):
            group_events_into_batches(ev, ln, max_len, nb)
        return
    got = group_events_into_batches(ev, ln, max_len, nb)
    assert got[0] == want[0]
    assert got[1] == want[1]
    # no event is split across batches, and no trajectory longer than max_len is kept
    seen = {}
    for b, idx in enumerate(got[0]):
        assert idx
        for i in idx:
            assert ln[i] <= max_len
            assert seen.setdefault(ev[i], b) == b


def test_group_events_all_too_long():
    from larndsim_amd.dataio import group_events_into_batches

    with pytest.raises(ValueError):
        group_events_into_batches([1, 2], np.array([5.0, 6.0], np.float32), 1.0)


def test_chop_entry_point_argument_checks():
    from larndsim_amd import _lib
    from larndsim_amd.dataio import chop_layout
    from larndsim_amd.synthetic import SEGMENT_FIELDS

    lib = _lib.load()
    ws = C.c_size_t()
    assert lib.lsim_chop_query(1000, C.byref(ws)) == 0 and ws.value >= 8000
    assert lib.lsim_chop_query(-1, C.byref(ws)) < 0
    lay = chop_layout(SEGMENT_FIELDS)
    bad = chop_layout(SEGMENT_FIELDS)
    bad.dE = 99
    info = C.c_void_p(8)  # never dereferenced: every call below fails its host checks
    assert lib.lsim_chop_count(C.byref(bad), None, 0, 0.1, None, 0, info, None) < 0
    assert b"column" in lib.lsim_last_error()
    assert lib.lsim_chop_count(C.byref(lay), None, 0, 0.0, None, 0, info, None) < 0  # precision 0
    assert lib.lsim_chop_count(C.byref(lay), None, 10, 0.1, None, 0, info, None) < 0  # NULL tracks
    assert lib.lsim_chop_fill(C.byref(lay), None, 0, 0.1, None, 64, 5, 4, None, None, 0, -1, None, info, None) < 0
    assert b"n_steps" in lib.lsim_last_error()
