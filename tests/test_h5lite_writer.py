"""The HDF5 writer of larndsim_amd/h5lite.py (the hit files of optimize/simulate.py:135-155).

No HDF5 library is installed here, so a written file cannot be opened by h5py in these tests.
What pins it instead: the reader that decodes the reference's own h5py-written input files reads
every written tree back exactly (groups with more entries than one symbol-table node and one
B-tree node hold, so the multi-node and two-level B-trees are used); the written structures are
compared field by field with those of the reference's file (superblock version and node sizes,
the dataset header's message set, flags and encodings); and the B-tree's ordering rules, which
the library's name lookup relies on, are checked by walking the written file."""
import struct
from pathlib import Path

import numpy as np
import pytest

from larndsim_amd import h5lite

GOLDEN = Path(__file__).parent / "golden" / "segments_input_0.h5"


def _eq(a, b):
    if isinstance(a, dict):
        assert sorted(a) == sorted(b)
        for k in a:
            _eq(a[k], b[k])
    else:
        a = np.asarray(a)
        assert b.dtype == a.dtype and b.shape == (a.shape or (1,)) and np.array_equal(b.reshape(a.shape), a)


def _tree():
    rng = np.random.default_rng(3)
    return {
        "batch_0": {f"event_{i}": {"adc": rng.random(i % 7 + 1).astype(np.float32),
                                   "pixels": rng.integers(-5, 10**12, i % 5 + 1),
                                   "eventID": np.full(i % 5 + 1, i, np.int64)} for i in range(300)},
        "batch_1": {},
        "wfs": rng.random((3, 5)),
        "empty": np.zeros(0, np.float32),
        "small": {"u8": np.array([0, 7, 255], np.uint8), "i16": np.array([-3, 3], np.int16),
                  "u32": np.array([1, 2 ** 32 - 1], np.uint32), "scalar": np.float64(2.5)},
    }


def test_round_trip(tmp_path):
    t = _tree()
    p = tmp_path / "t.h5"
    h5lite.write_tree(p, t)
    _eq(t, h5lite.read_tree(p))
    assert np.array_equal(h5lite.read_dataset(p, "batch_0/event_299/eventID"), [299, 299, 299, 299, 299])


def test_structure_matches_the_reference_file(tmp_path):
    ref = h5lite._File(GOLDEN.read_bytes())
    p = tmp_path / "t.h5"
    h5lite.write_tree(p, {"segments": np.arange(168, dtype=np.float32)})
    f = h5lite._File(p.read_bytes())
    # superblock: version, offset and length sizes, group node sizes K, flags, root entry cache type
    assert f.d[8:16] == ref.d[8:16]
    assert f.u("HHI", 16) == ref.u("HHI", 16)
    assert f.u("I", 56 + 16) == ref.u("I", 56 + 16) == 1
    # the dataset header: same messages in the same order with the same flags
    a, b = f.find("segments"), ref.find("segments")
    assert [(mt, fl) for mt, fl, _, _ in f.messages(a)] == [(mt, fl) for mt, fl, _, _ in ref.messages(b)]
    fm = {mt: (q, n) for mt, _, q, n in f.messages(a)}
    rm = {mt: (q, n) for mt, _, q, n in ref.messages(b)}
    # dataspace (the reference's segments dataset also has 168 elements) and fill value: byte equal
    for mt in (0x1, 0x5):
        (q1, n1), (q2, n2) = fm[mt], rm[mt]
        assert f.d[q1:q1 + n1] == ref.d[q2:q2 + n2]
    # layout: version 3, contiguous, and the size of the data
    q, _ = fm[0x8]
    assert f.d[q:q + 2] == ref.d[rm[0x8][0]:rm[0x8][0] + 2] == bytes([3, 1])
    assert f.u("Q", q + 10) == 168 * 4
    # group B-tree and symbol-table node sizes as in the reference file (the library reads the full
    # node size given by K): the reference's root heap starts right after its root B-tree node
    assert ref.root_heap - ref.root_btree == 24 + (4 * 16 + 1) * 8
    assert f.d[f.root_btree:f.root_btree + 4] == b"TREE"


def _walk(f, node, names):
    """Entries of a group B-tree in order, checking each node's keys: key 0 is the empty name and
    key i + 1 the last name under child i."""
    d = f.d
    assert d[node:node + 4] == b"TREE" and d[node + 4] == 0
    level, used = d[node + 5], f.u("H", node + 6)
    left, right = f.u("QQ", node + 8)
    for sib, back in ((left, 16), (right, 8)):  # sibling links point back at this node
        assert sib == h5lite._UNDEF or f.u("Q", sib + back) == node
    heap_data = f.u("Q", names + 24)
    name = lambda off: d[heap_data + off:d.index(b"\0", heap_data + off)]  # noqa: E731
    p = node + 24
    assert name(f.u("Q", p)) == b""
    out = []
    for i in range(used):
        child, key = f.u("QQ", p + 8 + 16 * i)
        if level > 0:
            sub = _walk(f, child, names)
        else:
            assert d[child:child + 4] == b"SNOD"
            n = f.u("H", child + 6)
            assert 1 <= n <= 8
            sub = [name(f.u("Q", child + 8 + 40 * j)) for j in range(n)]
        assert name(key) == sub[-1]
        out += sub
    return out


def test_btree_ordering(tmp_path):
    p = tmp_path / "t.h5"
    names = [f"event_{i}" for i in range(1000)] + ["a", "Z", "event_", "batch_9"]
    h5lite.write_tree(p, {"g": {n: np.zeros(1, np.int8) for n in names}})
    f = h5lite._File(p.read_bytes())
    hdr = f.find("g")
    (bt, hp), = [f.u("QQ", q) for mt, _, q, _ in f.messages(hdr) if mt == 0x11]
    assert f.d[bt + 5] >= 1  # 1004 names need 126 symbol-table nodes: more than one B-tree node
    got = _walk(f, bt, hp)
    assert got == sorted(n.encode() for n in names)  # byte order, as the library's strcmp


def test_refusals(tmp_path):
    with pytest.raises(NotImplementedError):
        h5lite.write_tree(tmp_path / "c.h5", {"c": np.zeros(2, np.complex64)})
    with pytest.raises(ValueError):
        h5lite.write_tree(tmp_path / "s.h5", {"a/b": np.zeros(1)})
