"""Writes tests/golden/chop_tracks_vectors.npz: input tracks and the output of the reference's
``chop_tracks`` (optimize/dataio.py:63-106) on them, for three step lengths.

The reference module imports h5py, which is not installed here, so this script parses
optimize/dataio.py and executes only the ``chop_tracks`` function definition (pure NumPy) from
it.  Only the vectors are committed; run from the repository root:

    python tests/golden/make_chop_vectors.py
"""
import ast
import os
import sys

import numpy as np

REF = "/root/reference/optimize/dataio.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chop_tracks_vectors.npz")

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "larnd-sim-jax_amd"))
from larndsim_amd.synthetic import SEGMENT_FIELDS  # noqa: E402


def reference_chop():
    tree = ast.parse(open(REF).read())
    fn = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "chop_tracks"]
    assert len(fn) == 1
    mod = ast.Module(body=fn, type_ignores=[])
    scope = {"np": np}
    exec(compile(mod, REF, "exec"), scope)
    return scope["chop_tracks"]


def make_tracks(rng, n, max_len):
    f = list(SEGMENT_FIELDS)
    t = np.zeros((n, len(f)), np.float32)
    start = rng.uniform(-15, 15, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    length = rng.uniform(0, max_len, size=n)
    end = start + d * length[:, None]
    for k, c in enumerate("xyz"):
        t[:, f.index(f"{c}_start")] = start[:, k]
        t[:, f.index(f"{c}_end")] = end[:, k]
        t[:, f.index(c)] = 0.5 * (start[:, k] + end[:, k])
    t[:, f.index("dE")] = rng.uniform(0.01, 2.0, size=n)
    t[:, f.index("dEdx")] = rng.uniform(1.5, 5.0, size=n)
    t[:, f.index("dx")] = length
    t[:, f.index("eventID")] = np.sort(rng.integers(0, 3, size=n))
    t[:, f.index("trackID")] = np.arange(n)
    t[:, f.index("pdgId")] = 13
    return t


def main():
    chop = reference_chop()
    fields = tuple(SEGMENT_FIELDS)
    rng = np.random.default_rng(11)
    cases = {}
    for name, n, max_len, prec in (("p004", 40, 1.5, 0.04), ("p01", 40, 3.0, 0.1), ("p0001", 6, 0.05, 0.001)):
        t = make_tracks(rng, n, max_len)
        # edge cases: a zero-length track, a track of exactly 3 steps, a track shorter than one step
        t[0, [fields.index("x_end"), fields.index("y_end"), fields.index("z_end")]] = \
            t[0, [fields.index("x_start"), fields.index("y_start"), fields.index("z_start")]]
        t[1, fields.index("x_end")] = t[1, fields.index("x_start")] + np.float32(3 * prec)
        t[1, fields.index("y_end")] = t[1, fields.index("y_start")]
        t[1, fields.index("z_end")] = t[1, fields.index("z_start")]
        t[2, fields.index("z_end")] = t[2, fields.index("z_start")] - np.float32(0.3 * prec)
        t[2, fields.index("x_end")] = t[2, fields.index("x_start")]
        t[2, fields.index("y_end")] = t[2, fields.index("y_start")]
        out = chop(t.copy(), fields, precision=prec)
        cases[f"{name}_in"] = t
        cases[f"{name}_out"] = out.astype(np.float32)
        cases[f"{name}_precision"] = np.float64(prec)
        print(name, t.shape, "->", out.shape, out.dtype)
    np.savez_compressed(OUT, fields=np.array(fields), **cases)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
