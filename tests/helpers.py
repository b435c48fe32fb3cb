"""Shared comparison helpers for the parity tests."""
import numpy as np

REC_DTYPE = np.dtype([("k0", "<i4"), ("uxy", "<i4"), ("ne", "<f4"), ("f", "<f4"), ("bx", "<i4"), ("by", "<i4"),
                      ("idx", "<i4"), ("la", "<f4"), ("lb", "<f4"), ("lc", "<f4"), ("wx", "<f4", 5), ("wy", "<f4", 5)])


def wfs_error(gpu, ref):
    """Max over rows of max|gpu - ref| / max|ref row| (rows with a zero reference use 1)."""
    scale = np.maximum(np.abs(ref).max(axis=1), 1e-30)
    err = np.abs(gpu - ref).max(axis=1)
    return float((err / scale).max()), float(np.abs(gpu - ref).max())


def segment_records(sim, st):
    """Per-segment records of a forward state, sorted back into input order."""
    import ctypes as C

    import torch

    from larndsim_amd import _lib

    n = st.n
    rec = torch.empty(n * 80, dtype=torch.uint8, device=sim.device)
    keys = torch.empty(n, dtype=torch.int32, device=sim.device)
    order = torch.empty(n, dtype=torch.int32, device=sim.device)
    _lib.check(sim.lib.lsim_segment_records(sim._ctx, n, st.n_events, C.c_void_p(st.ws.data_ptr()), st.ws.numel(),
                                            C.c_void_p(rec.data_ptr()), C.c_void_p(keys.data_ptr()),
                                            C.c_void_p(order.data_ptr()), sim._stream()))
    torch.cuda.synchronize()
    r = np.frombuffer(rec.cpu().numpy().tobytes(), dtype=REC_DTYPE)
    return r, keys.cpu().numpy().view(np.uint32), order.cpu().numpy().view(np.uint32)


def make_tracks(rows, fields=None):
    """(N, F) float32 segments from dicts of field values (unset fields 0)."""
    from larndsim_amd.synthetic import SEGMENT_FIELDS

    fields = fields or SEGMENT_FIELDS
    t = np.zeros((len(rows), len(fields)), np.float32)
    for i, r in enumerate(rows):
        for k, v in r.items():
            t[i, fields.index(k)] = v
    return t, fields


def mixed_event(params, n_c1=200):
    """C1 segments, their mirror images in TPC 1, border and out-of-TPC segments, padding rows
    (eventID -1 at the origin, as the reference's pad_batch leaves them)."""
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    tr = tr[:n_c1].copy()
    mir = tr[::3].copy()
    for k in ("z", "z_start", "z_end"):
        mir[:, f.index(k)] *= -1
    b = np.asarray(params.tpc_borders)
    extra, _ = make_tracks([
        dict(x=b[0, 0, 0] + 0.03, y=b[0, 1, 0] + 0.25, z=14.0, dE=1.0, dEdx=2.1, dx=0.4),      # pixel-plane corner
        dict(x=b[0, 0, 1] - 0.01, y=3.0, z=7.0, dE=1.0, dEdx=2.1, dx=0.4),                    # far x border
        dict(x=b[0, 0, 0] - 0.015, y=1.0, z=9.0, dE=1.0, dEdx=2.1, dx=0.4),                   # inside the margin, off-plane
        dict(x=2.0, y=2.0, z=31.5, dE=1.0, dEdx=2.1, dx=0.4),                                 # beyond the anode: no charge
        dict(x=2.0, y=2.0, z=30.5, dE=1.0, dEdx=2.1, dx=0.4),                                 # next to the anode
        dict(x=-3.0, y=4.0, z=0.17, dE=1.0, dEdx=2.1, dx=0.4),                                # next to the cathode
        dict(x=80.0, y=2.0, z=10.0, dE=1.0, dEdx=2.1, dx=0.4),                                # off every TPC and plane
    ], f)
    pad = np.zeros((5, len(f)), np.float32)
    pad[:, f.index("eventID")] = -1
    return np.concatenate([tr, mir, extra, pad]), f


def compare_hits(st, ref, tag=""):
    """The compacted hit arrays of a forward state against the oracle's ``parse_output`` (order
    (row, hit slot), sim_jax.py:619-647): ids, ticks and events bit-exact, hit_prob exact, ADC to
    1e-5 relative (1e-4 counts absolute), pixel x, y and z to 1e-5 cm (float32 of a float64 value).
    Rows whose hit ticks differ from the oracle's (threshold ties) are left out of the value
    comparison; their share of slots is at most 1e-3.  Returns the number of hits compared."""
    from larndsim_amd.sim import hits_dict

    H = st.out["adc"].shape[1]
    P = st.max_rows
    U = ref["unique_pixels"]
    valid = U >= 0
    n_inv = int((~valid).sum())
    assert valid[n_inv:].all(), "negative ids lead the oracle's rows"
    g = ref["grid"]
    ticks = st.out["ticks"][:P].cpu().numpy()
    same = ticks == g["ticks"][valid].astype(np.int64)
    assert same.mean() >= 1 - 1e-3, tag
    rows_ok = same.all(axis=1)
    # oracle hits: flat (row over U, slot) positions of its compaction mask, moved onto the P rows
    mask = (g["hit_prob"] > 1e-5) & (ref["unique_pixels"][:, None] >= 0)
    o_flat = np.nonzero(mask.ravel())[0] - n_inv * H
    h = {k: v.cpu().numpy() for k, v in hits_dict(st.out).items()}
    g_flat = h["hit_index"].astype(np.int64)
    assert np.all(np.diff(g_flat) > 0), f"{tag}: hits not in (row, slot) order"
    keep_g = rows_ok[g_flat // H]
    keep_o = rows_ok[o_flat // H]
    assert np.array_equal(g_flat[keep_g], o_flat[keep_o]), f"{tag}: compacted hit positions differ"
    n = int(keep_g.sum())
    for k in ("ticks", "event", "hit_pixels"):
        assert np.array_equal(h[k][keep_g].astype(np.int64), ref[k][keep_o].astype(np.int64)), f"{tag}: {k}"
    assert np.array_equal(h["hit_prob"][keep_g], ref["hit_prob"][keep_o].astype(np.float32)), f"{tag}: hit_prob"
    np.testing.assert_allclose(h["adcs"][keep_g], ref["adcs"][keep_o], rtol=1e-5, atol=1e-4, err_msg=f"{tag}: adcs")
    for k in ("pixel_x", "pixel_y", "pixel_z"):
        np.testing.assert_allclose(h[k][keep_g], ref[k][keep_o], rtol=0, atol=1e-5, err_msg=f"{tag}: {k}")
    print(f"[{tag}] {n} hits compared ({len(g_flat) - n} on tie rows)", flush=True)
    return n


def compare_forward(sim, params, lut, tracks, fields, tables=None, n_events=None, tag=""):
    """Run the HIP forward and the FP64 oracle on the same segments and check: pixel rows
    bit-exact, waveforms within 1e-5 of each row's peak, hit ticks exact on >= 99.9 % of slots
    and ADC within 1e-5 on rows whose ticks agree.  Returns (gpu state, oracle dict)."""
    import torch

    from oracle import oracle as O

    print(f"[gpu] forward {tag} n={len(tracks)}", flush=True)
    st = sim.forward(params, tracks, fields, n_events=n_events, want_tape=True)
    torch.cuda.synchronize()
    ref = O.simulate(params, lut, tracks, fields, tables=tables)
    U = ref["unique_pixels"]
    valid = U >= 0
    P = st.max_rows
    pids = st.out["pixel_ids"][:P].cpu().numpy()
    assert np.array_equal(pids, U[valid]), f"{tag}: pixel rows differ"
    if P == 0:
        return st, ref
    wfs = st.out["wfs"][:P].cpu().numpy().astype(np.float64)
    rel, ab = wfs_error(wfs, ref["wfs"][valid])
    print(f"[{tag}] rows={P} wfs rel err {rel:.3e} abs {ab:.3e}")
    assert rel < 1e-5, tag
    ticks = st.out["ticks"][:P].cpu().numpy()
    rt = ref["grid"]["ticks"][valid].astype(np.int64)
    same = ticks == rt
    assert same.mean() >= 1 - 1e-3, tag
    ok = same.all(axis=1)
    np.testing.assert_allclose(st.out["adc"][:P].cpu().numpy()[ok], ref["grid"]["adcs"][valid][ok], rtol=1e-5,
                               atol=1e-4, err_msg=tag)
    return st, ref
