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
