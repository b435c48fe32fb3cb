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
