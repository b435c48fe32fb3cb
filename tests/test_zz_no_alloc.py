"""lsim_fee_forward_noalloc, whose scratch is a file-scope device array, against lsim_fee_forward
with a caller workspace, through the C ABI: two calls in a row on the same GPU, the second with
fewer rows (the array still holds the first call's values), every output bit for bit; and a row
count above LSIM_NOALLOC_MAX_ROWS is an error, not a launch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20  # LSIM_NOALLOC_MAX_ROWS


def _fee(sim, params, wfs, pids, noalloc):
    from larndsim_amd import _lib
    from larndsim_amd.sim import params_struct

    P = wfs.shape[0]
    out = sim.alloc_out(P, want_wfs=False)
    for k in ("adc", "hit_adc", "hit_x", "hit_z"):
        out[k].fill_(-7.0)
    oc = sim.out_struct(out)
    pc = params_struct(params)
    if noalloc:
        rc = sim.lib.lsim_fee_forward_noalloc(sim._ctx, C.byref(pc), C.c_void_p(wfs.data_ptr()),
                                              C.c_void_p(pids.data_ptr()), P, C.byref(oc), sim._stream())
    else:
        sz = C.c_size_t()
        _lib.check(sim.lib.lsim_fee_query(sim._ctx, P, C.byref(sz)))
        ws = torch.empty(int(sz.value), dtype=torch.uint8, device=sim.device)
        rc = sim.lib.lsim_fee_forward(sim._ctx, C.byref(pc), C.c_void_p(wfs.data_ptr()), C.c_void_p(pids.data_ptr()), P,
                                      C.c_void_p(ws.data_ptr()), ws.numel(), C.byref(oc), sim._stream())
    _lib.check(rc)
    torch.cuda.synchronize()
    return out


def _same(a, b, P):
    for k in ("adc", "ticks"):
        assert torch.equal(a[k][:P], b[k][:P]), k
    nh = int(a["n_hits"].item())
    assert nh == int(b["n_hits"].item()) and nh > 0
    for k in ("hit_adc", "hit_x", "hit_y", "hit_z", "hit_tick", "hit_prob", "hit_event", "hit_pixel", "hit_index"):
        assert torch.equal(a[k][:nh], b[k][:nh]), k
    return nh


def test_noalloc_equals_workspace_entry_point(params4, lut):
    from larndsim_amd import synthetic
    from larndsim_amd.sim import Simulator

    sim = Simulator(params4, lut)
    tr, f = synthetic.workload("c2")
    print("[gpu] waveforms of the 200k-segment event", flush=True)
    st = sim.forward(params4, tr, f, n_events=1)
    P = st.max_rows
    wfs = st.out["wfs"][:P].contiguous()
    pids = st.out["pixel_ids"][:P].contiguous()
    torch.cuda.synchronize()
    for n in (P, 1000):  # the second call has fewer rows than the first
        print(f"[gpu] no-allocation self-trigger on {n} rows", flush=True)
        a = _fee(sim, params4, wfs[:n], pids[:n], noalloc=True)
        b = _fee(sim, params4, wfs[:n], pids[:n], noalloc=False)
        nh = _same(a, b, n)
        print(f"[noalloc] {n} rows, {nh} hits equal", flush=True)


def test_noalloc_row_limit(params4, lut):
    from larndsim_amd.sim import Simulator, params_struct

    sim = Simulator(params4, lut)
    pc = params_struct(params4)
    rc = sim.lib.lsim_fee_forward_noalloc(sim._ctx, C.byref(pc), None, None, MAX_ROWS + 1, None, sim._stream())
    assert rc < 0
    assert b"LSIM_NOALLOC_MAX_ROWS" in sim.lib.lsim_last_error()
    assert np.int64(MAX_ROWS) > 0
