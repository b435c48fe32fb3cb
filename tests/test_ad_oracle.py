"""Pins of the torch-float64 autograd oracle (CPU only): its forward equals the NumPy oracle,
and its gradient equals central finite differences of the same frozen-decision function."""
import numpy as np
import pytest
import torch

from oracle import ad_oracle as A


@pytest.fixture(scope="module")
def frozen(params4, lut, c1, tables):
    tr, f = c1
    return A.Frozen(params4, lut, tr[:120], f, tables=tables)


def test_forward_matches_numpy_oracle(frozen):
    with torch.no_grad():
        adc, z, wfs = frozen.forward(frozen.theta0())
    ref = frozen.out
    vr = frozen.valid_rows
    scale = np.abs(ref["wfs"]).max()
    np.testing.assert_allclose(wfs.numpy(), ref["wfs"], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(adc.numpy()[vr], ref["grid"]["adcs"][vr], rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(z.numpy()[vr], ref["grid"]["z"][vr], rtol=1e-12, atol=1e-12)


def test_gradient_matches_central_differences(frozen):
    vr = frozen.valid_rows
    hits = frozen.ticks[vr] < 1997
    assert hits.sum() > 5
    rng = np.random.default_rng(11)
    w_adc = rng.normal(size=hits.shape) * hits
    w_z = rng.normal(size=hits.shape) * hits
    g, _ = frozen.vjp(w_adc, w_z)
    th = frozen.theta0().numpy()
    for i in range(7):
        h = 1e-6 * abs(th[i])
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (frozen.value(tp, w_adc, w_z) - frozen.value(tm, w_adc, w_z)) / (2 * h)
        if i == 4:  # vdrift is never read by the simulation
            assert g[i] == 0.0 and fd == 0.0
            continue
        assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-9 * abs(fd) + 1e-12), A.GRAD_PARAMS[i]


def _wfs_loss(U, W, tU, tW):
    u, inv = np.unique(np.concatenate([U, tU]), return_inverse=True)
    a = torch.zeros((len(u), W.shape[1]), dtype=torch.float64).index_add(0, torch.as_tensor(inv[: len(U)]), W)
    b = torch.zeros_like(a).index_add(0, torch.as_tensor(inv[len(U):]), torch.as_tensor(tW))
    return ((a - b) ** 2).sum() / (b ** 2).sum()


def test_waveform_loss_central_differences(params4, lut, c1, tables):
    """A loss on the waveforms holds no hit decision.  Central differences of the whole chain (the
    NumPy oracle, decisions free) agree with the frozen chain's gradient at h = 1e-4 for every
    parameter that leaves the drift time alone.  eField moves every segment's drift time, and the
    waveform is linear in it between whole ticks (a kink at each tick crossing): differences over
    an interval holding crossings average the slopes on both sides, and shrinking h brings them to
    the gradient."""
    from oracle import oracle as O

    tr, f = c1
    tr = tr[:120]
    p = params4
    names = A.GRAD_PARAMS
    v0 = {n: float(getattr(p, n)) for n in names}
    scale = {"Ab": 1.02, "kb": 0.97, "eField": 1.01, "lifetime": 0.9, "vdrift": 1.0, "long_diff": 1.1,
             "tran_diff": 0.92}
    t = O.simulate(p.replace(**{n: v0[n] * scale[n] for n in names}), lut, tr, f, tables=tables)
    tU, tW = t["unique_pixels"], np.asarray(t["wfs"], np.float64)
    fr = A.Frozen(p, lut, tr, f, tables=tables)
    th = fr.theta0().requires_grad_(True)
    _wfs_loss(fr.U, fr.forward(th)[2], tU, tW).backward()
    g = th.grad.numpy()

    def fd(n, hr):
        h = hr * abs(v0[n])
        out = []
        for s in (1, -1):
            o = O.simulate(p.replace(**{n: v0[n] + s * h}), lut, tr, f, tables=tables)
            out.append(float(_wfs_loss(o["unique_pixels"], torch.as_tensor(np.asarray(o["wfs"], np.float64)), tU, tW)))
        return (out[0] - out[1]) / (2 * h)

    for n in ("Ab", "kb", "lifetime", "long_diff", "tran_diff"):
        i = names.index(n)
        assert g[i] == pytest.approx(fd(n, 1e-4), rel=1e-4), n
    i = names.index("eField")
    errs = [abs(fd("eField", hr) - g[i]) / abs(g[i]) for hr in (1e-3, 1e-6)]
    print("eField relative difference at h = 1e-3, 1e-6:", errs)
    assert errs[1] < 1e-3 and errs[1] < errs[0] / 5


def test_rows_subset_equals_full_chain(params4, lut, c1, tables):
    """Frozen over a subset of rows (for large events) returns the full chain's values on those
    rows, and the gradient of a loss on them equals the full chain's gradient of the same loss."""
    tr, f = c1
    tr = tr[:300]
    full = A.Frozen(params4, lut, tr, f, tables=tables)
    U = full.U
    rows = np.nonzero(U >= 0)[0]
    rng = np.random.default_rng(3)
    pick = np.sort(rng.choice(rows, size=12, replace=False))
    sub = A.Frozen(params4, lut, tr, f, tables=tables, rows_subset_pids=U[pick])
    assert np.array_equal(sub.U, U[pick])
    assert np.array_equal(sub.ticks, full.ticks[pick])
    th = full.theta0()
    with torch.no_grad():
        a_f, z_f, w_f = full.forward(th)
        a_s, z_s, w_s = sub.forward(th)
    scale = float(w_f.abs().max())
    np.testing.assert_allclose(w_s.numpy(), w_f.numpy()[pick], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(a_s.numpy(), a_f.numpy()[pick], rtol=1e-11, atol=1e-10)
    np.testing.assert_allclose(z_s.numpy(), z_f.numpy()[pick], rtol=1e-13, atol=1e-13)
    hits = sub.ticks < 1997
    assert hits.sum() > 3
    w_adc = rng.normal(size=hits.shape) * hits
    w_z = rng.normal(size=hits.shape) * hits
    g_s, _ = sub.vjp(w_adc, w_z)
    W_adc = np.zeros(full.ticks.shape)
    W_z = np.zeros(full.ticks.shape)
    W_adc[pick], W_z[pick] = w_adc, w_z
    vr = full.valid_rows
    g_f, _ = full.vjp(W_adc[vr], W_z[vr])
    np.testing.assert_allclose(g_s, g_f, rtol=1e-10, atol=1e-14)
