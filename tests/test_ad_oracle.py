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
