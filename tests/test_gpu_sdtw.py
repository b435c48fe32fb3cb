"""Soft-DTW on the GPU (lsim_sdtw: tiled anti-diagonal wavefront) against the float64
restatement (oracle/sdtw.py): values and gradients on sequences that are and are not whole
numbers of 64-wide tiles, single values, and the sdtw losses with autograd."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,g,seed", [
    (200, 257, 1.0, 0), (257, 200, 0.1, 1), (64, 64, 1.0, 2), (65, 129, 0.5, 3), (1, 50, 1.0, 4), (50, 1, 1.0, 5),
    (1, 1, 1.0, 6), (700, 450, 1.0, 7),
])
def test_value_and_gradient(n, m, g, seed):
    from larndsim_amd.losses import _sdtw_hip
    from oracle import sdtw as SD

    rng = np.random.default_rng(seed)
    x = rng.normal(size=n).astype(np.float32)
    y = (rng.normal(size=m) + 0.3).astype(np.float32)
    v, gx = _sdtw_hip(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), g, True)
    rv, rg = SD.soft_dtw_grad(x.astype(np.float64), y.astype(np.float64), g)
    print("sdtw", n, m, g, float(v), rv, flush=True)
    assert float(v) == pytest.approx(rv, rel=2e-5, abs=1e-4)
    gx = gx.cpu().numpy()
    assert np.allclose(gx, rg, rtol=1e-2, atol=1e-3 * max(1.0, np.abs(rg).max()))


def test_sdtw_losses_autograd():
    from larndsim_amd import losses
    from oracle import sdtw as SD

    rng = np.random.default_rng(9)
    adcs = np.where(rng.uniform(size=300) < 0.3, 0.0, rng.uniform(70, 120, size=300)).astype(np.float32)
    ref = np.where(rng.uniform(size=280) < 0.3, 0.0, rng.uniform(70, 120, size=280)).astype(np.float32)
    ticks = rng.uniform(0, 30, size=300).astype(np.float32)
    ticks_ref = rng.uniform(0, 30, size=280).astype(np.float32)
    a = torch.from_numpy(adcs).cuda().requires_grad_(True)
    dstw = losses.SoftDTW(gamma=1.0)
    L, _ = losses.sdtw_time_adc(None, a, None, torch.from_numpy(ticks).cuda(), torch.from_numpy(ref).cuda(), None,
                                torch.from_numpy(ticks_ref).cuda(), dstw, alpha=0.5)
    L.backward()
    va, ga = SD.soft_dtw_grad(adcs[adcs > 0].astype(np.float64), ref[ref > 0].astype(np.float64), 1.0)
    vt, _ = SD.soft_dtw_grad(ticks[ticks > 0].astype(np.float64), ticks_ref[ticks_ref > 0].astype(np.float64), 1.0)
    print("sdtw_time_adc", float(L), 0.5 * va + 0.5 * vt, flush=True)
    assert float(L) == pytest.approx(0.5 * va + 0.5 * vt, rel=1e-5)
    want = np.zeros(300)
    want[adcs > 0] = 0.5 * ga
    got = a.grad.cpu().numpy()
    assert np.allclose(got, want, rtol=1e-2, atol=1e-3 * np.abs(want).max())
    # the debiased form: sdtw(x, y) - (sdtw(x, x) + sdtw(y, y)) / 2, gradient through both x terms
    x = torch.from_numpy(ref[ref > 0][:60].copy()).cuda().requires_grad_(True)
    yy = torch.from_numpy(adcs[adcs > 0][:70].copy()).cuda()
    d = losses.SoftDTW(gamma=1.0, debiased=True).pairwise(x, yy)
    d.backward()
    xs, ys = ref[ref > 0][:60].astype(np.float64), adcs[adcs > 0][:70].astype(np.float64)
    vxy, gxy = SD.soft_dtw_grad(xs, ys, 1.0)
    vxx, gxx = SD.soft_dtw_grad(xs, xs, 1.0)
    vyy, _ = SD.soft_dtw_grad(ys, ys, 1.0)
    assert float(d) == pytest.approx(vxy - 0.5 * (vxx + vyy), rel=1e-4, abs=1e-2)
    assert np.allclose(x.grad.cpu().numpy(), gxy - gxx, rtol=1e-2, atol=1e-3 * np.abs(gxy - gxx).max())
