"""The grid form of the event-blocked Gaussian-kernel sums (lsim_rbf_sums_grid, used by the MMD of
mse_adc and by nll_loss) against the span form that forms every pair of an event (lsim_rbf_sums):
the grid leaves out only pairs whose float32 kernel value is exactly zero, so the two differ only
by summation order.  Clouds spread over 200 cm with clusters, so most grid columns are culled."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _span_sums(x, ex, y, ey, c, sigma):
    """Every pair of each event through the span kernel (3-D callers: fourth coordinate zero)."""
    from larndsim_amd import losses

    ox, oy = torch.argsort(ex, stable=True), torch.argsort(ey, stable=True)
    exs, eys = ex[ox].contiguous(), ey[oy].contiguous()
    lo = torch.searchsorted(eys, exs, right=False)
    hi = torch.searchsorted(eys, exs, right=True)
    out = torch.empty((x.shape[0], 4), dtype=torch.float64, device=x.device)
    out[ox] = losses._rbf_sums_hip(x[ox], lo, hi, y[oy], c[oy], sigma)
    return out


def _cloud(g, sizes, spread):
    ev = np.repeat(np.arange(len(sizes)), sizes)
    n = len(ev)
    p = g.uniform(-spread, spread, (n, 4))
    # half the points in a few tight clusters (dense columns next to empty ones)
    k = n // 2
    centres = g.uniform(-spread, spread, (5, 4))
    p[:k] = centres[g.integers(0, 5, k)] + g.normal(0, 1.5, (k, 4))
    return torch.tensor(p, dtype=torch.float32), torch.tensor(ev)


@pytest.mark.parametrize("sigma,dims,sizes,ref_sizes", [
    (1.0, 3, [3000, 1, 2500, 700], [2800, 40, 0, 900]),
    (0.5, 3, [4000], [3500]),
    (3.0, 4, [1500, 1800], [1700, 1600]),
])
def test_grid_sums_equal_all_pair_sums(sigma, dims, sizes, ref_sizes):
    from larndsim_amd import losses

    g = np.random.default_rng(int(sigma * 10) + dims)
    x, ex = _cloud(g, sizes, 100.0)
    y, ey = _cloud(g, ref_sizes, 100.0)
    if dims == 3:
        x[:, 3] = 0.0
        y[:, 3] = 0.0
    c = torch.tensor(g.normal(size=(y.shape[0], 4)))
    x, ex, y, ey, c = (t.cuda() for t in (x, ex, y, ey, c))
    got = losses._block_sums(x, ex, y, ey, c, sigma, dims=dims)
    want = _span_sums(x, ex, y, ey, c, sigma)
    scale = _span_sums(x, ex, y, ey, c.abs(), sigma)  # sum of |terms| per point and column
    err = ((got - want).abs() / scale.clamp(min=1e-300)).max().item()
    print(f"[rbf-grid] sigma {sigma} dims {dims}: max error / sum|terms| {err:.2e}, "
          f"points with a non-zero sum {(scale[:, 0] > 0).float().mean().item():.3f}", flush=True)
    assert err < 4e-6
    # points whose event has no candidates sum to exactly zero in both forms
    assert torch.equal(got[scale[:, 0] == 0], want[scale[:, 0] == 0])


def test_grid_keeps_the_smallest_non_zero_terms():
    """A pair at exp(-80) (1.8e-35, still a normal float32) with a large weight is kept; the grid
    spans 400 cm so that pair lies in another cell column than its query."""
    from larndsim_amd import losses

    sigma = 1.0
    d = float(np.sqrt(160.0))  # d^2 / (2 sigma^2) = 80
    x = torch.tensor([[0.0, 0.0, 0.0, 0.0], [200.0, 200.0, 200.0, 0.0]], dtype=torch.float32)
    y = torch.tensor([[d, 0.0, 0.0, 0.0], [0.0, 0.0, 15.0, 0.0], [-200.0, -200.0, -200.0, 0.0]],
                     dtype=torch.float32)
    c = torch.tensor([[1e30, 1.0, 0.0, 0.0], [1e30, 0.0, 1.0, 0.0], [1.0, 0.0, 0.0, 1.0]], dtype=torch.float64)
    ex, ey = torch.zeros(2, dtype=torch.int64), torch.zeros(3, dtype=torch.int64)
    x, y, c, ex, ey = (t.cuda() for t in (x, y, c, ex, ey))
    got = losses._block_sums(x, ex, y, ey, c, sigma).cpu()
    want = _span_sums(x, ex, y, ey, c, sigma).cpu()
    print("[rbf-grid] tail", got[0].tolist(), want[0].tolist(), flush=True)
    assert got[0, 0] > 0 and torch.equal(got, want)
