"""The target-only MMD term is computed once per target cloud: a fit rebuilds the target's
position and weight tensors on every loss call, so the cache compares values, not identity
(CPU tensors, no GPU)."""
import torch


def test_target_term_reused_for_equal_values(monkeypatch):
    from larndsim_amd import losses as L

    calls = []
    real = L._target_term

    def counting(*a, **k):
        calls.append(1)
        return real(*a, **k)

    monkeypatch.setattr(L, "_target_term", counting)
    L._TARGET_CACHE.clear()
    g = torch.Generator().manual_seed(0)
    x, y = torch.rand((40, 3), generator=g, dtype=torch.float64), torch.rand((30, 3), generator=g, dtype=torch.float64)
    px, py = torch.rand(40, generator=g, dtype=torch.float64), torch.rand(30, generator=g, dtype=torch.float64)
    ex, ey = torch.zeros(40, dtype=torch.int64), torch.zeros(30, dtype=torch.int64)
    v1 = L.mmd_blockwise(x, y, px, py, 1.0, ex, ey)
    v2 = L.mmd_blockwise(x, y.clone(), px, py.clone(), 1.0, ex.clone(), ey.clone())  # new tensors, same values
    assert len(calls) == 1
    assert float(v1) == float(v2)
    L.mmd_blockwise(x, y + 1e-3, px, py, 1.0, ex, ey)  # another target
    L.mmd_blockwise(x, y, px, py, 2.0, ex, ey)  # another sigma
    assert len(calls) == 3
    direct = L.mmd(x, y, px, py, 1.0)
    assert abs(float(v1) - float(direct)) <= 1e-6 * abs(float(direct))  # float32 kernel values
