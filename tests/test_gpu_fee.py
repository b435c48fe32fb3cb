"""The self-trigger readout on the GPU (lsim_fee_forward) on hand-built waveforms whose hits are
known in closed form (fee_jax.py:168-279), and on the oracle's waveforms of a real event."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CASES = [
    ([(100, 6000.0)], [99], [6000.0]),
    ([(100, 6000.0), (300, 8000.0)], [99, 299], [6000.0, 8000.0]),
    ([(100, 6000.0), (105, 2000.0)], [99], [8000.0]),
    ([(100, 6000.0), (119, 9000.0)], [99], [6000.0]),
    ([(100, 6000.0), (120, 9000.0)], [99, 119], [6000.0, 9000.0]),
    ([(100, 4000.0)], [], []),
    ([(0, 9000.0)], [], []),
    ([(1990, 6000.0)], [1989], [6000.0]),
    ([(1998, 6000.0)], [1997], [6000.0]),
    ([(50 + 40 * k, 6000.0 + k) for k in range(11)], [49 + 40 * k for k in range(10)], [6000.0 + k for k in range(10)]),
]


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


def test_hand_built_waveforms(sim, params4):
    from oracle import oracle as O

    H = params4.MAX_ADC_VALUES
    w = np.zeros((len(CASES), 2000), np.float32)
    for r, (spec, _, _) in enumerate(CASES):
        for t, q in spec:
            w[r, t] += np.float32(q / params4.t_sampling)
    pids = torch.arange(len(CASES), dtype=torch.int64) * 7 + 100
    print("[gpu] fee on hand-built rows", flush=True)
    out = sim.fee(params4, torch.from_numpy(w), pids)
    torch.cuda.synchronize()
    ticks = out["ticks"].cpu().numpy()
    adc = out["adc"].cpu().numpy()
    n_exp = 0
    for r, (_, tk, q) in enumerate(CASES):
        exp_t = tk + [1998] * (H - len(tk))
        exp_q = np.array(q + [0.0] * (H - len(q)))
        assert ticks[r].tolist() == exp_t, r
        np.testing.assert_allclose(adc[r], O.digitize(params4, exp_q), rtol=1e-6, err_msg=str(r))
        n_exp += sum(t < 1997 for t in tk)
    # compacted hits: tick < 1997 (hit_prob), in (row, k) order
    nh = int(out["n_hits"].item())
    assert nh == n_exp
    hp = out["hit_pixel"][:nh].cpu().numpy()
    ht = out["hit_tick"][:nh].cpu().numpy()
    exp = [(100 + 7 * r, t) for r, (_, tk, _) in enumerate(CASES) for t in tk if t < 1997]
    assert list(zip(hp.tolist(), ht.astype(int).tolist())) == exp


def test_fee_on_oracle_waveforms(sim, params4, lut, c1, tables):
    """Same decisions as the oracle's self-trigger when fed the oracle's own waveforms."""
    from oracle import oracle as O

    tr, f = c1
    ref = O.simulate(params4, lut, tr, f, tables=tables)
    U = ref["unique_pixels"]
    w = ref["wfs"].astype(np.float32)
    out = sim.fee(params4, torch.from_numpy(w), torch.from_numpy(U))
    torch.cuda.synchronize()
    a_ref, t_ref, _ = O.get_adc_values(params4, w.astype(np.float64))
    ticks = out["ticks"].cpu().numpy()
    same = ticks == t_ref
    print(f"[c1] fee tick mismatches {(~same).sum()} of {same.size}")
    assert same.mean() >= 1 - 1e-3
    ok = same.all(axis=1)
    np.testing.assert_allclose(out["adc"].cpu().numpy()[ok], O.digitize(params4, a_ref)[ok], rtol=1e-5, atol=1e-4)
    # the garbage row (-1) yields no compacted hit
    hp = out["hit_pixel"][: int(out["n_hits"].item())].cpu().numpy()
    assert (hp >= 0).all()


def test_rbf_sums_kernel_matches_torch():
    """lsim_rbf_sums (event-blocked Gaussian-kernel sums of the fit loss) against the chunked
    torch evaluation of the same sums on the CPU."""
    from larndsim_amd import losses

    g = torch.Generator().manual_seed(3)
    nx, ny = 3000, 2500
    ex = torch.randint(0, 7, (nx,), generator=g)
    ey = torch.randint(0, 7, (ny,), generator=g)
    x = torch.randn(nx, 3, generator=g, dtype=torch.float64) * 5
    y = torch.randn(ny, 3, generator=g, dtype=torch.float64) * 5
    c = torch.rand(ny, 4, generator=g, dtype=torch.float64)
    ref = torch.zeros(nx, 4, dtype=torch.float64)
    for e in range(7):
        ix, iy = (ex == e).nonzero().flatten(), (ey == e).nonzero().flatten()
        if len(ix) and len(iy):
            ref[ix] = torch.exp(-torch.cdist(x[ix], y[iy]).square() / (2 * 1.5 ** 2)) @ c[iy]
    got = losses._block_sums(x.cuda(), ex.cuda(), y.cuda(), ey.cuda(), c.cuda(), 1.5).cpu()
    torch.testing.assert_close(got, ref, rtol=2e-5, atol=1e-6)
    # the MMD with gradients on the GPU equals the CPU path
    px = torch.rand(nx, generator=g, dtype=torch.float64)
    py = torch.rand(ny, generator=g, dtype=torch.float64)
    xg = x.cuda().requires_grad_(True)
    pg = px.cuda().requires_grad_(True)
    mg = losses.mmd_blockwise(xg, y.cuda(), pg, py.cuda(), 1.5, ex.cuda(), ey.cuda())
    mg.backward()
    xc, pcpu = x.clone().requires_grad_(True), px.clone().requires_grad_(True)
    mc = losses.mmd_blockwise(xc, y, pcpu, py, 1.5, ex, ey)
    mc.backward()
    assert float(mg) == pytest.approx(float(mc), rel=1e-5)
    torch.testing.assert_close(xg.grad.cpu(), xc.grad, rtol=1e-4, atol=1e-8)
    torch.testing.assert_close(pg.grad.cpu(), pcpu.grad, rtol=1e-4, atol=1e-8)
