"""The chamfer and NLL fit losses on the GPU (lsim_nearest, lsim_rbf_sums over 4-D points),
event-blocked, against the dense restatements of the reference (oracle/losses.py, every
cross-event pair formed): values and gradients with respect to the simulated x, y, z, Q and
hit_prob, with the padding that pad_size adds."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _clouds(seed, sizes, ref_sizes):
    g = np.random.default_rng(seed)

    def cloud(sz):
        ev = np.repeat(np.arange(len(sz)), sz)
        n = len(ev)
        x, y = g.uniform(-15, 15, n), g.uniform(-15, 15, n)
        z = g.uniform(0, 30, n)
        # a few event-0 hits close to the origin, where the zero padding points sit
        k = min(5, sz[0])
        x[:k], y[:k], z[:k] = g.uniform(-0.5, 0.5, (3, k))
        q = g.uniform(5, 100, n)
        p = g.uniform(0, 1, n)
        return [torch.tensor(v, dtype=torch.float64) for v in (q, x, y, z, p)] + [torch.tensor(ev)]

    return cloud(sizes), cloud(ref_sizes)


@pytest.mark.parametrize("seed,sizes,ref_sizes", [
    (0, [40, 25, 60], [35, 30, 50]),
    (1, [300, 5, 200, 1], [250, 0, 220, 3]),
])
def test_chamfer_matches_dense_reference(seed, sizes, ref_sizes):
    from larndsim_amd import losses, sim
    from oracle import losses as OL

    (q, x, y, z, p, ev), (rq, rx, ry, rz, rp, rev) = _clouds(seed, sizes, ref_sizes)
    sim.size_history_dict.clear()
    dev = [t.cuda().requires_grad_(True) for t in (q, x, y, z)]
    L, _ = losses.chamfer_3d(None, dev[0], dev[1], dev[2], dev[3], None, p.cuda(), ev.cuda(), rq.cuda(), rx.cuda(),
                             ry.cuda(), rz.cuda(), None, rp.cuda(), rev.cuda())
    L.backward()
    new_ref, new_cur = sim.get_size_history()["chamfer_3d"][-1]
    sim.size_history_dict.clear()
    assert new_ref > len(rq) and new_cur > len(q)  # both sides padded: the origin points take part
    cpu = [t.clone().requires_grad_(True) for t in (q, x, y, z)]
    R = OL.chamfer_3d(cpu[0], cpu[1], cpu[2], cpu[3], ev, rq, rx, ry, rz, rev, new_cur, new_ref, dtype=torch.float64)
    R.backward()
    print("chamfer", seed, float(L), float(R), flush=True)
    assert float(L) == pytest.approx(float(R), rel=1e-5)
    for a, b, name in zip(dev, cpu, "qxyz"):
        ga, gb = a.grad.cpu(), b.grad
        assert torch.allclose(ga, gb, rtol=1e-4, atol=1e-6 * float(gb.abs().max())), name


@pytest.mark.parametrize("seed,sizes,ref_sizes,sigma", [
    (2, [40, 25, 60], [35, 30, 50], 1.0),
    (3, [300, 5, 200, 1], [250, 0, 220, 3], 2.5),
])
def test_nll_matches_dense_reference(seed, sizes, ref_sizes, sigma):
    from larndsim_amd import losses, sim
    from oracle import losses as OL

    (q, x, y, z, p, ev), (rq, rx, ry, rz, rp, rev) = _clouds(seed, sizes, ref_sizes)
    # shrink the clouds so that many kernel values are far from 0
    x, y, z, rx, ry, rz = (v * 0.2 for v in (x, y, z, rx, ry, rz))
    sim.size_history_dict.clear()
    dev = [t.cuda().requires_grad_(True) for t in (q, x, y, z, p)]
    L, _ = losses.nll_loss(None, dev[0], dev[1], dev[2], dev[3], None, dev[4], ev.cuda(), rq.cuda(), rx.cuda(),
                           ry.cuda(), rz.cuda(), None, rp.cuda(), rev.cuda(), sigma=sigma)
    L.backward()
    sim.size_history_dict.clear()
    cpu = [t.clone().requires_grad_(True) for t in (q, x, y, z, p)]
    R = OL.nll_loss(cpu[0], cpu[1], cpu[2], cpu[3], cpu[4], ev, rq, rx, ry, rz, rp, rev, sigma=sigma, pad=7,
                    dtype=torch.float64)
    R.backward()
    print("nll", seed, float(L), float(R), flush=True)
    assert float(L) == pytest.approx(float(R), rel=1e-5)
    for a, b, name in zip(dev, cpu, ["Q", "x", "y", "z", "hit_prob"]):
        ga, gb = a.grad.cpu(), b.grad
        assert torch.allclose(ga, gb, rtol=1e-4, atol=1e-5 * float(gb.abs().max())), name


def test_rbf_sums_fourth_coordinate():
    """lsim_rbf_sums measures distances over all four coordinates (zero for 3-D callers)."""
    from larndsim_amd import losses

    g = torch.Generator().manual_seed(0)
    xs = torch.rand((50, 4), generator=g)
    ys = torch.rand((70, 4), generator=g)
    c = torch.rand((70, 4), generator=g, dtype=torch.float64)
    lo = torch.zeros(50, dtype=torch.int32)
    hi = torch.full((50,), 70, dtype=torch.int32)
    got = losses._rbf_sums_hip(xs.cuda(), lo.cuda(), hi.cuda(), ys.cuda(), c.cuda(), 0.7).cpu()
    d = ((xs.double()[:, None, :] - ys.double()[None, :, :]) ** 2).sum(-1)
    want = torch.exp(-d / (2 * 0.49)) @ c.float().double()
    assert torch.allclose(got, want, rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("name", ["chamfer_3d", "nll"])
def test_named_loss_drives_the_adjoint(name, params4, lut, c1):
    """The fit driver's named losses on simulated hits: the loss gradient reaches the detector
    parameters through the simulation's adjoint (finite, and non-zero for Ab)."""
    from larndsim_amd import synthetic
    from larndsim_amd.losses import loss_functions
    from larndsim_amd.strategies import GenericLossStrategy, HipLUTSimulation

    tracks, fields = c1
    trk = torch.from_numpy(tracks).cuda()
    strat = HipLUTSimulation(lut)
    with torch.no_grad():
        target = strat.predict(synthetic.perturbed_params(params4, seed=1), trk, fields)
    fn, kw = loss_functions()[name]
    th = {n: torch.tensor(float(getattr(params4, n)), dtype=torch.float64, requires_grad=True)
          for n in ("Ab", "kb", "lifetime", "tran_diff")}
    p = params4.replace(**th)
    loss, _ = GenericLossStrategy(fn, **kw).compute(p, strat.predict(p, trk, fields), target)
    loss.backward()
    grads = {n: float(t.grad) if t.grad is not None else 0.0 for n, t in th.items()}
    print(name, float(loss), grads, flush=True)
    assert np.isfinite(float(loss)) and all(np.isfinite(v) for v in grads.values())
    assert grads["Ab"] != 0.0


def test_mse_adc_blockwise_keeps_float32_event_offset():
    """mse_adc_blockwise equals the dense mse_adc on float32 hits of several events, where the
    reference's x + event * 1e5 rounds x in float32 (to 1/128 cm at event 1, 1/16 cm at event 9)."""
    from larndsim_amd import losses

    (q, x, y, z, p, ev), (rq, rx, ry, rz, rp, rev) = _clouds(4, [30, 20, 25] + [10] * 7, [25, 30, 20] + [12] * 7)
    f32 = [t.to(torch.float32) for t in (q, x, y, z, p, rq, rx, ry, rz, rp)]
    q, x, y, z, p, rq, rx, ry, rz, rp = f32
    dense, _ = losses.mse_adc(None, q, x, y, z, None, p, ev, rq, rx, ry, rz, None, rp, rev)
    blk, _ = losses.mse_adc_blockwise(None, q.cuda(), x.cuda(), y.cuda(), z.cuda(), None, p.cuda(), ev.cuda(),
                                      rq.cuda(), rx.cuda(), ry.cuda(), rz.cuda(), None, rp.cuda(), rev.cuda())
    print("mse_adc", float(dense), float(blk), flush=True)
    assert float(blk) == pytest.approx(float(dense), rel=2e-4)
