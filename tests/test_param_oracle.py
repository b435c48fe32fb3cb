"""The oracle of the parametrized current model (oracle/param.py) pinned on the CPU: the tick
integrals of the exponentially modified Gaussian tails against scipy.stats.exponnorm, the
no-diffusion tails against scipy.stats.expon, the transverse-bin weights against the normal CDF,
and charge conservation of the electrons and of the scatter into waveforms."""
import numpy as np
import pytest
from scipy import stats

from oracle import param as PA


def test_emg_tick_integrals_are_exponnorm_cdf_differences():
    """upper - lower of integrated_expon_diff (before its normalisation) is the EMG CDF over the tick:
    0.5 erf - emg_pdf / lambda = CDF - 1/2 (detsim_jax.py:441-476)."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        mu, sigma, scale = rng.uniform(-4, 0), rng.uniform(0.02, 0.6), rng.uniform(0.2, 3.0)
        lam = 1 / scale
        x = -np.linspace(0, 5, 51)
        dt = 0.1
        f = lambda X: 0.5 * PA.erf((X - mu) / (np.sqrt(2) * sigma)) - PA.emg_pdf(X, mu, sigma, lam) / lam  # noqa: E731
        ref = stats.exponnorm(1 / (sigma * lam), loc=mu, scale=sigma)
        got = f(x + dt / 2) - f(x - dt / 2)
        want = ref.cdf(x + dt / 2) - ref.cdf(x - dt / 2)
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-13)
        # and the normalised ticks are those differences over the same normaliser
        t = PA.integrated_expon_diff(x[None], mu, scale, sigma, dt)[0]
        norm = (ref.cdf(x[0] - dt / 2) - 0.5) - (ref.cdf(x[-1] + dt / 2) - 0.5)
        np.testing.assert_allclose(t, want / norm / dt, rtol=1e-9, atol=1e-13)


def test_no_diffusion_tails_are_exponential_cdf_differences():
    """integrated_expon: the tick integral of an exponential starting at loc, plus the constant share
    of the mass before the window (detsim_jax.py:534-540)."""
    for loc, scale in ((-3.0, 0.7), (-0.5, 2.0), (-4.2, 0.15)):
        x = -np.linspace(0, 5, 51)
        dt = 0.1
        got = PA.integrated_expon(x, loc, scale, dt)
        # X = loc + Exp(scale): P(x - dt/2 < X <= x + dt/2)
        ref = stats.expon(loc=loc, scale=scale)
        tick = ref.cdf(x + dt / 2) - ref.cdf(x - dt / 2)
        extra = np.exp(min(0.0, (loc - dt / 2) / scale)) / len(x)
        np.testing.assert_allclose(got * dt, tick + extra, rtol=1e-12, atol=1e-14)


def test_sigma_to_zero_limit_of_the_diffusion_tail():
    """As sigma -> 0 the EMG tail tends to the plain exponential of the same decay: the normalised
    diffusion ticks approach the exponential ticks over the same normaliser."""
    x = -np.linspace(0, 5, 51)
    dt, loc, scale = 0.1, -2.37, 0.8
    t = PA.integrated_expon_diff(x[None], loc, scale, 1e-5, dt)[0]
    cdf = stats.expon(loc=loc, scale=scale).cdf  # the EMG with sigma = 0: loc + Exp(scale)
    want = (cdf(x + dt / 2) - cdf(x - dt / 2)) / (cdf(x[0] - dt / 2) - cdf(x[-1] + dt / 2)) / dt
    np.testing.assert_allclose(t, want, rtol=1e-6, atol=1e-9)


def test_transverse_bin_weights():
    edges = np.array([-0.25, -0.15, -0.05, 0.05, 0.15, 0.25])
    std = np.array([0.01, 0.05, 0.2, 3.0])
    w = PA.gaussian_1d_integral(edges, std)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    cdf = stats.norm.cdf(edges[None, 1:-1] / std[:, None])
    want = np.diff(np.concatenate([np.zeros((4, 1)), cdf, np.ones((4, 1))], axis=1), axis=1)
    np.testing.assert_allclose(w, want, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(w, w[:, ::-1], rtol=1e-12, atol=1e-16)  # symmetric edges


@pytest.fixture(scope="module")
def small(params4):
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    return tr[:120], f


@pytest.mark.parametrize("mc,nn", [(True, 0), (True, 2), (False, 1)])
def test_electrons_and_scatter_conserve_charge(params4, small, mc, nn):
    from oracle import oracle as O

    tr, f = small
    edges = np.linspace(-0.2, 0.2, 6)
    p = params4.replace(mc_diff=mc, number_pix_neighbors=nn)
    normals = np.random.default_rng(3).normal(size=(len(tr), 3))
    e = PA.electrons(p, tr, f, normals=normals, bin_edges=edges)
    q = O.segment_quantities(p, tr, f)
    per = 1 if mc else 25
    assert len(e["x"]) == len(tr) * per
    np.testing.assert_allclose(e["ne"].reshape(len(tr), per).sum(axis=1), q["n_e"], rtol=1e-12)
    pids = PA.get_pixels(p, e)
    assert pids.shape == (len(e["x"]), (2 * nn + 1) ** 2)
    U = np.unique(pids)
    with np.errstate(all="ignore"):
        t0_tick, sig = PA.current_mc(p, e, pids)
    # a decay surface crossing zero sends lambda = 1 / scale to infinity: those few nearest-neighbour
    # pairs overflow (inf * 0) in float64 too; the central pixel never does
    fin = np.isfinite(sig).all(axis=2)
    K = (2 * nn + 1) ** 2
    assert fin[:, K // 2].all() and fin.mean() > 0.97, fin.mean()
    sig = np.where(fin[..., None], sig, 0.0)
    wfs = PA.accumulate(p, U, pids, t0_tick, sig)
    # every value lands in one tick: the waveforms (column 0 = the dropped out-of-range values) hold it all
    np.testing.assert_allclose(wfs.sum(), sig.sum(), rtol=1e-12)
    T = sig.shape[2]
    inside = (t0_tick[:, None] - T + np.arange(T)[None] >= 0) & (t0_tick[:, None] - T + np.arange(T)[None] < wfs.shape[1] - 1)
    np.testing.assert_allclose(wfs[:, 1:].sum(), (sig * inside[:, None, :]).sum(), rtol=1e-12)
