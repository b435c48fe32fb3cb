"""Pins of the FP64 oracle (CPU only).

Every function of ``oracle.oracle`` is checked here against something other than its own
formula: the known answers of the reference's fixture ``output/jax_ref/output_lut_1.npz``
(extracted to ``tests/golden/output_lut_1_known.npz``), closed forms, geometric brute force,
an independent library routine, hand-built inputs with known outcomes, or the independent
scalar loop oracle on small events.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "output_lut_1_known.npz")


@pytest.fixture(scope="module")
def known():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="module")
def p1():
    from larndsim_amd import consts

    return consts.default_params(RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)


def _tracks(rows, fields=None):
    """(N, F) float32 segments from dicts of x, y, z, dE, dEdx, eventID (other fields 0)."""
    from larndsim_amd.synthetic import SEGMENT_FIELDS

    fields = fields or SEGMENT_FIELDS
    t = np.zeros((len(rows), len(fields)), np.float32)
    for i, r in enumerate(rows):
        for k, v in r.items():
            t[i, fields.index(k)] = v
    return t, fields


# ------------------------------------------------------------------------------------------
# known answers of the reference fixture (fee_jax.py:55-71, losses_jax.py:380-383,
# detsim_jax.py:296-319, consts_jax.py:192-216)
# ------------------------------------------------------------------------------------------

def test_digitize_zero_known_answer(p1, known):
    """Entries without a hit carry digitize(0): 75.82347 in the fixture."""
    d0 = np.float32(O.digitize(p1, 0.0))
    assert d0 == np.float32(75.82347)
    assert d0 in known["adcs"]


def test_digitize_clamps(p1):
    assert O.digitize(p1, -1e9) == 0.0
    assert O.digitize(p1, 1e9) == p1.ADC_COUNTS
    # strictly linear in between: slope GAIN * ADC_COUNTS / (V_REF - V_CM)
    q = np.array([1e4, 2e4])
    d = O.digitize(p1, q)
    assert d[1] - d[0] == pytest.approx(1e4 * p1.GAIN * p1.ADC_COUNTS / (p1.V_REF - p1.V_CM), rel=1e-12)


def test_adc2charge_known_answers(p1, known):
    """Q stored in the fixture is adc2charge of the stored ADC (float32 arithmetic there)."""
    q = O.adc2charge(known["adcs"].astype(np.float64), p1)
    np.testing.assert_allclose(q, known["Q"], rtol=1e-6, atol=2e-6)


def test_adc2charge_inverts_digitize(p1):
    q = np.linspace(0, 40e3, 11)
    np.testing.assert_allclose(O.adc2charge(O.digitize(p1, q), p1) * 1e3, q, rtol=1e-12, atol=1e-7)


def test_threshold_adc_absent_from_fixture(p1, known):
    """simulate.py:129-131 drops hits whose ADC equals digitize(threshold)."""
    assert np.float32(O.digitize(p1, p1.DISCRIMINATION_THRESHOLD)) not in known["adcs"]


def test_pixel_centres_known_answers(p1, known):
    """Every pixel x / y of the fixture is a pixel centre of one of the two TPCs."""
    b = np.asarray(p1.tpc_borders)
    for key, ax, n in (("x", 0, p1.n_pixels_x), ("y", 1, p1.n_pixels_y)):
        v = known[key].astype(np.float64)
        ok = np.zeros(len(v), bool)
        for t in range(b.shape[0]):
            u = np.rint((v - b[t, ax, 0]) / p1.pixel_pitch - 0.5)
            X, Y = O.get_pixel_coordinates(p1, u if ax == 0 else 0 * u, u if ax == 1 else 0 * u, np.full(len(u), t))
            c = X if ax == 0 else Y
            ok |= (np.abs(c - v) < 2e-5) & (u >= 0) & (u < n)
        assert ok.all(), key
    X, _ = O.get_pixel_coordinates(p1, np.array([4]), np.array([0]), np.array([0]))
    assert np.float32(X[0]) == np.float32(-29.0427)
    assert np.float32(-0.22170042) in known["x"]


def test_hit_z_known_answers(p1, known):
    """z = z_anode + tick * dt * v(E) * sign(cathode - anode) for one of the planes."""
    t = known["ticks"].astype(np.float64)
    z0 = O.get_hit_z(p1, t, np.zeros(len(t), np.int64))
    z1 = O.get_hit_z(p1, t, np.ones(len(t), np.int64))
    err = np.minimum(np.abs(z0 - known["z"]), np.abs(z1 - known["z"]))
    assert err.max() < 1e-5
    # the reference evaluates in float32 (ulp of the 31.9 cm drift term is 3.8e-6)
    assert O.get_hit_z(p1, 1998.0, 0) == pytest.approx(-1.3073711, abs=2e-6)


def test_vdrift_known_answer(p1):
    """get_vdrift at 0.5 kV/cm and 87.17 K (the vdrift_static default 0.159645)."""
    assert O.vdrift(p1) == pytest.approx(0.1596452, abs=5e-8)


# ------------------------------------------------------------------------------------------
# quench and drift closed forms (quenching_jax.py:18-75, drifting_jax.py:18-58)
# ------------------------------------------------------------------------------------------

def test_birks_closed_form(p1):
    dEdx, dE = np.array([2.1, 5.0]), np.array([0.4, 1.0])
    assert np.allclose(O.quench(p1.replace(kb=0.0), dEdx, dE), p1.Ab * dE * p1.MeVToElectrons, rtol=1e-15)
    # R(dEdx) is 1/(1+x) shaped: R(0) = Ab, and the half-value point is at kb*dEdx = eField*rho
    half = p1.eField * p1.lArDensity / p1.kb
    r = O.quench(p1, np.array([half]), np.array([1.0])) / p1.MeVToElectrons
    assert r[0] == pytest.approx(p1.Ab / 2, rel=1e-14)


def test_box_closed_form(p1):
    from larndsim_amd.consts import RecombinationMode

    p = p1.replace(recombination_mode=RecombinationMode.BOX)
    # csi = beta dEdx / (E rho); choose dEdx so that csi = e - alpha -> R = 1 / csi
    csi = math.e - p.alpha
    dEdx = csi * p.eField * p.lArDensity / p.beta
    r = O.quench(p, np.array([dEdx]), np.array([1.0])) / p.MeVToElectrons
    assert r[0] == pytest.approx(1.0 / csi, rel=1e-13)
    # log(alpha + csi) < 0 is clamped to zero recombination
    dEdx0 = 1e-3 * p.eField * p.lArDensity / p.beta
    assert O.quench(p, np.array([dEdx0]), np.array([1.0]))[0] == 0.0


def test_ellipsoid_reduces_to_box_along_field(p1):
    from larndsim_amd.consts import RecombinationMode

    pb = p1.replace(recombination_mode=RecombinationMode.BOX)
    pe = p1.replace(recombination_mode=RecombinationMode.ELLIPSOID)
    dEdx = np.array([2.1, 4.0])
    # cos(phi) = 0: b_phi = beta, only the 1e-10 regulariser differs
    np.testing.assert_allclose(O.quench(pe, dEdx, np.ones(2), cosphi=np.zeros(2)), O.quench(pb, dEdx, np.ones(2)),
                               rtol=1e-9)
    # cos(phi) = 1: b_phi = beta * R_param
    pr = pb.replace(beta=p1.beta * p1.R_param)
    np.testing.assert_allclose(O.quench(pe, dEdx, np.ones(2), cosphi=np.ones(2)), O.quench(pr, dEdx, np.ones(2)),
                               rtol=1e-9)


def test_drift_closed_form(p1):
    b = np.asarray(p1.tpc_borders)
    x = np.array([0.0, 1.0, 0.0, 100.0])
    y = np.array([0.0, 2.0, 0.0, 0.0])
    z = np.array([10.0, 30.0, -5.0, 10.0])
    plane, mask, ne, td, sl, st = O.drift(p1, x, y, z, np.full(4, 1000.0))
    assert plane.tolist()[:3] == [0, 0, 1]
    assert mask.tolist() == [True, True, True, False]
    v = O.vdrift(p1)
    td_ref = (np.abs(z - b[plane, 2, 0]) + 1e-6) / v
    np.testing.assert_allclose(td, td_ref, rtol=1e-15)
    np.testing.assert_allclose(ne[:3], 1000.0 * np.exp(-td_ref[:3] / p1.lifetime), rtol=1e-14)
    assert ne[3] == 0.0
    np.testing.assert_allclose(sl ** 2 / st ** 2, p1.long_diff / p1.tran_diff, rtol=1e-12)


# ------------------------------------------------------------------------------------------
# transverse weights, sub-bins and pixels (detsim_jax.py:230-354, 493-512; sim_jax.py:394-452)
# ------------------------------------------------------------------------------------------

def test_erf_weights_sum_and_limits(p1):
    edges = O.bin_edges(p1)
    w = p1.pixel_pitch / p1.nb_sampling_bins_per_pixel
    x0 = np.linspace(0, w, 7, endpoint=False)
    W = O.gaussian_1d_integral(edges, x0, np.full(7, 0.05))
    np.testing.assert_allclose(W.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    # sigma -> 0: all weight in the centre bin [0, w)
    W0 = O.gaussian_1d_integral(edges, x0 + 1e-9, np.full(7, 1e-12))
    np.testing.assert_allclose(W0[:, 2], 1.0, atol=1e-15)
    # symmetric placement (segment at the bin centre) gives mirror-symmetric weights
    Ws = O.gaussian_1d_integral(edges, np.array([w / 2]), np.array([0.07]))[0]
    np.testing.assert_allclose(Ws, Ws[::-1], rtol=1e-12)
    # huge sigma: the two outer bins take (almost) everything, half each
    Wb = O.gaussian_1d_integral(edges, np.array([w / 2]), np.array([1e6]))[0]
    assert Wb[0] == pytest.approx(0.5, abs=1e-6) and Wb[-1] == pytest.approx(0.5, abs=1e-6)


def test_pixel_id_bijection(p1):
    nx, ny, ntpc = p1.n_pixels_x, p1.n_pixels_y, 2
    rng = np.random.default_rng(0)
    px, py = rng.integers(0, nx, 1000), rng.integers(0, ny, 1000)
    pl, ev = rng.integers(0, ntpc, 1000), rng.integers(0, 50, 1000)
    pid = O.pixel2id(p1, px, py, pl, ev)
    assert len(np.unique(pid)) == len(np.unique(np.stack([px, py, pl, ev]), axis=1).T)
    X, Y, PL, EV = O.id2pixel(p1, pid)
    assert np.array_equal(X, px) and np.array_equal(Y, py) and np.array_equal(PL, pl) and np.array_equal(EV, ev)
    # consecutive ids walk x fastest, then y, plane, event
    assert O.pixel2id(p1, 1, 0, 0, 0) == 1 and O.pixel2id(p1, 0, 1, 0, 0) == nx
    assert O.pixel2id(p1, 0, 0, 1, 0) == nx * ny and O.pixel2id(p1, 0, 0, 0, 1) == nx * ny * ntpc
    for bad in ((-1, 0), (nx, 0), (0, -1), (0, ny)):
        assert O.pixel2id(p1, bad[0], bad[1], 0, 0) == -1


def test_subbin_lut_index_is_geometric_distance(p1):
    """cx of the centre bin and of every neighbour = floor(|segment sub-bin centre - pixel
    centre| / sub-bin width): the table {4,3,2,1,0,0,1,2,3,4} and its neighbour extension."""
    p = p1.replace(number_pix_neighbors=2)
    b = np.asarray(p.tpc_borders)
    nsb = p.nb_sampling_bins_per_pixel
    w = p.pixel_pitch / nsb
    rows = []
    for s in range(nsb):
        rows.append(dict(x=b[0, 0, 0] + 30 * p.pixel_pitch + (s + 0.5) * w, y=b[0, 1, 0] + 40.5 * p.pixel_pitch,
                         z=10.0, dE=1.0, dEdx=2.0, eventID=0))
    tr, f = _tracks(rows)
    q = O.segment_quantities(p, tr, f)
    assert q["cx"][:, 2, 2].tolist() == [4, 3, 2, 1, 0, 0, 1, 2, 3, 4]
    n = p.number_pix_neighbors
    a_off = np.repeat(np.arange(-n, n + 1), 2 * n + 1)
    xs = b[0, 0, 0] + 30 * p.pixel_pitch + (np.arange(nsb) + 0.5) * w
    xc = b[0, 0, 0] + (30 + a_off[None, :] + 0.5) * p.pixel_pitch
    expect = np.floor(np.abs(xs[:, None] - xc) / w + 1e-9).astype(np.int64)
    assert np.array_equal(q["cx_nb"], expect)
    # every diffusion bin's pixel is the pixel containing that sub-bin
    sx = q["bx"][:, None] + np.arange(-2, 3)[None, :]
    assert np.array_equal(O.id2pixel(p, q["pid"][:, :, 2])[0], np.floor_divide(sx, nsb))


def test_neighbour_grid_brute_force(p1):
    p = p1.replace(number_pix_neighbors=1)
    b = np.asarray(p.tpc_borders)
    tr, f = _tracks([dict(x=b[0, 0, 0] + 0.2, y=b[0, 1, 0] + 5.0, z=10.0, dE=1.0, dEdx=2.0, eventID=3)])
    q = O.segment_quantities(p, tr, f)
    mx, my, _, _ = O.id2pixel(p, q["main_pid"])
    got = q["pid_nb"][0].tolist()
    want = []
    for a in (-1, 0, 1):
        for c in (-1, 0, 1):
            if a == 0 and c == 0:
                want.append(-999)
            elif mx[0] + a < 0:
                want.append(-1)
            else:
                want.append(int(O.pixel2id(p, mx[0] + a, my[0] + c, 0, 3)))
    assert got == want


def test_template_interpolation_weights(p1):
    """Lagrange weights sum to one and reproduce a quadratic in sigma exactly."""
    tr, f = _tracks([dict(x=0.1, y=0.2, z=z, dE=1.0, dEdx=2.0, eventID=0) for z in (1.0, 8.0, 20.0, 29.0)])
    q = O.segment_quantities(p1, tr, f)
    s = O._sigmas(p1)
    np.testing.assert_allclose(q["la"] + q["lb"] + q["lc"], 1.0, rtol=1e-12)
    i = q["idx"]
    quad = lambda x: 3 * x * x - 2 * x + 0.5  # noqa: E731
    np.testing.assert_allclose(q["la"] * quad(s[i - 1]) + q["lb"] * quad(s[i]) + q["lc"] * quad(s[i + 1]),
                               quad(q["lam"]), rtol=1e-9)
    assert ((s[i - 1] < q["lam"]) & (q["lam"] <= s[i])).all()


# ------------------------------------------------------------------------------------------
# templates (consts_jax.py:387-449)
# ------------------------------------------------------------------------------------------

def test_templates_against_numpy_convolve(p1, lut):
    tab = O.build_templates(p1, lut)
    g = O.gaussian_kernels(p1)
    np.testing.assert_allclose(g.sum(axis=1), 1.0, rtol=1e-13)
    assert np.array_equal(tab["T_main"][0], lut[:5, :5].astype(np.float64))
    for m, cx, cy in ((1, 0, 0), (23, 4, 1), (99, 2, 3)):
        ref = np.convolve(lut[cx, cy].astype(np.float64), g[m], mode="same")
        np.testing.assert_allclose(tab["T_main"][m, cx, cy], ref, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(tab["C_R"][7, 9], np.cumsum(lut[7, 9].astype(np.float64)), rtol=1e-13)


# ------------------------------------------------------------------------------------------
# signals (sim_jax.py:141-286, 689-736)
# ------------------------------------------------------------------------------------------

def _small_event(p):
    """Eight C1 segments plus hand-placed ones: a pixel-plane edge, TPC 1, outside the TPCs,
    near the anode (early arrival) and far from it (late arrival / window clipping)."""
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    b = np.asarray(p.tpc_borders)
    extra, _ = _tracks([
        dict(x=b[0, 0, 0] + 0.05, y=b[0, 1, 0] + 0.3, z=12.0, dE=0.8, dEdx=2.0, eventID=0),
        dict(x=3.1, y=-2.2, z=-12.0, dE=0.8, dEdx=2.0, eventID=0),
        dict(x=80.0, y=0.0, z=12.0, dE=0.8, dEdx=2.0, eventID=0),
        dict(x=1.0, y=1.0, z=30.5, dE=0.8, dEdx=2.0, eventID=0),
        dict(x=-1.0, y=1.0, z=0.2, dE=0.8, dEdx=2.0, eventID=0),
    ], f)
    return np.concatenate([tr[::125], extra]), f


def test_vectorised_matches_loop_oracle(p1, lut):
    p = p1.replace(number_pix_neighbors=2)
    tr, f = _small_event(p)
    wfs, U, _ = O.simulate_wfs(p, lut, tr, f)
    wl, Ul = O.simulate_wfs_loop(p, lut, tr, f)
    assert np.array_equal(U, Ul)
    scale = np.abs(wl).max()
    np.testing.assert_allclose(wfs, wl, rtol=0, atol=1e-11 * scale)


def test_linearity_in_charge(p1, lut, tables):
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    tr = tr[:40].copy()
    w1, U1, _ = O.simulate_wfs(p1, lut, tr, f, tables=tables)
    tr2 = tr.copy()
    tr2[:, f.index("dE")] *= 2
    w2, U2, _ = O.simulate_wfs(p1, lut, tr2, f, tables=tables)
    assert np.array_equal(U1, U2)
    np.testing.assert_allclose(w2, 2 * w1, rtol=1e-12, atol=1e-12 * np.abs(w1).max())


def test_permutation_invariance(p1, lut, tables):
    from larndsim_amd import synthetic

    tr, f = synthetic.workload("c1")
    tr = tr[:60]
    w1, U1, _ = O.simulate_wfs(p1, lut, tr, f, tables=tables)
    perm = np.random.default_rng(3).permutation(len(tr))
    w2, U2, _ = O.simulate_wfs(p1, lut, tr[perm], f, tables=tables)
    assert np.array_equal(U1, U2)
    np.testing.assert_allclose(w1, w2, rtol=0, atol=1e-12 * np.abs(w1).max())


def test_charge_conservation_identity(p1, lut, tables):
    """With no clipped column, the total over one segment's waveforms equals, per stripe,
    n * (sum_{t=k0+1}^{Nt-1} T[t] + T[Nt-L] - f T[k0+1]) computed from the raw template
    values (no cumulative tables): the window + boundary-correction construction."""
    p = p1.replace(number_pix_neighbors=2)
    tr, f = _tracks([dict(x=0.37, y=-1.21, z=9.0, dE=0.9, dEdx=2.2, eventID=0)])
    wfs, U, q = O.simulate_wfs(p, lut, tr, f, tables=tables)
    nt, L = lut.shape[-1], p.signal_length
    k0 = int(np.floor(q["tau"][0]))
    fr = q["tau"][0] - k0
    assert 2 <= nt - L - k0 - 2 and nt - L - k0 + L < p.n_ticks - 2  # nothing clipped
    raw = lut.astype(np.float64)

    def stripe_sum(T):
        return T[k0 + 1:].sum() + T[nt - L] - fr * T[k0 + 1]

    total = 0.0
    i, la, lb, lc = q["idx"][0], q["la"][0], q["lb"][0], q["lc"][0]
    g = O.gaussian_kernels(p)
    for a in range(5):
        for c in range(5):
            if not np.isin(q["pid"][0, a, c], U):
                continue
            cx, cy = q["cx"][0, a, c], q["cy"][0, a, c]
            Tm = [np.convolve(raw[cx, cy], g[m], mode="same") if m else raw[cx, cy] for m in (i - 1, i, i + 1)]
            n = q["n_bin"][0, a, c]
            # window uses the interpolated template, the correction the idx template only
            win = (la * Tm[0] + lb * Tm[1] + lc * Tm[2])[nt - L:].sum()
            corr = Tm[1][k0 + 1:nt - L].sum() + Tm[1][nt - L] - fr * Tm[1][k0 + 1]
            total += n * (win + corr)
    for j in range(q["pid_nb"].shape[1]):
        # neighbours absent from U land on row 0, which is -1 (present in U): still summed
        total += q["n_e"][0] * stripe_sum(raw[q["cx_nb"][0, j], q["cy_nb"][0, j]])
    assert wfs.sum() == pytest.approx(total, rel=1e-11)


def test_outside_tpc_contributes_nothing(p1, lut, tables):
    tr, f = _tracks([dict(x=80.0, y=0.0, z=12.0, dE=1.0, dEdx=2.0, eventID=0)])
    wfs, U, q = O.simulate_wfs(p1, lut, tr, f, tables=tables)
    assert q["n_e"][0] == 0.0 and not q["mask"][0]
    assert np.abs(wfs).max() == 0.0


# ------------------------------------------------------------------------------------------
# self-trigger (fee_jax.py:168-279), hit_prob / parse_output (sim_jax.py:619-647, 738-769)
# ------------------------------------------------------------------------------------------

def _deltas(p, spec, T=2000):
    """Waveform rows with charge q delivered entirely at tick t, for (t, q) in spec."""
    w = np.zeros((1, T))
    for t, q in spec:
        w[0, t] += q / p.t_sampling
    return w


@pytest.mark.parametrize("spec,ticks,adcs", [
    # one pulse: crossing between t-1 and t -> i* = t-1; charge read 19 ticks after i*
    ([(100, 6000.0)], [99], [6000.0]),
    # two separated pulses -> two hits
    ([(100, 6000.0), (300, 8000.0)], [99, 299], [6000.0, 8000.0]),
    # second pulse inside the integration window: one hit with both charges
    ([(100, 6000.0), (105, 2000.0)], [99], [8000.0]),
    # second pulse exactly at the reset index i*+20: read outside the window, but the reset
    # subtracts it too, so it never triggers
    ([(100, 6000.0), (119, 9000.0)], [99], [6000.0]),
    # ... one tick later it survives the reset and triggers on its own
    ([(100, 6000.0), (120, 9000.0)], [99, 119], [6000.0, 9000.0]),
    # below threshold: nothing
    ([(100, 4000.0)], [], []),
    # above threshold from tick 0: no up-crossing is ever seen
    ([(0, 9000.0)], [], []),
    # late pulse: read index clamped to the last tick
    ([(1990, 6000.0)], [1989], [6000.0]),
    # eleven pulses: the first ten are kept
    ([(50 + 40 * k, 6000.0 + k) for k in range(11)], [49 + 40 * k for k in range(10)],
     [6000.0 + k for k in range(10)]),
])
def test_self_trigger_hand_built(p1, spec, ticks, adcs):
    w = _deltas(p1, spec)
    a, t, _ = O.get_adc_values(p1, w)
    H = p1.MAX_ADC_VALUES
    exp_t = ticks + [1998] * (H - len(ticks))
    exp_a = adcs + [0.0] * (H - len(adcs))
    assert t[0].tolist() == exp_t
    np.testing.assert_allclose(a[0], exp_a, rtol=1e-12)


def test_self_trigger_ramp(p1):
    """A linear ramp of Q (constant current c) crosses the threshold at the first i with
    Q[i+1] >= thr, i.e. i* = ceil(thr / (c dt)) - 2."""
    c = 523.0
    w = np.full((1, 2000), c)
    a, t, _ = O.get_adc_values(p1, w)
    i_star = math.ceil(p1.DISCRIMINATION_THRESHOLD / (c * p1.t_sampling)) - 2
    assert t[0, 0] == i_star
    assert a[0, 0] == pytest.approx(c * p1.t_sampling * (i_star + 20), rel=1e-12)
    # after the reset the ramp restarts from zero at g = i*+20
    assert t[0, 1] == i_star + 20 + i_star + 1


def test_hit_prob_and_compaction_brute_force(p1):
    rng = np.random.default_rng(2)
    P, H = 7, p1.MAX_ADC_VALUES
    ticks = rng.choice([5.0, 1996.0, 1997.0, 1998.0], size=(P, H))
    hit_prob = np.where(ticks < 2000 - 3, 1.0, 0.0)
    adcs = rng.normal(size=(P, H))
    upix = np.array([-1, 3, 10, 11, 12 + 140 * 280 * 2, 50, 60])
    _, _, plane, ev = O.id2pixel(p1, upix)
    ev = np.where(upix < 0, -1, ev)
    X = rng.normal(size=P)
    Y = rng.normal(size=P)
    Z = rng.normal(size=(P, H))
    out = O.parse_output(p1, adcs, X, Y, Z, ticks, hit_prob, ev, upix)
    want = [(r, k) for r in range(P) for k in range(H) if ticks[r, k] < 1997 and ev[r] >= 0 and upix[r] >= 0]
    assert len(out["adcs"]) == len(want)
    for i, (r, k) in enumerate(want):
        assert out["adcs"][i] == adcs[r, k] and out["hit_pixels"][i] == upix[r]
        assert out["pixel_x"][i] == X[r] and out["pixel_z"][i] == Z[r, k] and out["event"][i] == ev[r]


def test_mse_loss_closed_form():
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    b = np.array([[1.0, 1.0], [5.0, 5.0]])
    # pixel 7 in both, pixel 8 only in a, pixel 9 only in b
    v = O.mse_loss(a, np.array([7, 8]), b, np.array([7, 9]))
    assert v == (0 + 1) + (9 + 16) + (25 + 25)


def test_main_pixel_ids_match_full_quantities(p1):
    """The light main-pixel function (for multi-million-segment pixel sets) equals the centre bin
    of segment_quantities on an event with both TPCs, borders, off-plane and padding rows, also
    with shifted positions and in small chunks."""
    import sys

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from helpers import mixed_event

    tr, f = mixed_event(p1, n_c1=600)
    for p in (p1, p1.replace(shift_x=0.1234, shift_y=-0.0777, shift_z=0.3125)):
        want = O.segment_quantities(p, tr, f)["main_pid"]
        assert (want < -1).any() and (want == -1).any() and (want >= 0).any()
        assert np.array_equal(O.main_pixel_ids(p, tr, f), want)
        assert np.array_equal(O.main_pixel_ids(p, tr, f, chunk=97), want)
