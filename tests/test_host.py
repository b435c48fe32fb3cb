"""Host-side checks that need no GPU: the C ABI library loads and exports every entry point
of include/lsim.h, the ctypes struct layouts equal the C ones, the product path refuses to run
without a GPU (no CPU fallback), and the host logic mirroring the reference (pad_size,
parameters, geometry, synthetic workloads, torch helpers) behaves as the reference does."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lsim.h")


def _header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lsim_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from larndsim_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        import sys

        sys.path.insert(0, os.path.join(ROOT, "larnd-sim-jax_amd"))
        import build

        build.build(verbose=False)
    return _lib.load()


def test_library_exports_every_header_entry_point(lib):
    from larndsim_amd import _lib

    names = _header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTS)


def test_library_calls_without_gpu_work(lib):
    """Entry points that touch no device memory answer on a host without a GPU."""
    import ctypes as C

    assert lib.lsim_version() > 0
    assert lib.lsim_tape_bytes() > 0
    assert lib.lsim_set_row_waves(3) < 0  # only 1, 2, 4, 8
    assert b"waves" in lib.lsim_last_error()
    assert lib.lsim_set_row_waves(4) == 0
    # a null context is an argument error, not a crash
    sz = C.c_size_t()
    bound = C.c_int64()
    assert lib.lsim_forward_query(None, 10, 1, C.byref(sz), C.byref(bound)) < 0


def test_struct_layouts_match_header(tmp_path):
    from larndsim_amd import _lib

    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "abi_layout.c")], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c = dict(ln.split() for ln in lines if ln.strip())
    types = {"lsim_geometry": _lib.Geometry, "lsim_config": _lib.Config, "lsim_params": _lib.ParamsC,
             "lsim_segments": _lib.Segments, "lsim_fwd_out": _lib.FwdOut, "lsim_chop_layout": _lib.ChopLayout}
    n = 0
    for key, val in c.items():
        parts = key.split(".")
        T = types[parts[1]]
        if parts[0] == "sizeof":
            assert C_sizeof(T) == int(val), key
        else:
            assert getattr(T, parts[2]).offset == int(val), key
        n += 1
    assert n == 87


def C_sizeof(T):
    import ctypes

    return ctypes.sizeof(T)


def test_no_cpu_fallback(params4, lut):
    """The product path raises without a GPU instead of computing on the host."""
    from larndsim_amd import _lib
    from larndsim_amd.sim import Simulator

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.LsimError):
        Simulator(params4, lut)


def test_noise_is_refused(params4):
    from larndsim_amd.sim import _check_noise

    with pytest.raises(NotImplementedError):
        _check_noise(params4.replace(RESET_NOISE_CHARGE=900.0))


def test_pad_size_history_rule():
    from larndsim_amd import sim

    sim.size_history_dict.clear()
    assert sim.pad_size(100, "t") == 102  # int(100 * 1.025 + 0.5), 100 * 1.025 rounds below 102.5
    assert sim.pad_size(101, "t") == 102  # reuses a size within +5 %
    assert sim.pad_size(102, "t") == 102
    assert sim.pad_size(90, "t") == 92  # 102 > 90 * 1.05: new size
    assert sim.pad_size((10, 20), "u") == (10, 21)
    assert sim.get_size_history()["t"] == [(92,), (102,)]
    sim.size_history_dict.clear()


def test_params_struct_drift_velocity(params4):
    from larndsim_amd import consts
    from larndsim_amd.sim import params_struct
    from oracle import oracle as O

    s = params_struct(params4)
    assert s.vdrift_eff == pytest.approx(O.vdrift(params4), rel=1e-15)
    h = 1e-6
    v = lambda E: consts.vdrift_of(E, params4.temperature, params4.ELECTRON_MOBILITY_PARAMS)  # noqa: E731
    fd = (v(params4.eField + h) - v(params4.eField - h)) / (2 * h)
    assert s.dvdrift_deField == pytest.approx(fd, rel=1e-8)
    assert s.threshold == params4.DISCRIMINATION_THRESHOLD and s.recombination_mode == 2
    assert s.vdrift == params4.vdrift  # carried, never read


def test_module0_geometry_values(params4):
    """Derived module0 + multi_tile_layout-2.4.16_v4 values (consts_jax.py:220-385)."""
    p = params4
    assert p.pixel_pitch == pytest.approx(0.4434)
    assert (p.n_pixels_x, p.n_pixels_y) == (140, 280)
    b = np.asarray(p.tpc_borders, np.float64)
    np.testing.assert_allclose(b[0], [[-31.038, 31.038], [-62.076, 62.076], [30.58975, 0.15875]], atol=1e-5)
    np.testing.assert_allclose(b[1], [[-31.038, 31.038], [-62.076, 62.076], [-30.58975, -0.15875]], atol=1e-5)
    assert p.drift_length == pytest.approx(30.431, abs=1e-6)
    assert p.n_ticks == 2001 and p.adc_interval == 18 and p.MAX_ADC_VALUES == 10
    assert (p.GAIN, p.V_PEDESTAL, p.V_CM, p.V_REF, p.DISCRIMINATION_THRESHOLD) == (
        4e-3, 580.0, 284.27734375, 1282.71484375, 5000)
    assert (p.eField, p.Ab, p.kb, p.lifetime, p.long_diff, p.tran_diff) == (0.5, 0.8, 0.0486, 2200.0, 4e-6, 8.8e-6)
    from larndsim_amd.consts import template_sigmas

    s = template_sigmas(p)
    assert s.dtype == np.float32 and len(s) == 100 and s[0] == np.float32(0.001) and s[-1] == np.float32(10.0)


def test_build_params_class_rejects_unknown():
    from larndsim_amd import consts

    with pytest.raises(ValueError):
        consts.build_params_class(["Ab", "not_a_param"])
    cls = consts.build_params_class(["Ab"])
    assert cls._grad_fields == ("Ab",)


def test_load_lut_formats(tmp_path, params4):
    from larndsim_amd import consts

    a = np.random.default_rng(0).normal(size=(3, 3, 7)).astype(np.float32)
    np.save(tmp_path / "r.npy", a)
    r, _ = consts.load_lut(str(tmp_path / "r.npy"), params4)
    assert np.array_equal(r, a)
    np.savez(tmp_path / "r.npz", response=a, drift_length=np.float64(30.0))
    r, p2 = consts.load_lut(str(tmp_path / "r.npz"), params4)
    assert np.array_equal(r, a) and p2.response_full_drift_t == pytest.approx(30.0 / params4.vdrift_static)
    np.savez(tmp_path / "bad.npz", other=a)
    with pytest.raises(ValueError):
        consts.load_lut(str(tmp_path / "bad.npz"), params4)


def test_synthetic_workloads_are_seeded_and_inside_the_cube():
    from larndsim_amd import synthetic

    a, f = synthetic.workload("c1")
    b, _ = synthetic.workload("c1")
    assert np.array_equal(a, b) and a.shape == (1000, len(f)) and a.dtype == np.float32
    for k, ax in (("x", 0), ("y", 1), ("z", 2)):
        v = a[:, f.index(k)]
        assert (v >= synthetic.CUBE_LO[ax] - 1e-4).all() and (v <= synthetic.CUBE_HI[ax] + 1e-4).all()
    np.testing.assert_allclose(a[:, f.index("dE")], a[:, f.index("dEdx")] * a[:, f.index("dx")], rtol=1e-5)
    np.testing.assert_allclose(a[:, f.index("dx")], 0.4, rtol=1e-5)
    assert (a[:, f.index("eventID")] == 0).all()
    sizes = synthetic.event_sizes("c3")
    assert len(sizes) == 128 and sizes.min() >= 2000 and sizes.max() <= 400000
    lut = synthetic.make_lut(0)
    assert lut.shape == (45, 45, 1950) and lut.dtype == np.float32
    # the collecting pixel's centre response integrates to one charge unit
    assert lut[0, 0].sum() * 0.1 == pytest.approx(1.0, rel=1e-4)


def test_torch_helpers_match_oracle(params4):
    from larndsim_amd import detsim, fee, losses
    from oracle import oracle as O

    q = torch.linspace(-1e4, 9e4, 17, dtype=torch.float64)
    np.testing.assert_allclose(fee.digitize(params4, q).numpy(), O.digitize(params4, q.numpy()), rtol=1e-15)
    d = torch.linspace(0, 256, 9, dtype=torch.float64)
    np.testing.assert_allclose(fee.adc2charge(d, params4).numpy(), O.adc2charge(d.numpy(), params4), rtol=1e-15)
    rng = np.random.default_rng(1)
    px, py = rng.integers(-2, 142, 50), rng.integers(-2, 282, 50)
    pl, ev = rng.integers(0, 2, 50), rng.integers(0, 9, 50)
    pid_t = detsim.pixel2id(params4, px, py, pl, ev).numpy()
    assert np.array_equal(pid_t, O.pixel2id(params4, px, py, pl, ev))
    ok = pid_t >= 0
    for a, bb in zip(detsim.id2pixel(params4, pid_t[ok]), O.id2pixel(params4, pid_t[ok])):
        assert np.array_equal(a.numpy(), bb)
    t = torch.tensor([0.0, 17.0, 1998.0], dtype=torch.float64)
    np.testing.assert_allclose(detsim.get_hit_z(params4, t, torch.tensor([0, 1, 0])).numpy(),
                               O.get_hit_z(params4, t.numpy(), np.array([0, 1, 0])), rtol=1e-14)
    xy = detsim.get_pixel_coordinates(params4, torch.tensor([3.0]), torch.tensor([5.0]), torch.tensor([1])).numpy()
    X, Y = O.get_pixel_coordinates(params4, np.array([3.0]), np.array([5.0]), np.array([1]))
    np.testing.assert_allclose(xy[0], [X[0], Y[0]], rtol=1e-6)
    a = torch.tensor([[1.0, 2.0], [3.0, 4.0]], dtype=torch.float64)
    b = torch.tensor([[1.0, 1.0], [5.0, 5.0]], dtype=torch.float64)
    got, _ = losses.mse_loss(a, torch.tensor([7, 8]), b, torch.tensor([7, 9]))
    assert float(got) == O.mse_loss(a.numpy(), np.array([7, 8]), b.numpy(), np.array([7, 9]))


def test_mmd_closed_form():
    """Identical weighted clouds have zero MMD; one point each has 2 - 2 exp(-d^2 / 2 s^2)."""
    from larndsim_amd import losses

    x = torch.randn(5, 3, dtype=torch.float64)
    w = torch.rand(5, dtype=torch.float64)
    assert float(losses.mmd(x, x, w, w, 1.0)) == pytest.approx(0.0, abs=1e-12)
    a = torch.zeros(1, 3, dtype=torch.float64)
    b = torch.tensor([[1.0, 2.0, 2.0]], dtype=torch.float64)
    one = torch.ones(1, dtype=torch.float64)
    assert float(losses.mmd(a, b, one, one, 1.5)) == pytest.approx(2 - 2 * np.exp(-9 / (2 * 1.5 ** 2)), rel=1e-12)


def test_new_entry_points_check_arguments(lib):
    """The readout, soft-DTW, nearest-neighbour and chop entry points reject bad arguments on the
    host (no device access): sizes, limits and NULL pointers."""
    import ctypes as C

    sz = C.c_size_t()
    assert lib.lsim_prob_fee_query(10, 2000, 20, 10, C.byref(sz)) == 0 and sz.value > 10 * 4 * 2000 * 4
    assert lib.lsim_prob_fee_query(10, 1, 20, 10, C.byref(sz)) < 0
    # n_ticks beyond 2048, more paths than 32, noise 0
    assert lib.lsim_prob_fee(None, 1, 4096, 20, 10, 18, C.c_float(7e3), C.c_float(900.0), C.c_float(0.1),
                             C.c_float(1e-9), None, 0, None, None, None) < 0
    assert lib.lsim_prob_fee(None, 1, 2000, 33, 10, 18, C.c_float(7e3), C.c_float(900.0), C.c_float(0.1),
                             C.c_float(1e-9), None, 0, None, None, None) < 0
    assert lib.lsim_prob_fee(None, 1, 2000, 20, 10, 18, C.c_float(7e3), C.c_float(0.0), C.c_float(0.1),
                             C.c_float(1e-9), None, 1 << 30, None, None, None) < 0
    assert b"noise_sigma" in lib.lsim_last_error()
    # the parametrized current model: modes, bin edges, sample count, NULLs (no device access)
    pp = C.c_void_p(8)
    assert lib.lsim_param_electrons(None, None, None, 0, 1, None, None, 6, None, None, None, None, None, None, None,
                                    None) < 0
    assert lib.lsim_param_rows(None, None, None, None, None, None, None, None, 0, None, None, 1, None, None) < 0
    assert lib.lsim_param_rows(pp, None, None, None, None, None, None, None, 0, None, None, 1, None, None) < 0
    # its adjoint: sizes and NULLs
    assert lib.lsim_param_backward_query(None, 1, 1, C.byref(sz)) < 0
    assert lib.lsim_param_backward_query(pp, -1, 1, C.byref(sz)) < 0
    assert lib.lsim_param_backward_query(pp, 10, 4, C.byref(sz)) == 0 and sz.value >= 10 * 128 + 4 * 7 * 8
    assert lib.lsim_param_backward(pp, None, None, 1, 1, None, None, 0, None, None, None, None, None, None, 0, None,
                                   None, None, None, None, None, None, 0, None, None) < 0
    assert b"NULL" in lib.lsim_last_error()
    # scheduling knobs: out-of-range values are refused on the host
    assert lib.lsim_set_row_tile(0) < 0 and lib.lsim_set_row_tile(4097) < 0 and b"tile" in lib.lsim_last_error()
    assert lib.lsim_set_row_tile(16) == 0
    assert lib.lsim_set_row_cost(0) < 0 and lib.lsim_set_row_cost(256) == 0
    # the readout's adjoint: same limits, a workspace of the forward's size, no rows is a no-op
    bwd = lambda P, T, V, ws_bytes: lib.lsim_prob_fee_backward(  # noqa: E731
        P, T, V, 10, 18, C.c_float(7e3), C.c_float(900.0), C.c_float(0.1), None, ws_bytes, None, None, None, None, None)
    assert bwd(1, 4096, 20, 1 << 30) < 0 and bwd(1, 2000, 33, 1 << 30) < 0
    assert lib.lsim_prob_fee_query(10, 2000, 20, 10, C.byref(sz)) == 0
    assert bwd(10, 2000, 20, sz.value - 1) == -4 and b"workspace" in lib.lsim_last_error()
    assert bwd(10, 2000, 20, sz.value) < 0 and b"NULL" in lib.lsim_last_error()
    assert bwd(0, 2000, 20, 1 << 20) == 0
    # grid kernel sums: the cull radius is where exp(-d^2 / 2 sigma^2) is 0 in float32 (e^-106), +1 %
    assert lib.lsim_rbf_cull_radius(C.c_float(2.0)) == pytest.approx(1.01 * np.sqrt(212.0) * 2.0, rel=1e-6)
    assert float(np.float32(np.exp(-106.0))) == 0.0
    grid = lambda nx, ny, dims, sigma: lib.lsim_rbf_sums_grid(  # noqa: E731
        C.c_void_p(8), C.c_void_p(8), nx, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), ny, C.c_void_p(8), dims,
        C.c_float(sigma), C.c_void_p(8), None)
    assert grid(10, 10, 5, 1.0) < 0 and b"dims" in lib.lsim_last_error()
    assert grid(10, 10, 3, 0.0) < 0 and grid(10, 10, 3, float("inf")) < 0
    assert grid(10, 1 << 31, 3, 1.0) < 0 and grid(-1, 10, 3, 1.0) < 0
    assert grid(0, 10, 3, 1.0) == 0
    assert lib.lsim_sdtw_query(100, 200, C.byref(sz)) == 0 and sz.value >= 2 * 100 * 200 * 4
    assert lib.lsim_sdtw_query(0, 5, C.byref(sz)) < 0
    assert lib.lsim_sdtw_query(1 << 20, 1 << 15, C.byref(sz)) < 0  # n * m beyond 2^34
    assert lib.lsim_sdtw(None, 10, None, 10, C.c_float(1.0), None, 0, None, None, None) < 0
    assert lib.lsim_sdtw(C.c_void_p(8), 10, C.c_void_p(8), 10, C.c_float(0.0), C.c_void_p(8), 1 << 20,
                         C.c_void_p(8), None, None) < 0
    assert b"gamma" in lib.lsim_last_error()
    assert lib.lsim_nearest(None, 5, None, None, None, 5, 6, None, None, None) < 0  # n_origin > nx
    assert lib.lsim_nearest(None, 0, None, None, None, 0, 0, None, None, None) == 0  # nothing to do
    assert lib.lsim_chop_query(0, C.byref(sz)) == 0


def test_needs_grad_follows_grad_mode():
    """Parameters that require gradients are recorded only in grad mode (under torch.no_grad()
    predict takes the plain path, as the reference's predict runs in both modes)."""
    import torch

    from larndsim_amd import consts
    from larndsim_amd.autograd import needs_grad

    p = consts.default_params(RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)
    assert not needs_grad(p)
    pg = p.replace(Ab=torch.tensor(float(p.Ab), dtype=torch.float64, requires_grad=True))
    assert needs_grad(pg)
    with torch.no_grad():
        assert not needs_grad(pg)
