import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "larnd-sim-jax_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built liblsim.so")


@pytest.fixture(scope="session")
def params4():
    from larndsim_amd import consts

    return consts.default_params(number_pix_neighbors=4, RESET_NOISE_CHARGE=0.0, UNCORRELATED_NOISE_CHARGE=0.0)


@pytest.fixture(scope="session")
def lut():
    from larndsim_amd import synthetic

    return synthetic.make_lut(0)


@pytest.fixture(scope="session")
def c1():
    from larndsim_amd import synthetic

    return synthetic.workload("c1")


@pytest.fixture(scope="session")
def tables(params4, lut):
    from oracle import oracle as O

    return O.build_templates(params4, lut)
