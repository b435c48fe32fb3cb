"""Edge cases of the HIP forward against the FP64 oracle: empty input, segments outside the
TPCs, padding segments (eventID -1, dE 0; dataio.py:340-373), both anode planes, the pixel-plane
border, the BOX and ELLIPSOID recombination models (quenching_jax.py:18-75), and the other
neighbour / diffusion-bin configurations (number_pix_neighbors 2, nb_tran_diff_bins 3 and 1)."""
import numpy as np
import pytest
import torch

from helpers import compare_forward, mixed_event

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(params4, lut):
    from larndsim_amd.sim import Simulator

    return Simulator(params4, lut)


def test_empty(sim, params4):
    from larndsim_amd.synthetic import SEGMENT_FIELDS

    st = sim.forward(params4, np.zeros((0, len(SEGMENT_FIELDS)), np.float32), SEGMENT_FIELDS, n_events=1)
    torch.cuda.synchronize()
    assert st.max_rows == 0
    assert int(st.out["n_hits"].item()) == 0


def test_mixed_event(sim, params4, lut, tables):
    tr, f = mixed_event(params4)
    compare_forward(sim, params4, lut, tr, f, tables=tables, n_events=1, tag="mixed")


@pytest.mark.parametrize("mode", ["BOX", "ELLIPSOID"])
def test_recombination_models(sim, params4, lut, tables, mode):
    from larndsim_amd.consts import RecombinationMode

    p = params4.replace(recombination_mode=RecombinationMode[mode])
    tr, f = mixed_event(params4, n_c1=300)
    compare_forward(sim, p, lut, tr, f, tables=tables, n_events=1, tag=mode)


@pytest.mark.parametrize("nn,nb", [(0, 5), (2, 5), (4, 3), (1, 1)])
def test_other_configurations(params4, lut, nn, nb):
    from larndsim_amd.sim import Simulator
    from oracle import oracle as O

    p = params4.replace(number_pix_neighbors=nn, nb_tran_diff_bins=nb)
    s = Simulator(p, lut)
    tr, f = mixed_event(p, n_c1=300)
    compare_forward(s, p, lut, tr, f, tables=O.build_templates(p, lut), n_events=1, tag=f"n{nn}b{nb}")


@pytest.mark.parametrize("kw", [dict(signal_length=187), dict(signal_length=100), dict(MAX_ADC_VALUES=4)],
                         ids=["L187", "L100", "H4"])
def test_window_length_and_hit_slots(params4, lut, kw):
    """The longest window the kernels take (187 samples: three 63-lane chunks), a shorter one, and
    fewer self-trigger passes than the compiled maximum."""
    from larndsim_amd.sim import Simulator
    from oracle import oracle as O

    p = params4.replace(**kw)
    s = Simulator(p, lut)
    tr, f = mixed_event(p, n_c1=300)
    compare_forward(s, p, lut, tr, f, tables=O.build_templates(p, lut), n_events=1, tag=str(kw))
