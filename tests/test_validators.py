"""Event-id validators of the batching step (detsim_jax.py:24-152): the int64 packing limits
follow from the packing formula, and the checks accept and reject exactly at them."""
import numpy as np
import pytest

from larndsim_amd import consts, detsim


@pytest.fixture(scope="module")
def params():
    return consts.default_params()


@pytest.mark.parametrize("kind", ["pixel", "bin"])
def test_packing_limit_is_tight(params, kind):
    top = detsim.max_safe_event_id(params, kind)
    assert detsim.max_possible_id(params, top, kind) <= detsim.INT64_MAX < detsim.max_possible_id(params, top + 1, kind)
    # the largest packed pixel id of an event equals pixel2id of its last pixel on its last plane
    if kind == "pixel":
        nx, ny, ntpc = int(params.n_pixels_x), int(params.n_pixels_y), detsim._ntpc(params)
        e = 12345
        last = int(detsim.pixel2id(params, nx - 1, ny - 1, ntpc - 1, e))
        assert last == detsim.max_possible_id(params, e, kind)
    detsim.validate_event_ids_for_packing(params, [-1, 0, top], kind)
    with pytest.raises(OverflowError):
        detsim.validate_event_ids_for_packing(params, [0, top + 1], kind)
    with pytest.raises(ValueError):
        detsim.validate_event_ids_for_packing(params, [0, -2], kind)
    detsim.validate_event_ids_for_packing(params, [], kind)
    detsim.validate_event_ids_for_packing(params, [-1, -1], kind)


@pytest.mark.parametrize("kind", ["pixel", "bin"])
def test_packed_id_decoding(params, kind):
    s = detsim._stride(params, kind)
    top = detsim.max_safe_event_id(params, kind)
    detsim.validate_packed_ids_for_decoding(params, [-1, 0, top * s + s - 1], kind)
    with pytest.raises(ValueError):
        detsim.validate_packed_ids_for_decoding(params, [5, -7], kind)
    if (top + 1) * s <= detsim.INT64_MAX:
        with pytest.raises(OverflowError):
            detsim.validate_packed_ids_for_decoding(params, [(top + 1) * s], kind)
    with pytest.raises(ValueError):
        detsim.validate_event_ids_for_packing(params, [0], "strip")


def test_local_event_ids():
    detsim.validate_local_event_ids([-1, 0, 2, 1, 1, -1])
    detsim.validate_local_event_ids([])
    detsim.validate_local_event_ids(np.full(4, -1))
    with pytest.raises(ValueError):
        detsim.validate_local_event_ids([0, 2])
    with pytest.raises(ValueError):
        detsim.validate_local_event_ids([1, 2])
    with pytest.raises(ValueError):
        detsim.validate_local_event_ids([0, -3])
