"""Regenerate ``module0_layout-2.4.16_v4.json`` from the detector YAML files.

Usage: python make_geometry.py <module0.yaml> <multi_tile_layout-2.4.16_v4.yaml>
(the two files live in larnd-sim-jax under src/larndsim/detector_properties/ and
src/larndsim/pixel_layouts/).  The JSON holds what ``load_detector_properties`` derives, so
that machines without the YAML files get identical parameters.
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from larndsim_amd.consts import GRAD_PARAMS, build_params_class, geometry_to_json, load_detector_properties  # noqa: E402

if __name__ == "__main__":
    detprop, layout = sys.argv[1], sys.argv[2]
    p = load_detector_properties(build_params_class(GRAD_PARAMS), detprop, layout)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "module0_layout-2.4.16_v4.json")
    with open(out, "w") as f:
        json.dump(geometry_to_json(p), f, indent=1)
    print("wrote", out)
