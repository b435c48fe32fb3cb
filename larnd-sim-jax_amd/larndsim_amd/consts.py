"""Simulation parameters, detector geometry and the response-LUT templates.

Mirrors the parameter interface of larnd-sim-jax (``src/larndsim/consts_jax.py``):

* ``Params_template`` field names, units and defaults follow ``consts_jax.py:25-160`` and the
  defaults dict of ``load_detector_properties`` (``consts_jax.py:244-298``);
* ``build_params_class`` (``consts_jax.py:162-190``) returns a dataclass type whose instances
  support ``.replace(**kw)`` like a flax struct dataclass; the set of differentiable
  fields is recorded on the class as ``_grad_fields``;
* ``get_vdrift`` follows ``consts_jax.py:192-216``;
* ``load_detector_properties`` follows ``consts_jax.py:220-385`` (YAML detector properties +
  pixel layout -> TPC borders, pixel counts);
* ``load_lut`` follows ``consts_jax.py:387-449``; here it returns the raw (Nx, Ny, Nt) response
  and the derived template tables are built on the GPU by the library (``lsim_lut_set``).

``tpc_borders`` is kept as float32 values (the reference stores it with ``jnp.asarray`` under
32-bit JAX); every other scalar is a Python float (float64).
"""
from __future__ import annotations

import dataclasses
import json
import os
from collections import defaultdict
from enum import Enum
from typing import Iterable

import numpy as np

_DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class RecombinationMode(Enum):
    """Recombination models (``consts_jax.py:20-23``)."""

    BOX = 1
    BIRKS = 2
    ELLIPSOID = 3


# The seven physics parameters the adjoint differentiates (order is the gradient layout).
GRAD_PARAMS = ("Ab", "kb", "eField", "lifetime", "vdrift", "long_diff", "tran_diff")


@dataclasses.dataclass(frozen=True)
class Params_template:
    """Parameter set of the LUT simulation (field names of ``consts_jax.py:80-160``)."""

    eField: float = 0.50
    Ab: float = 0.8
    kb: float = 0.0486
    lifetime: float = 2.2e3
    vdrift: float = 0.1648
    vdrift_static: float = 0.159645
    long_diff: float = 4.0e-6
    tran_diff: float = 8.8e-6
    tpc_borders: np.ndarray = None
    recombination_mode: RecombinationMode = RecombinationMode.BIRKS
    lArDensity: float = 1.38
    alpha: float = 0.93
    beta: float = 0.212
    R_param: float = 1.25
    MeVToElectrons: float = 4.237e04
    pixel_pitch: float = 0.4434
    n_pixels_x: int = 140
    n_pixels_y: int = 280
    max_radius: int = 0
    max_active_pixels: int = 0
    drift_length: float = 0.0
    t_sampling: float = 0.1
    time_interval: tuple = (0.0, 200.0)
    time_padding: float = 190
    min_step_size: float = 0.001
    time_max: float = 0
    time_window: float = 189.1
    e_charge: float = 1.602e-19
    temperature: float = 87.17
    response_bin_size: float = 0.04434
    number_pix_neighbors: int = 1
    electron_sampling_resolution: float = 0.001
    signal_length: int = 150
    response_full_drift_t: float = 190.61638
    MAX_ADC_VALUES: int = 10
    DISCRIMINATION_THRESHOLD: float = 7e3
    ADC_HOLD_DELAY: int = 15
    CLOCK_CYCLE: float = 0.1
    GAIN: float = 4e-3
    V_CM: float = 288.0
    V_REF: float = 1300.0
    V_PEDESTAL: float = 580.0
    ADC_COUNTS: int = 2 ** 8
    RESET_NOISE_CHARGE: float = 900.0
    UNCORRELATED_NOISE_CHARGE: float = 500.0
    ELECTRON_MOBILITY_PARAMS: tuple = (551.6, 7158.3, 4440.43, 4.29, 43.63, 0.2053)
    shift_x: float = 0.0
    shift_y: float = 0.0
    shift_z: float = 0.0
    size_margin: float = 2e-2
    diffusion_in_current_sim: bool = True
    mc_diff: bool = False
    nb_sampling_bins_per_pixel: int = 10
    long_diff_template: np.ndarray = None
    long_diff_extent: int = 20
    roi_threshold: float = 0.01
    roi_split_length: int = 400
    fee_paths_scaling: int = 20
    nb_tran_diff_bins: int = 5
    hit_prob_threshold: float = 1e-5
    tran_diff_bin_edges: np.ndarray = None

    # flax.struct API used throughout the reference (params.replace(...)).
    def replace(self, **kw):
        return dataclasses.replace(self, **kw)

    @property
    def n_ticks(self) -> int:
        """Array ticks including the garbage column 0 (``sim_jax.py:147``)."""
        return int(self.time_interval[1] / self.t_sampling) + 1

    @property
    def adc_interval(self) -> int:
        """Integration window of the self-trigger in ticks (``fee_jax.py:216``)."""
        return round((3 * self.CLOCK_CYCLE + self.ADC_HOLD_DELAY * self.CLOCK_CYCLE) / self.t_sampling)


def build_params_class(params_with_grad: Iterable[str]):
    """Return the parameter dataclass; ``params_with_grad`` names the differentiable fields.

    Reference: ``consts_jax.py:162-190``. Here every field is a plain attribute; the names of the
    differentiable ones are kept on the class so the autograd layer knows which to expose.
    """
    names = tuple(params_with_grad)
    fields = {f.name for f in dataclasses.fields(Params_template)}
    for n in names:
        if n not in fields:
            raise ValueError(f"Unknown parameter '{n}'")
    cls = type("Params", (Params_template,), {"_grad_fields": names})
    return dataclasses.dataclass(frozen=True)(cls)


def get_vdrift(params) -> float:
    """Electron drift velocity in cm/us from eField (kV/cm) and temperature (K).

    Reference: ``consts_jax.py:192-216`` (mobility parameterisation of
    https://doi.org/10.1016/j.nima.2016.01.073).
    """
    return vdrift_of(params.eField, params.temperature, params.ELECTRON_MOBILITY_PARAMS)


def vdrift_of(eField, temperature, mob):
    a0, a1, a2, a3, a4, a5 = mob
    num = a0 + a1 * eField + a2 * pow(eField, 1.5) + a3 * pow(eField, 2.5)
    denom = 1 + (a1 / a0) * eField + a4 * pow(eField, 2) + a5 * pow(eField, 3)
    temp_corr = pow(temperature / 89, -1.5)
    mu = num / denom * temp_corr / 1000
    return mu * eField


def _yaml_load(path):
    import yaml

    loader = getattr(yaml, "CSafeLoader", yaml.SafeLoader)
    with open(path) as f:
        return yaml.load(f, Loader=loader)


_DETPROP_KEYS = {f.name for f in dataclasses.fields(Params_template)} | {"tpc_centers"}


def load_detector_properties(params_cls, detprop_file, pixel_file):
    """Build a parameter instance from a detector-properties YAML and a pixel-layout YAML.

    Reference: ``consts_jax.py:220-385``. Unknown keys in the detector file raise ``ValueError``
    (``consts_jax.py:306-313``); the TPC borders are derived from the tile positions,
    orientations and drift length exactly as the reference does (``consts_jax.py:317-383``).
    """
    mm2cm = 0.1
    d = {}
    tpc_centers = np.zeros((2, 3))
    detprop = _yaml_load(detprop_file)
    for key, value in detprop.items():
        if key not in _DETPROP_KEYS:
            raise ValueError(f"Key '{key}' in detector properties file is not recognized.")
        if key == "tpc_centers":
            tpc_centers = np.array(value, dtype=np.float64)
        elif isinstance(value, list):
            d[key] = tuple(value)
        else:
            d[key] = value
    tpc_centers[:, [2, 0]] = tpc_centers[:, [0, 2]]

    layout = _yaml_load(pixel_file)
    pitch = layout["pixel_pitch"] * mm2cm
    d["pixel_pitch"] = pitch
    pos = np.array(list(layout["chip_channel_to_position"].values()))
    xs = pos[:, 0] * pitch
    ys = pos[:, 1] * pitch
    tile_borders = np.zeros((2, 2))
    tile_borders[0] = [-(max(xs) + pitch) / 2, (max(xs) + pitch) / 2]
    tile_borders[1] = [-(max(ys) + pitch) / 2, (max(ys) + pitch) / 2]

    tile_indeces = layout["tile_indeces"]
    tpc_ids = np.unique(np.array(list(tile_indeces.values()))[:, 0], axis=0)
    anodes = defaultdict(list)
    cathode_dir = {}
    for tpc_id in tpc_ids:
        dirs = []
        for tile in tile_indeces:
            if tile_indeces[tile][0] == tpc_id:
                anodes[tpc_id].append(layout["tile_positions"][tile])
                dirs.append(layout["tile_orientations"][tile][0])
        if len(set(dirs)) != 1:
            raise ValueError("Tiles in same anode plane have different drift directions.")
        if dirs[0] not in (1, -1):
            raise ValueError("Cathode direction should be either 1 or -1.")
        cathode_dir[tpc_id] = dirs[0]

    if "drift_length" in layout:
        d["drift_length"] = layout["drift_length"] * mm2cm
    else:
        mod_anodes = np.array(list(layout["tile_positions"].values()))[:, 0] * mm2cm
        d["drift_length"] = 0.5 * (max(mod_anodes) - min(mod_anodes)) * mm2cm

    borders = np.zeros((len(anodes), 3, 2))
    for itpc, tpc_id in enumerate(anodes):
        tiles = np.vstack(anodes[tpc_id]) * mm2cm
        c = tpc_centers[itpc] if itpc < len(tpc_centers) else np.zeros(3)
        borders[itpc, 0] = (min(tiles[:, 2]) + tile_borders[0][0] + c[0], max(tiles[:, 2]) + tile_borders[0][1] + c[0])
        borders[itpc, 1] = (min(tiles[:, 1]) + tile_borders[1][0] + c[1], max(tiles[:, 1]) + tile_borders[1][1] + c[1])
        borders[itpc, 2] = (min(tiles[:, 0]) + c[2], max(tiles[:, 0]) + d["drift_length"] * cathode_dir[tpc_id] + c[2])
    d["tpc_borders"] = borders.astype(np.float32)
    d["n_pixels_x"] = len(np.unique(xs)) * 2
    d["n_pixels_y"] = len(np.unique(ys)) * 4
    d["long_diff_template"] = np.linspace(0.001, 10, 100).astype(np.float32)
    return params_cls(**{k: v for k, v in d.items() if k in {f.name for f in dataclasses.fields(params_cls)}})


def geometry_to_json(params) -> dict:
    """Serialisable snapshot of what ``load_detector_properties`` derived (for machines
    without the YAML files)."""
    out = {}
    for f in dataclasses.fields(params):
        v = getattr(params, f.name)
        if isinstance(v, np.ndarray):
            out[f.name] = {"array": v.tolist(), "dtype": str(v.dtype)}
        elif isinstance(v, Enum):
            out[f.name] = {"enum": v.name}
        else:
            out[f.name] = v
    return out


def params_from_json(params_cls, blob: dict):
    kw = {}
    for k, v in blob.items():
        if isinstance(v, dict) and "array" in v:
            kw[k] = np.array(v["array"], dtype=v["dtype"])
        elif isinstance(v, dict) and "enum" in v:
            kw[k] = RecombinationMode[v["enum"]]
        elif isinstance(v, list):
            kw[k] = tuple(v)
        else:
            kw[k] = v
    return params_cls(**kw)


def default_params(params_with_grad: Iterable[str] = GRAD_PARAMS, **overrides):
    """module0 detector properties + the multi_tile_layout-2.4.16_v4 pixel layout.

    Loaded from the committed snapshot ``data/module0_layout-2.4.16_v4.json`` (written by
    ``data/make_geometry.py`` with ``load_detector_properties`` from the two YAML files).
    """
    cls = build_params_class(params_with_grad)
    with open(os.path.join(_DATA_DIR, "module0_layout-2.4.16_v4.json")) as f:
        p = params_from_json(cls, json.load(f))
    return p.replace(**overrides) if overrides else p


def template_sigmas(params) -> np.ndarray:
    """Longitudinal-diffusion template grid in ticks, float32 like ``jnp.linspace`` under
    32-bit JAX (``consts_jax.py:296``)."""
    if params.long_diff_template is not None:
        return np.asarray(params.long_diff_template, dtype=np.float32)
    return np.linspace(0.001, 10, 100).astype(np.float32)


def load_lut(lut_file, params):
    """Load the raw response LUT (``consts_jax.py:387-449``).

    Returns ``(response, params)`` where ``response`` is the raw (Nx, Ny, Nt) float32 array.
    The reference returns the 100 Gaussian-smeared templates here; this implementation
    builds the smeared window and cumulative tables on the device from the raw response
    (same mathematics: template m is the raw response convolved along time with a normalised
    Gaussian of sigma ``long_diff_template[m]`` over +-``long_diff_extent`` ticks, template 0
    is the raw response).  Only ``.npy`` / ``.npz`` files are read, with ``allow_pickle=False``.
    """
    updated = params
    resp = np.load(lut_file, allow_pickle=False)
    if isinstance(resp, np.lib.npyio.NpzFile):
        if "response" not in resp:
            raise ValueError("No 'response' key found in the npz file.")
        if "drift_length" in resp:
            updated = params.replace(response_full_drift_t=float(resp["drift_length"]) / params.vdrift_static)
        resp = resp["response"]
    elif not isinstance(resp, np.ndarray):
        raise ValueError("Unsupported response format. Expected npz or numpy array.")
    resp = np.ascontiguousarray(resp, dtype=np.float32)
    if resp.ndim != 3:
        raise ValueError(f"response LUT must be 3-D (Nx, Ny, Nt), got shape {resp.shape}")
    return resp, updated
