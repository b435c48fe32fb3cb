"""ctypes binding of ``liblsim.so`` (C ABI declared in ``include/lsim.h``).

The library is built in-tree by ``build.py`` / ``__graft_entry__.build()``.  There is no CPU
fallback: if the library is missing every simulation entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblsim.so")

MAX_TPC = 8
MAX_TEMPLATES = 128
MAX_ADC = 16

EXPORTS = (
    "lsim_last_error", "lsim_version", "lsim_tape_bytes", "lsim_ctx_create", "lsim_ctx_destroy", "lsim_lut_set",
    "lsim_forward_query", "lsim_forward_prepare", "lsim_forward_rows", "lsim_forward", "lsim_fee_query",
    "lsim_fee_forward", "lsim_fee_forward_noalloc", "lsim_negative_rows_query", "lsim_negative_rows",
    "lsim_backward_query", "lsim_backward", "lsim_backward_wfs", "lsim_set_row_waves", "lsim_set_row_cost",
    "lsim_set_row_tile",
    "lsim_segment_records", "lsim_timing", "lsim_timing_read", "lsim_rbf_sums", "lsim_rbf_sums_grid",
    "lsim_rbf_cull_radius", "lsim_nearest", "lsim_chop_query", "lsim_chop_count", "lsim_chop_fill",
    "lsim_prob_fee_query", "lsim_prob_fee", "lsim_prob_fee_backward", "lsim_sdtw_query", "lsim_sdtw",
    "lsim_param_electrons", "lsim_param_rows", "lsim_param_backward_query", "lsim_param_backward",
)


class Geometry(C.Structure):
    _fields_ = [
        ("n_tpc", C.c_int32), ("n_pixels_x", C.c_int32), ("n_pixels_y", C.c_int32),
        ("pixel_pitch", C.c_double), ("tpc_borders", C.c_double * 2 * 3 * MAX_TPC),
    ]


class Config(C.Structure):
    _fields_ = [
        ("number_pix_neighbors", C.c_int32), ("nb_tran_diff_bins", C.c_int32),
        ("nb_sampling_bins_per_pixel", C.c_int32), ("signal_length", C.c_int32), ("n_ticks", C.c_int32),
        ("max_adc_values", C.c_int32), ("adc_interval", C.c_int32), ("n_templates", C.c_int32),
        ("long_diff_extent", C.c_int32), ("long_diff_template", C.c_double * MAX_TEMPLATES),
    ]


class ParamsC(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "eField", "Ab", "kb", "lifetime", "vdrift", "long_diff", "tran_diff",
        "lArDensity", "alpha", "beta", "R_param", "MeVToElectrons", "vdrift_eff", "dvdrift_deField",
        "shift_x", "shift_y", "shift_z", "size_margin", "t_sampling",
        "threshold", "gain", "v_pedestal", "v_cm", "v_ref", "adc_counts")] + [
        ("recombination_mode", C.c_int32), ("pad_", C.c_int32)]


_fp = C.POINTER(C.c_float)


class Segments(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("dE", C.c_void_p), ("dEdx", C.c_void_p),
                ("event_id", C.c_void_p), ("z_start", C.c_void_p), ("z_end", C.c_void_p), ("dx", C.c_void_p),
                ("stride", C.c_int64), ("n", C.c_int64)]


class FwdOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "n_rows", "pixel_ids", "wfs", "adc", "ticks", "tape", "n_hits", "hit_adc", "hit_x", "hit_y", "hit_z",
        "hit_tick", "hit_prob", "hit_event", "hit_pixel", "hit_index", "status")]


class ChopLayout(C.Structure):
    _fields_ = [("n_fields", C.c_int32), ("start", C.c_int32 * 3), ("end", C.c_int32 * 3), ("mid", C.c_int32 * 3),
                ("dE", C.c_int32), ("dx", C.c_int32), ("event", C.c_int32),
                ("n_zero", C.c_int32), ("zero", C.c_int32 * 8), ("n_neg_one", C.c_int32), ("neg_one", C.c_int32 * 8)]


class LsimError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LsimError(f"{path} not found: build it with `python larnd-sim-jax_amd/build.py` "
                        "(or __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, i32, i64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t
    lib.lsim_last_error.restype = C.c_char_p
    lib.lsim_tape_bytes.restype = sz
    sig = {
        "lsim_ctx_create": [C.POINTER(vp), C.POINTER(Geometry), C.POINTER(Config)],
        "lsim_ctx_destroy": [vp],
        "lsim_lut_set": [vp, vp, i32, i32, i32, vp],
        "lsim_forward_query": [vp, i64, i32, C.POINTER(sz), C.POINTER(i64)],
        "lsim_forward_prepare": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, vp, sz, vp, vp, vp],
        "lsim_forward_rows": [vp, C.POINTER(ParamsC), i64, i32, vp, sz, i64, C.POINTER(FwdOut), vp],
        "lsim_forward": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, vp, sz, i64, C.POINTER(FwdOut), vp],
        "lsim_fee_query": [vp, i64, C.POINTER(sz)],
        "lsim_fee_forward_noalloc": [vp, C.POINTER(ParamsC), vp, vp, i64, C.POINTER(FwdOut), vp],
        "lsim_negative_rows_query": [vp, C.POINTER(sz), C.POINTER(i64)],
        "lsim_negative_rows": [vp, C.POINTER(ParamsC), C.POINTER(Segments), vp, sz, vp, i64, vp, vp],
        "lsim_fee_forward": [vp, C.POINTER(ParamsC), vp, vp, i64, vp, sz, C.POINTER(FwdOut), vp],
        "lsim_backward_query": [vp, i64, i64, C.POINTER(sz)],
        "lsim_backward": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, vp, sz, i64, C.POINTER(FwdOut), vp, vp,
                          vp, sz, vp, vp],
        "lsim_backward_wfs": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, vp, sz, i64, C.POINTER(FwdOut), vp,
                              vp, vp, vp, sz, vp, vp],
        "lsim_set_row_waves": [C.c_int],
        "lsim_set_row_cost": [C.c_int],
        "lsim_set_row_tile": [C.c_int],
        "lsim_segment_records": [vp, i64, i32, vp, sz, vp, vp, vp, vp],
        "lsim_timing": [C.c_int],
        "lsim_timing_read": [C.POINTER(C.c_double), C.POINTER(i64)],
        "lsim_rbf_sums": [vp, i64, vp, vp, vp, vp, i64, C.c_float, vp, vp],
        "lsim_rbf_sums_grid": [vp, vp, i64, vp, vp, vp, i64, vp, i32, C.c_float, vp, vp],
        "lsim_nearest": [vp, i64, vp, vp, vp, i64, i64, vp, vp, vp],
        "lsim_prob_fee_query": [i64, i32, i32, i32, C.POINTER(sz)],
        "lsim_prob_fee": [vp, i64, i32, i32, i32, i32, C.c_float, C.c_float, C.c_float, C.c_float, vp, sz, vp, vp, vp],
        "lsim_prob_fee_backward": [i64, i32, i32, i32, i32, C.c_float, C.c_float, C.c_float, vp, sz, vp, vp, vp, vp,
                                   vp],
        "lsim_param_electrons": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, i32, vp, vp, i32, vp, vp, vp, vp,
                                 vp, vp, vp, vp],
        "lsim_param_rows": [vp, C.POINTER(ParamsC), vp, vp, vp, vp, vp, vp, i64, vp, vp, i32, vp, vp],
        "lsim_param_backward_query": [vp, i64, i64, C.POINTER(sz)],
        "lsim_param_backward": [vp, C.POINTER(ParamsC), C.POINTER(Segments), i32, i32, vp, vp, i32, vp, vp, vp, vp, vp,
                                vp, i64, vp, vp, vp, vp, vp, vp, vp, sz, vp, vp],
        "lsim_sdtw_query": [i64, i64, C.POINTER(sz)],
        "lsim_sdtw": [vp, i64, vp, i64, C.c_float, vp, sz, vp, vp, vp],
        "lsim_chop_query": [i64, C.POINTER(sz)],
        "lsim_chop_count": [C.POINTER(ChopLayout), vp, i64, C.c_double, vp, sz, vp, vp],
        "lsim_chop_fill": [C.POINTER(ChopLayout), vp, i64, C.c_double, vp, sz, i64, i64, vp, vp, i64, i64, vp, vp,
                           vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.lsim_rbf_cull_radius.argtypes = [C.c_float]
    lib.lsim_rbf_cull_radius.restype = C.c_float
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = _lib.lsim_last_error().decode() if _lib is not None else "?"
        raise LsimError(f"lsim error {rc}: {msg}")
