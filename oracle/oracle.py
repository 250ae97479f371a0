"""ctypes binding of the CPU oracle (liboxav1_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
The oracle takes the same ABI structs as the HIP library's drop-in entry
points (include/oxav1_pf.h).
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oxideav-av1_amd"))

from oxav1 import abi  # noqa: E402
from oxav1.dropin import FilterCalls  # noqa: E402
from oxav1.frames import planes_struct  # noqa: E402

LIB_PATH = os.path.join(HERE, "liboxav1_oracle.so")
# Built by `make check` with -fsanitize=undefined (no recovery: the first
# signed overflow, bad shift or out-of-range index aborts the process);
# selected with OXAV1_ORACLE_SAN=ubsan.
UBSAN_PATH = os.path.join(HERE, "liboxav1_oracle_ubsan.so")

ORACLE_NAMES = {
    "lf": "oo_loop_filter_frame",
    "cdef": "oo_cdef_frame",
    "sr": "oo_upscale_frame",
    "lr": "oo_loop_restoration_frame",
    "fg": "oo_film_grain_synthesis",
}


class OoStrength(C.Structure):
    _fields_ = [("lvl", C.c_uint8), ("limit", C.c_uint8), ("blimit", C.c_uint8), ("thresh", C.c_uint8)]


class OoMasks(C.Structure):
    _fields_ = [("hev_mask", C.c_uint8), ("filter_mask", C.c_uint8), ("flat_mask", C.c_uint8),
                ("flat_mask2", C.c_uint8)]


class OoLrGeom(C.Structure):
    _fields_ = [("unit_row", C.c_uint32), ("unit_col", C.c_uint32), ("x", C.c_uint32),
                ("y", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("stripe_start_y", C.c_int32), ("stripe_end_y", C.c_int32),
                ("plane_end_x", C.c_int32), ("plane_end_y", C.c_int32)]


class OoChainGrids(C.Structure):
    _fields_ = [("lf", abi.LfGrids), ("cdef", abi.CdefGrids), ("lr", abi.LrGrids)]


_P = C.POINTER
_SIGS = {
    "oo_loop_filter_frame": (None, [_P(abi.LfCtx), _P(abi.LfGrids), _P(abi.Plane), C.c_uint32]),
    "oo_loop_filter_frame_parallel": (None, [_P(abi.LfCtx), _P(abi.LfGrids), _P(abi.Plane),
                                             C.c_uint32]),
    "oo_loop_filter_frame_reversed": (None, [_P(abi.LfCtx), _P(abi.LfGrids), _P(abi.Plane),
                                             C.c_uint32]),
    "oo_cdef_frame": (None, [_P(abi.CdefCtx), _P(abi.CdefGrids), _P(abi.Plane), C.c_uint32,
                             _P(abi.Plane), C.c_uint32]),
    "oo_upscale_frame": (C.c_int, [_P(abi.SrCtx), _P(abi.Plane), C.c_uint32, _P(abi.Plane),
                                   C.c_uint32]),
    "oo_loop_restoration_frame": (None, [_P(abi.LrCtx), _P(abi.LrGrids), _P(abi.Plane),
                                         C.c_uint32, _P(abi.Plane), C.c_uint32, _P(abi.Plane),
                                         C.c_uint32]),
    "oo_film_grain_synthesis": (None, [_P(abi.FilmGrainParams), C.c_uint8, C.c_uint8, C.c_uint8,
                                       C.c_uint8, C.c_uint8, _P(abi.Plane), C.c_uint32]),
    "oo_lf_filter_size": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8]),
    "oo_lf_adaptive_filter_strength": (OoStrength, [_P(abi.LfCtx), _P(abi.LfGrids), C.c_uint32,
                                                    C.c_uint32, C.c_uint8, C.c_uint8]),
    "oo_lf_filter_mask": (OoMasks, [_P(abi.Plane), C.c_int32, C.c_int32, C.c_uint8, C.c_uint8,
                                    C.c_uint8, C.c_uint8, C.c_int32, C.c_int32, C.c_uint32,
                                    C.c_uint8]),
    "oo_lf_narrow_filter": (None, [_P(abi.Plane), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_int, C.c_uint8]),
    "oo_lf_wide_filter": (None, [_P(abi.Plane), C.c_int32, C.c_int32, C.c_uint8, C.c_int32,
                                 C.c_int32, C.c_uint32]),
    "oo_get_plane_residual_size": (C.c_int32, [C.c_uint32, C.c_uint8, C.c_uint8, C.c_uint8]),
    "oo_get_tx_size": (C.c_int32, [C.c_uint8, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8]),
    "oo_cdef_direction": (None, [_P(abi.CdefCtx), _P(abi.Plane), C.c_uint32, C.c_uint32,
                                 C.c_uint32, _P(C.c_int32), _P(C.c_int32)]),
    "oo_cdef_filter_block": (None, [_P(abi.CdefCtx), _P(abi.Plane), C.c_uint32, _P(abi.Plane),
                                    C.c_uint32, C.c_uint8, C.c_uint32, C.c_uint32, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_int32]),
    "oo_cdef_constrain": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32]),
    "oo_upscale_sample": (C.c_int32, [_P(C.c_int32), C.c_uint32, C.c_int32, C.c_uint32,
                                      C.c_int32, C.c_int32, C.c_uint8]),
    "oo_count_units_in_frame": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "oo_lr_derive_block_geometry": (OoLrGeom, [_P(abi.LrCtx), C.c_uint8, C.c_uint32,
                                               C.c_uint32]),
    "oo_lr_get_source_sample": (C.c_int32, [_P(abi.Plane), C.c_uint32, _P(abi.Plane),
                                            C.c_uint32, C.c_uint8, C.c_int32, C.c_int32,
                                            _P(OoLrGeom)]),
    "oo_wiener_coefficients": (None, [_P(C.c_int32), _P(C.c_int32)]),
    "oo_lr_box_filter": (None, [_P(abi.Plane), C.c_uint32, _P(abi.Plane), C.c_uint32,
                                C.c_uint8, _P(OoLrGeom), C.c_uint8, C.c_int32, C.c_int32,
                                C.c_uint8, _P(C.c_int32)]),
    "oo_fg_random": (C.c_uint32, [_P(C.c_uint16), C.c_uint32]),
    "oo_fg_generate_grain": (C.c_uint32, [_P(abi.FilmGrainParams), C.c_uint8, C.c_uint8,
                                          C.c_uint8, _P(C.c_int32), _P(C.c_int32),
                                          _P(C.c_int32)]),
    "oo_fg_scaling_lookup_init": (None, [_P(abi.FilmGrainParams), C.c_uint8, _P(C.c_uint8)]),
    "oo_fg_scale_lut": (C.c_int32, [_P(C.c_uint8), C.c_uint8, C.c_int32, C.c_uint8]),
    "oo_chain": (C.c_int, [_P(abi.FrameDesc), _P(OoChainGrids), _P(abi.Plane), _P(abi.Plane),
                           _P(abi.Plane), _P(C.c_double)]),
    "oo_table_checksum": (C.c_int64, [C.c_char_p]),
    "oo_cdef_elect": (C.c_int, [_P(abi.CdefElectInput), _P(abi.CdefElection)]),
    "oo_superres_coded_width": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "oo_superres_candidate_denoms": (C.c_uint32, [C.c_uint32, _P(C.c_uint32)]),
    "oo_superres_probe": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint8]),
    "oo_superres_downscale": (None, [C.c_void_p * 3, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8, C.c_uint8,
                                     C.c_uint32, C.c_void_p * 3]),
    "oo_superres_upscale_recon": (C.c_int, [C.c_void_p * 3, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint8, C.c_uint8, C.c_uint8, C.c_uint8, C.c_void_p * 3]),
    "oo_fg_denoise": (None, [C.c_void_p * 3, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8, C.c_uint8,
                             C.c_void_p * 3]),
    "oo_fg_noise_estimate": (None, [C.c_void_p * 3, C.c_void_p * 3, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8,
                                    C.c_uint8, C.c_uint8, _P(abi.NoiseEstimate)]),
    "oo_fg_plane_sigma16": (C.c_uint64, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint8]),
    "oo_fg_apply_grain": (None, [_P(abi.FilmGrainParams), C.c_void_p * 3, C.c_uint32, C.c_uint32, C.c_uint8,
                                 C.c_uint8, C.c_uint8, C.c_uint8, C.c_uint8, C.c_void_p * 3]),
    "oo_cdef_run_units": (None, [_P(abi.CdefElectInput), _P(abi.CdefParams), C.c_int, _P(C.c_uint64)]),
    "oo_lr_elect": (C.c_int, [_P(abi.LrElectInput), _P(abi.LrElection)]),
    "oo_lr_apply_plan": (C.c_int, [_P(abi.LrElection), _P(abi.LrElectInput), C.c_void_p * 3, _P(C.c_uint64)]),
    "oo_lr_unit_rects": (C.c_uint32, [_P(abi.LrElectInput), C.c_uint8, C.c_uint32, C.c_uint32, _P(C.c_uint32),
                                      C.c_uint32]),
    "oo_lr_fit_wiener": (None, [_P(C.c_int32), _P(C.c_uint16), C.c_uint32, C.c_uint32, _P(C.c_uint32), C.c_uint32,
                                C.c_int, _P(C.c_int32), _P(C.c_double)]),
    "oo_log2_frac256": (C.c_uint32, [C.c_uint32]),
    "oo_recenter": (C.c_int64, [C.c_int64, C.c_int64]),
    "oo_cw_cost": (C.c_uint64, [C.c_uint32, C.c_int, C.c_int64, C.c_int64, C.c_uint32, C.c_int64, C.c_int64,
                                _P(C.c_uint16)]),
    "oo_cw_symbol_run": (C.c_uint64, [C.c_uint32, C.c_uint32, _P(C.c_uint16), C.c_uint32, C.c_uint32]),
    "oo_cw_signed_ref_bits": (C.c_uint32, [C.c_int64, C.c_int64, C.c_uint32, C.c_int64, C.c_int64,
                                           _P(C.c_uint8), C.c_uint32, _P(C.c_uint64)]),
    "oo_lr_price_unit": (C.c_uint64, [_P(C.c_int32), _P(C.c_int32), C.c_int, _P(abi.LrUnit), _P(C.c_uint16)]),
}

_lib = None


def _bind(lib):
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


def load_path(path):
    """A separately built oracle library (a sanitizer or mutant build) with
    the same signatures; symbols it lacks stay unbound."""
    return _bind(C.CDLL(path))


def build():
    """Compile the oracle with its Makefile (gcc, -O3)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = LIB_PATH
    san = os.environ.get("OXAV1_ORACLE_SAN", "")
    if san:
        if san != "ubsan":
            raise RuntimeError(f"OXAV1_ORACLE_SAN={san!r}: only 'ubsan' is built")
        path = UBSAN_PATH
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, "liboxav1_oracle_ubsan.so"], check=True)
    elif not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    for name in _SIGS:
        getattr(lib, name)  # every symbol must be there
    _lib = _bind(lib)
    return lib


def calls():
    return FilterCalls(load(), ORACLE_NAMES, status_returning=False)


def chain(frame):
    """The reference chain (frame_driver.rs:790-1116) on a Frame.

    Returns (ref_out planes, out planes, stage_ns[5]).
    """
    lib = load()
    rec = [p.copy() for p in frame.planes]
    ref = [np.zeros(frame.out_shape(p), dtype=np.int32) for p in range(frame.num_planes)]
    out = [np.zeros(frame.out_shape(p), dtype=np.int32) for p in range(frame.num_planes)]
    _, lfg, k1 = frame.lf_args()
    _, cg, k2 = frame.cdef_args()
    _, lrg, k3 = frame.lr_args()
    g = OoChainGrids(lf=lfg, cdef=cg, lr=lrg)
    ns = (C.c_double * 5)()
    st = lib.oo_chain(C.byref(frame.desc), C.byref(g), planes_struct(rec), planes_struct(ref),
                      planes_struct(out), ns)
    if st != 0:
        raise RuntimeError(f"oracle chain status {st}")
    return ref, out, list(ns)
