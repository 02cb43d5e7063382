"""ctypes binding of ``libfasterlmm_amd.so`` (declared in include/fasterlmm_amd.h).

The library is the only compute path: if it is missing, or no ROCm GPU is visible, every entry point raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libfasterlmm_amd.so"

PREC_SPLIT_BF16 = 0
PREC_FP32 = 1
PREC_SPLIT_FP16 = 2
ROTATE_UY = 0
ROTATE_US = 1
ROTATE_US_X2 = 2
DELTA_GOLDEN = 0
DELTA_NEWTON = 1
DELTA_FIXED = 2
DELTA_GRID = 4
BED_I8 = 0
BED_F64 = 1
BED_F32 = 2
SCAN_SQUARE = 0
SCAN_WIDE = 1
SCAN_WIDE_FACTORED = 2

_P = C.c_void_p
_I64 = C.c_int64
_I = C.c_int
_D = C.c_double

NOALLOC_SCRATCH_BYTES = 64 << 20  # FLMM_NOALLOC_SCRATCH_BYTES

# flmm_wald_scan_screened and its two halves
_SCREENED = [_P, _P, _P, _P, _I64, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P, _I64, _P, _P, _I64,
             _D, _D, _P, _P, _D, _P, _P, _P, _P, _P, _P]

# name -> argtypes, in header order
SIGNATURES: dict[str, list] = {
    "flmm_abi_version": [],
    "flmm_last_error": [],
    "flmm_scan_table_bytes": [_I],
    "flmm_screen_work_bytes": [_I64, _I64],
    "flmm_screen_order_bytes": [_I64],
    "flmm_screen_q_rows": [_I64, _I64, _I64],
    "flmm_scan_table_fill": [_I, _I, _P, _I, _P],
    "flmm_wald_scan_grouped": [_I, _I, _P, _I, _P],
    "flmm_max_reduce_grouped": [_P, _I, _I64, _P],
    "flmm_standardise": [_I, _P, _I64, _I64, _I64, _P, _I64, _P, _P, _P],
    "flmm_bed_decode": [_P, _I64, _P, _I64, _P, _I64, _I, _I, _P, _I64, _P, _P, _P],
    "flmm_row_norms": [_P, _I64, _I64, _I64, _I64, _P, _P, _P],
    "flmm_pack_planes": [_I, _P, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _I64, _I64, _P],
    "flmm_pack_columns_projected": [_I, _P, _I64, _I64, _P, _I, _P, _P, _P, _P, _P, _I64, _I64, _P],
    "flmm_rotate": [_I, _I, _P, _P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P],
    "flmm_delta_weights": [_P, _I64, _P, _I64, _I64, _I, _D, _D, _P, _P, _P, _P],
    "flmm_delta_grid": [_P, _P, _I, _D, _P, _I64, _I64, _I64, _P, _P],
    "flmm_delta_grid_split": [_P, _P, _I, _D, _P, _P, _D, _D, _D, _I64, _I64, _I64, _P, _P, _P],
    "flmm_delta_refine": [_I, _I, _P, _I64, _I64, _I64, _P, _I64, _P, _I64, _D, _I, _D, _D, _P, _P, _P, _P, _P,
                          _P, _P, _P, _P, _P, _P, _I64, _I64, _D, _D, _I, _P, _P],
    "flmm_gram": [_I, _P, _P, _P, _I64, _I64, _I64, _I, _P, _P, _I64, _P],
    "flmm_gram_f64": [_P, _I64, _I64, _I64, _I, _P, _P, _I64, _P],
    "flmm_rotate_f64": [_P, _I64, _P, _I64, _I64, _P, _I64, _I64, _P, _I64, _I64, _P],
    "flmm_gemm_f64": [_P, _I64, _I64, _P, _I64, _P, _P],
    "flmm_rescore_f64": [_P, _P, _I64, _P, _I64, _P, _P, _P, _P, _I64, _I64, _P, _D, _P, _P, _P, _P, _I64, _P],
    "flmm_select_above": [_P, _I64, _I64, _I64, _D, _P, _P, _I64, _P, _P],
    "flmm_factored_weight_planes": [_I, _I64, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P, _I64, _P, _P],
    "flmm_delta_refine_learned": [_I, _P, _I64, _I64, _I64, _P, _I64, _P, _I64, _D, _I, _D, _D, _P, _P, _D, _D, _P,
                                  _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _D, _D, _I, _P, _P],
    "flmm_mlp_features": [_P, _I64, _I64, _I64, _P, _D, _D, _P, _P],
    "flmm_mlp_dense_fwd": [_P, _I64, _I, _P, _P, _I, _I, _P, _P],
    "flmm_mlp_dense_bwd": [_P, _P, _P, _P, _I64, _I, _I, _I, _P, _P, _P, _P, _P],
    "flmm_mlp_mse_grad": [_P, _P, _I64, _P, _P],
    "flmm_adam": [_P, _P, _P, _P, _I64, _D, _D, _D, _D, _I, _P],
    "flmm_wald_scan_screened": _SCREENED,
    "flmm_wald_screen": _SCREENED,
    "flmm_screen_resolve": _SCREENED,
    "flmm_wald_scan": [_I, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P,
                       _P, _D, _P, _P, _P, _P, _P, _P, _P, _P],
    "flmm_wald_scan_lowrank": [_I, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64,
                               _I64, _I64, _P, _P, _D, _P, _P, _P, _P, _P, _P, _P, _P],
    "flmm_wald_scan_lowrank_noalloc": [_I, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64,
                                       _I64, _I64, _I64, _P, _P, _D, _P, _P, _P, _P, _P, _P, _P, _P],
    "flmm_max_reduce": [_P, _I64, _I64, _I64, _P, _P],
    "flmm_neglog10p_f1": [_P, _I64, _D, _P, _P],
    "flmm_effect_size": [_P, _I64, _I64, _P, _P, _P, _P],
}

# entry points that do not return an int status
RESTYPES = {"flmm_last_error": C.c_char_p, "flmm_scan_table_bytes": C.c_int64, "flmm_screen_work_bytes": C.c_int64,
            "flmm_screen_order_bytes": C.c_int64, "flmm_screen_q_rows": C.c_int64}


class ScanGroup(C.Structure):
    """Mirror of ``flmm_scan_group``: one SNP block x column set of a grouped scan launch (device pointers)."""
    _fields_ = [
        ("s1_hi", _P), ("s1_lo", _P), ("s2_hi", _P), ("s2_lo", _P), ("s1_scale", _P), ("s2_scale", _P),
        ("mc_pad", _I64),
        ("q1_hi", _P), ("q1_lo", _P), ("q2_hi", _P), ("q2_lo", _P), ("q1_scale", _P), ("q2_scale", _P),
        ("p_pad", _I64), ("kpad", _I64), ("kpad2", _I64), ("mc", _I64), ("P", _I64), ("n_real", _I64),
        ("rWr", _P), ("h2", _P), ("nu", _D), ("row_index", _P),
        ("f", _P), ("beta", _P), ("se", _P), ("sfve", _P), ("nullh2", _P), ("pmax", _P), ("max_F", _P),
    ]


_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Open the shared library (once) and type its entry points."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m fasterlmm_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


class FlmmError(RuntimeError):
    pass


def call(name: str, *args) -> None:
    """Call an entry point; a non-zero status raises ValueError (bad argument) or FlmmError (HIP failure)."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.flmm_last_error().decode()
        if rc == 1:
            raise ValueError(msg)
        raise FlmmError(msg)


def device() -> torch.device:
    """The GPU every kernel runs on; raises when there is none (no CPU fallback exists)."""
    if not torch.cuda.is_available():
        raise RuntimeError("fasterlmm_amd needs a ROCm GPU (MI355X / gfx950); none is visible")
    load()
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()
