"""ctypes binding of libexageostat_gpu.so (the C ABI in include/exageostat_gpu.h).

The structures below mirror the header field for field; ``tests`` check
that every symbol the header declares is exported.  Loading fails loudly
when the shared library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexageostat_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "exageostat_gpu.h")

c_double_p = ctypes.POINTER(ctypes.c_double)

ChamUpper = 121
ChamLower = 122
ChamUpperLower = 123
ChamNoTrans = 111
ChamTrans = 112


class Location(ctypes.Structure):
    _fields_ = [("x", c_double_p), ("y", c_double_p), ("z", c_double_p)]


class Desc(ctypes.Structure):
    _fields_ = [
        ("mat", ctypes.c_void_p), ("dtyp", ctypes.c_int), ("mb", ctypes.c_int), ("nb", ctypes.c_int),
        ("bsiz", ctypes.c_int), ("lm", ctypes.c_int), ("ln", ctypes.c_int), ("m", ctypes.c_int),
        ("n", ctypes.c_int), ("mt", ctypes.c_int), ("nt", ctypes.c_int), ("uplo", ctypes.c_int),
    ]


class MLEData(ctypes.Structure):
    _fields_ = [
        ("variance", ctypes.c_double), ("variance1", ctypes.c_double), ("variance2", ctypes.c_double),
        ("computation", ctypes.c_char_p), ("test", ctypes.c_int), ("async_", ctypes.c_int),
        ("iter_count", ctypes.c_int), ("l1", Location), ("lmiss", Location), ("lobs", Location),
        ("lm", Location), ("descC", ctypes.c_void_p), ("descZ", ctypes.c_void_p),
        ("descZcpy", ctypes.c_void_p), ("descdet", ctypes.c_void_p), ("descproduct", ctypes.c_void_p),
        ("descZmiss", ctypes.c_void_p), ("descC12", ctypes.c_void_p), ("descC22", ctypes.c_void_p),
        ("descZactual", ctypes.c_void_p), ("descZobs", ctypes.c_void_p), ("descmse", ctypes.c_void_p),
        ("sequence", ctypes.c_void_p), ("request", ctypes.c_void_p), ("verbose", ctypes.c_int),
        ("log", ctypes.c_int), ("avg_exec_time_per_iter", ctypes.c_double),
        ("total_exec_time", ctypes.c_double), ("avg_flops_per_iter", ctypes.c_double),
        ("avg_generation_per_iter", ctypes.c_double), ("avg_solve_per_iter", ctypes.c_double),
        ("avg_cholesky_per_iter", ctypes.c_double), ("final_loglik", ctypes.c_double),
        ("recovery_file", ctypes.c_char_p), ("checkpoint_file", ctypes.c_char_p),
        ("det", ctypes.c_double), ("dotp", ctypes.c_double), ("mserror", ctypes.c_double),
        ("mserror1", ctypes.c_double), ("mserror2", ctypes.c_double), ("dm", ctypes.c_char_p),
        ("diag_thick", ctypes.c_int), ("opt_tol", ctypes.c_int), ("opt_max_iters", ctypes.c_int),
        ("kernel_fun", ctypes.c_char_p), ("precision", ctypes.c_int), ("mspe", ctypes.c_int),
        ("seed", ctypes.c_int), ("mloe", ctypes.c_double), ("mmom", ctypes.c_double),
        ("gpu", ctypes.c_void_p), ("silent", ctypes.c_int),
        ("status", ctypes.c_int), ("logdet", ctypes.c_double), ("time_gen", ctypes.c_double),
        ("time_facto", ctypes.c_double), ("time_logdet", ctypes.c_double), ("time_solve", ctypes.c_double),
        ("time_total", ctypes.c_double), ("mp_fp32_corr", ctypes.c_double), ("mp_bf16", ctypes.c_int),
        ("mp_ir_steps", ctypes.c_int), ("mp_fp64_corr", ctypes.c_double), ("mp_lolo", ctypes.c_int),
    ]


_P = ctypes.POINTER
_SIGNATURES = {
    "exageostat_init": (None, [_P(ctypes.c_int)] * 4),
    "exageostat_finalize": (None, []),
    "EXAGEOSTAT_dmle_Call": (None, [_P(MLEData)] + [ctypes.c_int] * 8),
    "MLE_Finalize": (None, [_P(MLEData)]),
    "MLE_get_zobs": (None, [_P(MLEData), c_double_p, ctypes.c_int]),
    "GenerateXYLoc": (_P(Location), [ctypes.c_int, ctypes.c_int]),
    "pick_random_points": (None, [_P(MLEData), c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "MLE_zvg": (None, [_P(MLEData), c_double_p, c_double_p] + [ctypes.c_int] * 5),
    "EXAGEOSTAT_MLE_dzvg_Tile": (None, [_P(MLEData), c_double_p, c_double_p] + [ctypes.c_int] * 3),
    "MLE_alg": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_dmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "MORSE_dmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_sdmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_MLE_dcmg_Tile_Async": (ctypes.c_int, [ctypes.c_int, _P(Desc), _P(Location), _P(Location),
                                                      _P(Location), c_double_p, ctypes.c_char_p,
                                                      ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]),
    "MORSE_MLE_dcmg_Tile_Async": (ctypes.c_int, [ctypes.c_int, _P(Desc), _P(Location), _P(Location),
                                                 _P(Location), c_double_p, ctypes.c_char_p, ctypes.c_char_p,
                                                 ctypes.c_void_p, ctypes.c_void_p]),
    "EXAGEOSTAT_MLE_dmdet_Tile_Async": (ctypes.c_int, [_P(Desc), ctypes.c_void_p, ctypes.c_void_p, _P(Desc)]),
    "MORSE_MLE_dmdet_Tile_Async": (ctypes.c_int, [_P(Desc), ctypes.c_void_p, ctypes.c_void_p, _P(Desc)]),
    "EXAGEOSTAT_dmle_Predict_Allocate": (None, [_P(MLEData)] + [ctypes.c_int] * 6),
    "EXAGEOSTAT_dmle_Predict_Tile": (ctypes.c_double, [_P(MLEData), c_double_p, ctypes.c_int, ctypes.c_int,
                                                       c_double_p, c_double_p, c_double_p, ctypes.c_int]),
    "MORSE_dmle_Predict_Tile": (ctypes.c_double, [_P(MLEData), c_double_p, ctypes.c_int, ctypes.c_int,
                                                  c_double_p, c_double_p, c_double_p, ctypes.c_int]),
    "prediction_init": (None, [_P(MLEData)] + [ctypes.c_int] * 6),
    "prediction_finalize": (None, [_P(MLEData)]),
    "EXAGEOSTAT_Desc_Create": (ctypes.c_int, [_P(_P(Desc)), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "EXAGEOSTAT_Desc_Create_User": (ctypes.c_int, [_P(_P(Desc)), c_double_p, ctypes.c_int, ctypes.c_int]),
    "EXAGEOSTAT_Desc_Destroy": (ctypes.c_int, [_P(_P(Desc))]),
    "EXAGEOSTAT_Sequence_Create": (ctypes.c_int, [_P(ctypes.c_void_p)]),
    "EXAGEOSTAT_Sequence_Wait": (ctypes.c_int, [ctypes.c_void_p]),
    "EXAGEOSTAT_Sequence_Destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "EXAGEOSTAT_Lapack_to_Tile": (ctypes.c_int, [c_double_p, ctypes.c_int, _P(Desc)]),
    "EXAGEOSTAT_Tile_to_Lapack": (ctypes.c_int, [_P(Desc), c_double_p, ctypes.c_int]),
    "EXAGEOSTAT_dpotrf_Tile": (ctypes.c_int, [ctypes.c_int, _P(Desc)]),
    "EXAGEOSTAT_dpotri_Tile": (ctypes.c_int, [ctypes.c_int, _P(Desc)]),
    "EXAGEOSTAT_dtrsv_Tile": (ctypes.c_int, [ctypes.c_int, _P(Desc), c_double_p]),
    "exageostat_dlarnv": (None, [ctypes.c_int, _P(ctypes.c_int), ctypes.c_int, c_double_p]),
    "exageostat_gpu_synchronize": (None, []),
    "exageostat_gpu_profile_enable": (None, [ctypes.c_int]),
    "exageostat_gpu_profile_reset": (None, []),
    "exageostat_mixed_prec_map": (ctypes.c_long, [_P(MLEData), _P(ctypes.c_ubyte), ctypes.c_long]),
    "exageostat_fisher": (ctypes.c_int, [_P(MLEData), _P(ctypes.c_double), _P(ctypes.c_double),
                                         _P(ctypes.c_double)]),
    "EXAGEOSTAT_Fisher_Tile": (_P(ctypes.c_double), [_P(MLEData), ctypes.c_int, _P(ctypes.c_double), ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int]),
    "EXAGEOSTAT_dmle_mloe_mmom_Allocate": (None, [_P(MLEData)] + [ctypes.c_int] * 5),
    "mloe_mmom_init": (None, [_P(MLEData)] + [ctypes.c_int] * 5),
    "EXAGEOSTAT_dmle_mloe_mmom_Tile": (None, [_P(MLEData), c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int]),
    "EXAGEOSTAT_dmle_mloe_mmom_Tile_Async": (None, [_P(MLEData), c_double_p, c_double_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int]),
    "mloe_mmom_finalize": (None, [_P(MLEData)]),
    "MLOE_MMOM_Finalize": (None, [_P(MLEData)]),
    "exageostat_mloe_mmom": (ctypes.c_int, [_P(MLEData), c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                            c_double_p, c_double_p]),
    "exageostat_gpu_profile_get": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_long), c_double_p, c_double_p,
                                                  c_double_p]),
    "exageostat_gpu_last_status": (ctypes.c_int, []),
    "exageostat_gpu_last_error": (ctypes.c_char_p, []),
    "exageostat_gpu_device_count": (ctypes.c_int, []),
}


def header_symbols(path: str = HEADER_PATH) -> list[str]:
    """Function names declared in include/exageostat_gpu.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


_lib = None


def load() -> ctypes.CDLL:
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C exageostat_amd`; there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
