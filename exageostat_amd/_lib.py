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


_D, _I, _F = ctypes.c_double, ctypes.c_int, ctypes.c_float
_S, _V, _DP = ctypes.c_char_p, ctypes.c_void_p, c_double_p


class MLEData(ctypes.Structure):
    """MLE_data: the reference's fields in its order (FILE* as void*), then
    this implementation's fields; `async_` is the C field `async`."""
    _fields_ = [
        ("variance", _D), ("variance1", _D), ("variance2", _D), ("variance3", _D), ("computation", _S),
        ("c_fun", _S), ("test", _I), ("async_", _I), ("iter_count", _I), ("l1", Location), ("lmiss", Location),
        ("lobs", Location), ("lm", Location), ("descC", _V), ("descsubC11", _V), ("descsubC12", _V),
        ("descsubC21", _V), ("descsubC22", _V), ("descZ", _V), ("descZ1", _V), ("descZ2", _V), ("descZ3", _V),
        ("Adense", _DP), ("Adense2", _DP), ("descZcpy", _V), ("descdet", _V), ("descsum", _V),
        ("descproduct", _V), ("descproduct1", _V), ("descproduct2", _V), ("descproduct3", _V),
        ("descZmiss", _V), ("descC11", _V), ("descC21", _V), ("descC12", _V), ("descC22", _V),
        ("descZactual", _V), ("descZobs", _V), ("descmse", _V), ("descmse1", _V), ("descmse2", _V),
        ("descmse3", _V), ("descZtrace", _V), ("sequence", _V), ("request", _V), ("verbose", _I),
        ("check", _I), ("log", _I), ("avg_exec_time_per_iter", _D), ("total_exec_time", _D),
        ("avg_flops_per_iter", _D), ("avg_generation_per_iter", _D), ("avg_solve_per_iter", _D),
        ("avg_cholesky_per_iter", _D), ("avg_exec_time_gen_stage", _D), ("avg_flops_gen_stage", _D),
        ("final_loglik", _D), ("locsFPath", _S), ("timeFPath", _S), ("obsFPath", _S), ("obsFPath2", _S),
        ("obsFPath3", _S), ("actualZFPath", _S), ("actualZFPath2", _S), ("actualZFPath3", _S),
        ("actualZLocFPath", _S), ("actualTimeFPath", _S), ("det", _D), ("sum", _D), ("dotp", _D),
        ("dotp1", _D), ("dotp2", _D), ("dotp3", _D), ("sdotp", _F), ("mserror", _D), ("mserror1", _D),
        ("mserror2", _D), ("mserror3", _D), ("dm", _S), ("diag_thick", _I), ("nFileLog", _S), ("pFileLog", _V),
        ("hicma_maxrank", _I), ("hicma_data_type", _I), ("hicma_descC", _V), ("hicma_descZ", _V),
        ("hicma_descCD", _V), ("hicma_descCUV", _V), ("hicma_descCrk", _V), ("hicma_descZcpy", _V),
        ("hicma_descdet", _V), ("hicma_descproduct", _V), ("hicma_descC12D", _V), ("hicma_descC12UV", _V),
        ("hicma_descC12rk", _V), ("hicma_descC22D", _V), ("hicma_descC22UV", _V), ("hicma_descC22rk", _V),
        ("hicma_acc", _D), ("hsequence", _V), ("hrequest", _V), ("opt_tol", _I), ("opt_max_iters", _I),
        ("ooc", _I), ("kernel_fun", _S), ("precision", _I), ("desctemp", _V), ("desctemp22", _V),
        ("desck_t", _V), ("desck_a", _V), ("desck_ttmp", _V), ("desck_atmp", _V), ("descK_t", _V),
        ("descK_a", _V), ("descexpr1", _V), ("descexpr2", _V), ("descexpr3", _V), ("descexpr4", _V),
        ("descestimatedalpha", _V), ("desctruthalpha", _V), ("desc_mloe_mmom", _V), ("descmloe", _V),
        ("descmmom", _V), ("expr1", _D), ("expr2", _D), ("expr3", _D), ("expr4", _D), ("mloe", _D),
        ("mmom", _D), ("mloe_mmom", _I), ("mloe_mmom_async", _I), ("mspe", _I), ("recovery_file", _S),
        ("checkpoint_file", _S), ("time_slots", _I), ("idw", _I), ("fisher", _I), ("descr", _V),
        ("descrcpy", _V), ("seed", _I), ("gpu", _V), ("silent", _I), ("status", _I), ("logdet", _D),
        ("time_gen", _D), ("time_facto", _D), ("time_logdet", _D), ("time_solve", _D), ("time_total", _D),
        ("mp_fp32_corr", _D), ("mp_bf16", _I), ("mp_ir_steps", _I), ("mp_fp64_corr", _D), ("mp_lolo", _I),
        ("mp_ir_tol", _D), ("mp_ir_done", _I), ("mp_ir_change", _D), ("mp_lean", _I),
        ("mp_grad_f32", _I), ("mp_grad_check", _D),
    ]


_P = ctypes.POINTER
_IP = ctypes.POINTER(ctypes.c_int)
# exageostat_func_t: the NLOPT objective type (MLE_alg's signature)
OBJECTIVE = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p)
_SIGNATURES = {
    "exageostat_init": (None, [_P(ctypes.c_int)] * 4),
    "exageostat_finalize": (None, []),
    "EXAGEOSTAT_dmle_Call": (None, [_P(MLEData)] + [ctypes.c_int] * 8),
    "MLE_Finalize": (None, [_P(MLEData)]),
    "MLE_get_zobs": (None, [_P(MLEData), c_double_p, ctypes.c_int]),
    "GenerateXYLoc": (_P(Location), [ctypes.c_int, ctypes.c_int]),
    "GenerateXYLoc_ST": (_P(Location), [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pick_random_points": (None, [_P(MLEData), c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "MLE_zvg": (None, [_P(MLEData), c_double_p, c_double_p] + [ctypes.c_int] * 5),
    "EXAGEOSTAT_MLE_dzvg_Tile": (None, [_P(MLEData), c_double_p, c_double_p] + [ctypes.c_int] * 3),
    "MLE_alg": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_dmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "MORSE_dmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_sdmle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_dmle_Tile_Async": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_sdmle_Tile_Async": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_smle_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_smle_Tile_Async": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_dmle_diag_Tile": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_dmle_diag_Tile_Async": (ctypes.c_double, [ctypes.c_uint, c_double_p, c_double_p, ctypes.c_void_p]),
    "EXAGEOSTAT_smle_Call": (None, [_P(MLEData)] + [ctypes.c_int] * 8),
    "EXAGEOSTAT_sdmle_Call": (None, [_P(MLEData)] + [ctypes.c_int] * 8),
    "EXAGEOSTAT_dmle_diag_Call": (None, [_P(MLEData)] + [ctypes.c_int] * 8),
    "EXAGEOSTAT_MLE_dzvg_Tile_Async": (None, [_P(MLEData), c_double_p, c_double_p] + [ctypes.c_int] * 3),
    "EXAGEOSTAT_dmle_Predict_Tile_Async": (ctypes.c_double, [_P(MLEData), c_double_p] + [ctypes.c_int] * 3),
    "EXAGEOSTAT_smle_Predict_Tile": (ctypes.c_double, [_P(MLEData), c_double_p, ctypes.c_int, ctypes.c_int,
                                                       c_double_p, c_double_p, c_double_p, ctypes.c_int]),
    "EXAGEOSTAT_sdmle_Predict_Tile": (ctypes.c_double, [_P(MLEData), c_double_p, ctypes.c_int, ctypes.c_int,
                                                        c_double_p, c_double_p, c_double_p, ctypes.c_int]),
    "EXAGEOSTAT_dmle_diag_Predict_Tile": (ctypes.c_double, [_P(MLEData), c_double_p, ctypes.c_int, ctypes.c_int,
                                                            c_double_p, c_double_p, c_double_p, ctypes.c_int]),
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
    "EXAGEOSTAT_dtrsv_Tile_noalloc": (ctypes.c_int, [ctypes.c_int, _P(Desc), c_double_p]),
    "exageostat_dlarnv": (None, [ctypes.c_int, _P(ctypes.c_int), ctypes.c_int, c_double_p]),
    "exageostat_bobyqa": (ctypes.c_int, [ctypes.c_uint, OBJECTIVE, ctypes.c_void_p, c_double_p, c_double_p,
                                         c_double_p, ctypes.c_double, ctypes.c_int, c_double_p,
                                         _P(ctypes.c_int)]),
    "init_data_values": (None, [_P(MLEData)]),
    "readLocsFile": (_P(Location), [ctypes.c_char_p, ctypes.c_int]),
    "readLocsFile3d": (_P(Location), [ctypes.c_char_p, ctypes.c_int]),
    "readObsFile": (c_double_p, [ctypes.c_char_p, ctypes.c_int]),
    "readTimeFile": (c_double_p, [ctypes.c_char_p, ctypes.c_int]),
    "rexageostat_init": (None, [_IP] * 4),
    "rexageostat_finalize": (None, []),
    "gen_z_givenlocs_exact": (None, [c_double_p, _IP, c_double_p, _IP, _IP, c_double_p] + [_IP] * 9 + [c_double_p]),
    "gen_z_exact": (None, [_IP, c_double_p] + [_IP] * 10 + [c_double_p]),
    "mle_exact": (None, [c_double_p, _IP] * 5 + [_IP] * 3 + [c_double_p] + [_IP] * 6 + [c_double_p]),
    "mle_dst": (None, [c_double_p, _IP] * 3 + [_IP, c_double_p, _IP, c_double_p, _IP, _IP, _IP, _IP, c_double_p]
                + [_IP] * 6 + [c_double_p]),
    "mle_mp": (None, [c_double_p, _IP] * 3 + [_IP, c_double_p, _IP, c_double_p, _IP, _IP, _IP, _IP, c_double_p]
               + [_IP] * 6 + [c_double_p]),
    "exact_predict": (None, [c_double_p, _IP] * 5 + [_IP, _IP, _IP, c_double_p] + [_IP] * 8 + [c_double_p]),
    "exact_mloe_mmom": (None, [c_double_p, _IP] * 5 + [_IP, _IP, _IP, c_double_p, c_double_p] + [_IP] * 7
                        + [c_double_p]),
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
    names = re.findall(r"^(?!typedef)[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
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
