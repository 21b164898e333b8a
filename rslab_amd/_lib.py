"""ctypes binding of librslab_amd.so (C ABI declared in include/rslab_amd.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C rslab_amd/csrc``). There is no fallback: if the library is missing,
importing this module raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librslab_amd.so")

RSL_OK = 0
RSL_NUMERICALLY_RANK_DEFICIENT = 1
RSL_INVALID_INPUT = 2
RSL_DIMENSION_MISMATCH = 3
RSL_DEVICE_ERROR = 4
RSL_OUT_OF_MEMORY = 5


class Settings(C.Structure):
    """Mirror of ``rsl_settings``."""

    _fields_ = [
        ("ordering", C.c_int32),
        ("nemin", C.c_int32),
        ("relax", C.c_int32),
        ("relax_width", C.c_int32),
        ("relax_rows", C.c_int32),
        ("panel_nb", C.c_int32),
        ("pivot", C.c_int32),
        ("abs_floor", C.c_double),
        ("scaling", C.c_int32),
        ("precision", C.c_int32),
        ("nd_seed", C.c_uint64),
        ("wide_flops", C.c_double),
        ("drop_tol", C.c_double),
        ("cmod_cap", C.c_int64),
        ("graphs", C.c_int32),
        ("solve_wide", C.c_int32),
        ("method", C.c_int32),
        ("solve_dep", C.c_int32),
        ("solve_split", C.c_int32),
        ("pivot_u", C.c_double),
    ]


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)
_U8P = C.POINTER(C.c_uint8)

# name -> (restype, argtypes)
SIGNATURES = {
    "rsl_settings_default": (None, [C.POINTER(Settings)]),
    "rsl_analyze": (C.c_int, [C.c_int64, _I64P, _I64P, C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_symbolic_info": (C.c_int, [_P, _I64P, _DP]),
    "rsl_plan_create": (C.c_int, [C.c_int64, _I64P, _I64P, _I32P, C.c_int32, _I32P, _I32P, _I32P,
                                  C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_symbolic_replan": (C.c_int, [_P, C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_symbolic_export": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P, _I64P, _I32P]),
    "rsl_factor": (C.c_int, [_P, C.c_int32, _DP, C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_refactor": (C.c_int, [_P, _DP]),
    "rsl_solve": (C.c_int, [_P, _DP, C.c_int64, _DP]),
    "rsl_solve_device": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "rsl_solve_device_noalloc": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "rsl_factor_info": (C.c_int, [_P, _I64P, _I64P, _DP]),
    "rsl_factor_nnz": (C.c_int, [_P, _I64P]),
    "rsl_factor_download": (C.c_int, [_P, _I64P, _I32P, _DP, _DP, _DP, _U8P, _I32P]),
    "rsl_refine": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_double, _DP, _DP]),
    "rsl_refactor_device": (C.c_int, [_P, _P, _P]),
    "rsl_refine_device": (C.c_int, [_P, _P, _P, C.c_int64, C.c_double, _P, _DP, _P]),
    "rsl_free_factor": (None, [_P]),
    "rsl_free_symbolic": (None, [_P]),
    "rsl_last_error": (C.c_char_p, []),
    "rsl_order": (C.c_int, [C.c_int32, C.c_int64, _I64P, _I64P, C.c_uint64, _I32P, _DP]),
    "rsl_rng_uniform": (None, [C.c_uint64, C.c_double, C.c_double, C.c_int64, _DP]),
    "rsl_profile_enable": (C.c_int, [_P, C.c_int32]),
    "rsl_analyze_batch": (C.c_int, [C.c_int32, _I64P, C.POINTER(_I64P), C.POINTER(_I64P), C.POINTER(Settings),
                                    C.POINTER(_P)]),
    "rsl_factor_inertia_batch": (C.c_int, [_P, C.c_int32, _I64P]),
    "rsl_gmres_block": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_double, C.c_int32, C.c_int32, C.c_int32, _DP,
                                  _DP, _DP]),
    "rsl_cocg": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_double, C.c_int32, C.c_int32, _DP, _DP]),
    "rsl_cocr": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_double, C.c_int32, C.c_int32, _DP, _DP]),
    "rsl_factor_save": (C.c_int, [_P, C.c_char_p]),
    "rsl_factor_load": (C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    "rsl_profile_read": (C.c_int, [_P, C.c_int32, _DP]),
    "rsl_profile_kind_name": (C.c_char_p, [C.c_int32]),
    "rsl_lu_analyze": (C.c_int, [C.c_int64, _I64P, _I64P, C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_lu_symbolic_info": (C.c_int, [_P, _I64P, _DP]),
    "rsl_lu_symbolic_export": (C.c_int, [_P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P, _I32P, _I64P, _I32P]),
    "rsl_lu_factor": (C.c_int, [_P, C.c_int32, _DP, C.POINTER(Settings), C.POINTER(_P)]),
    "rsl_lu_refactor": (C.c_int, [_P, _DP]),
    "rsl_lu_solve": (C.c_int, [_P, _DP, C.c_int64, _DP]),
    "rsl_lu_solve_device": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "rsl_lu_refine": (C.c_int, [_P, _DP, _DP, C.c_int64, C.c_int32, _DP, _DP]),
    "rsl_lu_info": (C.c_int, [_P, _I64P, _DP]),
    "rsl_lu_factor_download": (C.c_int, [_P, _I64P, _I32P, _DP, _I64P, _I32P, _DP, _I32P, _I32P, _DP, _DP]),
    "rsl_lu_free_factor": (None, [_P]),
    "rsl_lu_free_symbolic": (None, [_P]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C rslab_amd/csrc"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
