"""ctypes binding of libmhdct.so (declarations in include/mhdct.h).

The library is loaded from this package directory.  There is no CPU fallback: if the
library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libmhdct.so"
_lib = None


class MhdctError(RuntimeError):
    pass


class Settings(C.Structure):
    """mhdct_settings (include/mhdct.h): the [hydro]/[mhd] keys the time step reads."""

    _fields_ = [
        ("gamma0", C.c_double),
        ("cfl", C.c_double),
        ("slope_type", C.c_double),
        ("smallr", C.c_double),
        ("smallc", C.c_double),
        ("scaling_factor", C.c_double),
        ("riemann", C.c_int),
        ("prolongation", C.c_int),
        ("abort_when_negative_eint", C.c_int),
        ("time_integrator", C.c_int),
        ("gravity_enabled", C.c_int),
        ("gravity", C.c_double * 3),
        ("viscosity_enabled", C.c_int),
        ("viscosity_mu", C.c_double),
        ("viscosity_hancock_predictor", C.c_int),
    ]


class AmrParams(C.Structure):
    """mhdct_amr_params (include/mhdct.h): the [amr] keys of the refinement criterion."""

    _fields_ = [
        ("level_min", C.c_int),
        ("level_max", C.c_int),
        ("epsilon_refine", C.c_double),
        ("epsilon_coarsen", C.c_double),
        ("epsilon_lohner", C.c_double),
        ("ivar", C.c_int),
        ("criterion", C.c_int),
        ("nvars", C.c_int),
        ("vars", C.c_int * 4),
    ]


# exported symbols and their signatures (also checked by tests/test_capi_and_mesh.py)
SIGNATURES = {
    "mhdct_mesh_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "mhdct_mesh_set_bc": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mhdct_mesh_adapt": (C.c_int, [C.c_void_p, C.POINTER(C.c_int8), C.c_int]),
    "mhdct_mesh_balance": (C.c_int, [C.c_void_p]),
    "mhdct_mesh_set_leaves": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhdct_mesh_finalize": (C.c_int, [C.c_void_p]),
    "mhdct_mesh_num_leaves": (C.c_int, [C.c_void_p]),
    "mhdct_mesh_num_outside": (C.c_int, [C.c_void_p]),
    "mhdct_mesh_leaves": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhdct_mesh_outside": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mhdct_mesh_last_error": (C.c_char_p, [C.c_void_p]),
    "mhdct_mesh_destroy": (None, [C.c_void_p]),
    "mhdct_create": (C.c_int, [C.POINTER(Settings), C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_void_p)]),
    "mhdct_set_mesh": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mhdct_mem_bytes": (C.c_uint64, [C.c_void_p]),
    "mhdct_do_time_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_void_p]),
    "mhdct_compute_dt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "mhdct_fill_outside": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mhdct_linear_combination": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                           C.c_double, C.c_void_p]),
    "mhdct_last_error": (C.c_char_p, [C.c_void_p]),
    "mhdct_destroy": (None, [C.c_void_p]),
    "mhdct_batch_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "mhdct_batch_add": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.POINTER(C.c_int)]),
    "mhdct_batch_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "mhdct_batch_step": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "mhdct_batch_step_timed": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_double)]),
    "mhdct_stage_name": (C.c_char_p, [C.c_int]),
    "mhdct_batch_slot_state": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int64)]),
    "mhdct_batch_check": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mhdct_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mhdct_batch_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "mhdct_batch_set_tune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "mhdct_batch_set_front": (C.c_int, [C.c_void_p, C.c_int]),
    "mhdct_batch_set_stage_threads": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mhdct_batch_set_class_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "mhdct_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "mhdct_refine_flags": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(AmrParams),
                                     C.POINTER(C.c_int8), C.POINTER(C.c_double), C.c_void_p]),
    "mhdct_refine_flags_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(AmrParams),
                                             C.POINTER(C.c_int8), C.POINTER(C.c_double), C.c_void_p]),
    "mhdct_mesh_clone": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mhdct_remap_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "mhdct_remap_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mhdct_remap_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mhdct_remap_last_error": (C.c_char_p, [C.c_void_p]),
    "mhdct_remap_destroy": (None, [C.c_void_p]),
    "mhdct_config_parse": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mhdct_config_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mhdct_config_get": (C.c_char_p, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "mhdct_config_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "mhdct_config_amr_params": (C.c_int, [C.c_void_p, C.POINTER(AmrParams)]),
    "mhdct_config_mesh": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mhdct_config_last_error": (C.c_char_p, [C.c_void_p]),
    "mhdct_config_destroy": (None, [C.c_void_p]),
    "mhdct_batch_num_slots": (C.c_int, [C.c_void_p]),
    "mhdct_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "mhdct_batch_destroy": (None, [C.c_void_p]),
}


def lib() -> C.CDLL:
    """Load libmhdct.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise MhdctError(
            f"{_LIB_PATH} is missing: build it with `python -m godunov_mhd_ct_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(str(_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def lib_path() -> Path:
    return _LIB_PATH
