"""ctypes binding of libccs.so (include/ccs.h).

The library is built in place by ``__graft_entry__.build()`` (or ``make -C
conformal_cubed_sphere_amd/csrc``).  Importing this module without the built
library raises ImportError: there is no CPU fallback for the time-step path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccs.so")

_D = ctypes.POINTER(ctypes.c_double)

# status codes (enum ccs_status)
CCS_OK = 0
CCS_ERR_INVALID = 1
CCS_ERR_DEVICE = 2
CCS_ERR_ORDER = 3
CCS_ERR_NONFINITE = 4
CCS_ERR_NOMEM = 5
CCS_ERR_SINGULAR = 6

# enum ccs_precision / ccs_scheme
CCS_FP64_PARITY = 0
CCS_FP32_SPLIT = 1
CCS_FP64_FAST = 2
CCS_RK3_TVD = 0
CCS_RK4 = 1
# enum ccs_spatial_op
CCS_OP_WENO = 0
CCS_OP_CENTRED2 = 1
CCS_OP_CENTRED4 = 2
CCS_OP_CENTRED6 = 3
CCS_OP_WENO_EXT = 4

STATUS_NAMES = {
    0: "CCS_OK", 1: "CCS_ERR_INVALID", 2: "CCS_ERR_DEVICE", 3: "CCS_ERR_ORDER",
    4: "CCS_ERR_NONFINITE", 5: "CCS_ERR_NOMEM", 6: "CCS_ERR_SINGULAR",
}


class CcsMetrics(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("halo", ctypes.c_int),
        ("radius", ctypes.c_double), ("omega", ctypes.c_double), ("dx", ctypes.c_double),
        ("lon", _D), ("lat", _D), ("sinlon", _D), ("coslon", _D), ("sinlat", _D), ("coslat", _D),
        ("sqrtG", _D), ("f", _D), ("areaCell", _D), ("G", _D), ("IG", _D), ("A", _D), ("IA", _D),
    ]


class CcsCaseParams(ctypes.Structure):
    _fields_ = [
        ("case_num", ctypes.c_int), ("n_bumps", ctypes.c_int), ("mountain", ctypes.c_int),
        ("pad_", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("gravity", ctypes.c_double),
        ("u0", ctypes.c_double), ("h_mean", ctypes.c_double), ("bump_amp", ctypes.c_double),
        ("bump_rmin", ctypes.c_double), ("bump_rmax", ctypes.c_double),
        ("mountain_h", ctypes.c_double), ("mountain_r", ctypes.c_double),
    ]


class CcsConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("halo", ctypes.c_int), ("precision", ctypes.c_int),
        ("scheme", ctypes.c_int), ("eflux", ctypes.c_int), ("device", ctypes.c_int),
        ("use_graph", ctypes.c_int), ("spatial_op", ctypes.c_int),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("radius", ctypes.c_double),
        ("gravity", ctypes.c_double), ("omega", ctypes.c_double),
    ]


class CcsEnsembleConfig(ctypes.Structure):
    _fields_ = [
        ("n_members", ctypes.c_int), ("halo", ctypes.c_int), ("precision", ctypes.c_int),
        ("scheme", ctypes.c_int), ("device", ctypes.c_int), ("use_graph", ctypes.c_int),
        ("tile_y", ctypes.c_int), ("persistent", ctypes.c_int),
        ("radius", ctypes.c_double), ("gravity", ctypes.c_double), ("omega", ctypes.c_double),
    ]


class CcsHistoryDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("halo", ctypes.c_int), ("case_num", ctypes.c_int), ("pad_", ctypes.c_int),
                ("dx_deg", ctypes.c_double), ("dt", ctypes.c_double)]


_I = ctypes.POINTER(ctypes.c_int)

# Every entry point declared in include/ccs.h, with its ctypes signature.
SIGNATURES = {
    "ccs_mesh_geometry": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, _D]),
    "ccs_mesh_metrics": (ctypes.c_int, [_D, _D, _D, ctypes.POINTER(CcsMetrics)]),
    "ccs_mesh_ext_metric": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _D, _D]),
    "ccs_case_defaults": (None, [ctypes.POINTER(CcsCaseParams)]),
    "ccs_init_case": (ctypes.c_int, [ctypes.POINTER(CcsMetrics), ctypes.POINTER(CcsCaseParams),
                                     _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_create": (ctypes.c_int, [ctypes.POINTER(CcsConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "ccs_set_mesh": (ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_set_mesh_ext": (ctypes.c_int, [ctypes.c_void_p, _D]),
    "ccs_set_state": (ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D, _D]),
    "ccs_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ccs_get_state": (ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_tendency": (ctypes.c_int, [ctypes.c_void_p, _D, _D, _D]),
    "ccs_diagnostics": (ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    "ccs_time_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D]),
    "ccs_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "ccs_destroy": (None, [ctypes.c_void_p]),
    "ccs_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ccs_version": (ctypes.c_char_p, []),
    "ccs_mesh_file_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "ccs_mesh_file_read": (ctypes.c_int, [ctypes.c_char_p, _I, _D, _D, _D]),
    "ccs_mesh_file_read_ext": (ctypes.c_int, [ctypes.c_char_p, _I, _D]),
    "ccs_history_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CcsHistoryDesc), _D, _D, _D, _D,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "ccs_history_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_history_close": (ctypes.c_int, [ctypes.c_void_p]),
    "ccs_diagnostics_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 11
                                + [ctypes.c_char_p, ctypes.c_size_t]),
    "ccs_ensemble_create": (ctypes.c_int, [ctypes.POINTER(CcsEnsembleConfig), _I, _D, _D,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "ccs_ensemble_set_mesh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_ensemble_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D]),
    "ccs_ensemble_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D]),
    "ccs_ensemble_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D]),
    "ccs_ensemble_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ccs_ensemble_step_members": (ctypes.c_int, [ctypes.c_void_p, _I]),
    "ccs_ensemble_run_until": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "ccs_ensemble_steps_done": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]),
    "ccs_ensemble_autotune_source": (ctypes.c_int, [ctypes.c_void_p, _I]),
    "ccs_ensemble_autotune": (ctypes.c_int, [ctypes.c_void_p, _I, _I, _D]),
    "ccs_ensemble_time_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    "ccs_ensemble_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "ccs_ensemble_destroy": (None, [ctypes.c_void_p]),
    "ccs_ensemble_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run __graft_entry__.build() or "
            "`make -C conformal_cubed_sphere_amd/csrc` (no CPU fallback exists)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_D)
