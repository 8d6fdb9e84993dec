"""Loader for libmixsolver_hip.so (the HIP/CDNA4 product library).

There is no CPU fallback: if the library is missing or cannot be loaded the
import raises, naming what to build.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmixsolver_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mixsolver.h")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_longlong)
_V = C.c_void_p
_lib = None


class MixSolverError(RuntimeError):
    pass


def header_symbols():
    """Entry points declared in include/mixsolver.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ms_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MixSolverError(f"{LIB_PATH} not built: run python -m mixsolver_amd.build")
    L = C.CDLL(LIB_PATH)
    sig = {
        "ms_init": (C.c_int, [C.c_int]),
        "ms_last_error": (C.c_char_p, []),
        "ms_version": (C.c_int, []),
        "ms_system_create": (C.c_int, [C.POINTER(_V), C.c_int, _D, C.c_int, _L, _I, _I, _I, _I, C.c_int, _V, _V]),
        "ms_system_destroy": (C.c_int, [_V]),
        "ms_system_info": (C.c_int, [_V, _L]),
        "ms_set_field": (C.c_int, [_V, C.c_char_p, _D, C.c_longlong]),
        "ms_get_field": (C.c_int, [_V, C.c_char_p, _D, C.c_longlong]),
        "ms_set_ts_order": (C.c_int, [_V, C.c_int]),
        "ms_set_cell_types": (C.c_int, [_V, _I]),
        "ms_solid_temperature_assembly": (C.c_int, [_V, _D, _D, _D]),
        "ms_get_topology": (C.c_int, [_V, _I, _I, _I]),
        "ms_simple_run": (C.c_int, [_V, C.c_int, _I, _I]),
        "ms_piso_run": (C.c_int, [_V, C.c_int, _I]),
        "ms_pimple_run": (C.c_int, [_V, C.c_int, _I]),
        "ms_piso_continue": (C.c_int, [_V, C.c_int, _I]),
        "ms_pimple_continue": (C.c_int, [_V, C.c_int, _I]),
        "ms_get_history": (C.c_int, [_V, _D, C.c_int, _I]),
        "ms_solver_stats": (C.c_int, [_V, _L, C.c_int]),
        "ms_set_stall_window": (C.c_int, [C.c_int]),
        "ms_spmv_probe": (C.c_int, [C.c_int, _L, _D, _L]),
        "ms_compute_forces": (C.c_int, [_V, _D]),
        "ms_set_graphs": (C.c_int, [C.c_int]),
        "ms_set_fused_momentum": (C.c_int, [C.c_int]),
        "ms_set_schur_matfree": (C.c_int, [C.c_int]),
        "ms_set_fused_p_update": (C.c_int, [C.c_int]),
        "ms_get_ts_order": (C.c_int, [_V, _I]),
        "ms_nan_scan": (C.c_int, [_V, _I, _I, _D, _I, _I]),
        "ms_set_interface": (C.c_int, [_V, C.c_char_p, _D, C.c_longlong]),
        "ms_set_interface_modes": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ms_dump_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, _D, _D, _D, _D, _D, _D, _D]),
        "ms_dump_read": (C.c_int, [C.c_char_p, C.c_int, _D, _D, _D, _D, _D, _D, _D, _I, C.c_char_p, C.c_int, _I]),
        "ms_vtu_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _D, C.c_int, C.POINTER(C.c_longlong), _I, _I, _I,
                                   C.c_int, _D, _D, _D, _D, _D, _D]),
        "ms_cp_distribution": (C.c_int, [_V, _D, C.c_int, _I]),
        "ms_separation": (C.c_int, [_V, _D]),
        "ms_compute_gradients": (C.c_int, [_V]),
        "ms_momentum_assembly": (C.c_int, [_V, C.c_int, _D, _D, _D]),
        "ms_temperature_assembly": (C.c_int, [_V, _D, _D, _D]),
        "ms_flux_rhiechow": (C.c_int, [_V]),
        "ms_outflow_mass_rescale": (C.c_int, [_V]),
        "ms_ppe_assembly": (C.c_int, [_V, _D, _D]),
        "ms_correct_fields": (C.c_int, [_V, _D, _D]),
        "ms_cg_jacobi_noalloc": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_int, _I, _D, _I]),
        "ms_cg_jacobi": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_int, _I, _D, _I]),
        "ms_cg_ic0": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_int, _I, _D, _I]),
        "ms_bicgstab_ilu0": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_int, _I, _D, _I]),
        "ms_solve": (C.c_int, [C.c_int, _I, _I, _D, _D, _D, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, _I, _D, _I]),
        "ms_time_kernel": (C.c_int, [_V, C.c_int, C.c_int, _D, _D]),
        "ms_time_copy": (C.c_int, [C.c_longlong, C.c_int, _D]),
        "ms_time_stream": (C.c_int, [C.c_longlong, C.c_int, C.c_int, _D]),
        "ms_coarse_info": (C.c_int, [_V, _I, _I, _I]),
        "ms_coarse_aggregates": (C.c_int, [C.c_int, _D, C.c_int, _I]),
        "ms_cell_ordering": (C.c_int, [C.c_int, _I, _I, _D, C.c_int, C.c_int, C.c_int, _I, _I, _I]),
        "ms_periodic_match": (C.c_int, [C.c_int, _D, C.c_int, _L, _I, _I, _I, _I, C.c_int, _V, _I]),
        "ms_server_create": (C.c_int, [C.POINTER(_V), _V]),
        "ms_server_destroy": (C.c_int, [_V]),
        "ms_server_add_pattern": (C.c_int, [_V, C.c_int, _I, _I, _D, _I]),
        "ms_server_submit": (C.c_int, [_V, C.c_int, _D, _D, C.c_double, C.c_int, _L]),
        "ms_server_result": (C.c_int, [_V, C.c_longlong, C.c_double, _D, _I, _D, _I, _D, _D, _I]),
        "ms_server_stats": (C.c_int, [_V, _L]),
        "ms_sweep_create": (C.c_int, [C.POINTER(_V), _V, C.c_longlong, C.c_int]),
        "ms_sweep_destroy": (C.c_int, [_V]),
        "ms_sweep_submit": (C.c_int, [_V, C.c_int, _D, C.c_double, C.c_int, C.c_int, _L]),
        "ms_sweep_result": (C.c_int, [_V, C.c_longlong, C.c_double, _I, _I, _I, _D, _D, _D, _L, _D, _D, _D, _D]),
        "ms_sweep_stats": (C.c_int, [_V, _L]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().ms_last_error()
        raise MixSolverError(f"libmixsolver_hip error {rc}: {msg.decode() if msg else ''}")
    return rc


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)
