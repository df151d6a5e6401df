"""ctypes binding of libdompap_hip.so (the C-ABI of include/dompap_hip.h).

The product path has no CPU fallback: if the library is missing, importing a
module that needs it raises ImportError with the build instruction.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdompap_hip.so")

DP_OK = 0
DP_ERR_ARG = -1
DP_ERR_HIP = -2
DP_ERR_RTC = -3
DP_ERR_OVERFLOW = -4
DP_ERR_STATE = -5
DP_ERR_NONFINITE = -6

DP_FP64 = 0
DP_MIXED = 1

FIELDS = {
    "positions": (0, "f8", True),
    "velocities": (1, "f8", True),
    "betas": (2, "f8", True),
    "image_positions": (3, "i4", True),
    "masses": (4, "f8", False),
    "types": (5, "i4", False),
    "positions_neighbour_list": (6, "f8", True),
}

DP_INTEGRATOR_GJF = 1
DP_INTEGRATOR_LEAPFROG = 2
DP_ORDER_X_FASTEST = 0
DP_ORDER_LAST_FASTEST = 1

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_p_i64 = ctypes.POINTER(ctypes.c_int64)
_p_i32 = ctypes.POINTER(ctypes.c_int32)
_p_dbl = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol of include/dompap_hip.h
SIGNATURES = {
    "dp_version": (_c_int, []),
    "dp_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_i64, _c_int, _c_int]),
    "dp_destroy": (_c_int, [_vp]),
    "dp_last_error": (ctypes.c_char_p, [_vp]),
    "dp_set_replicas": (_c_int, [_vp, _vp, _vp]),
    "dp_set_integrator": (_c_int, [_vp, _vp, _vp, _vp]),
    "dp_set_neighbor_params": (_c_int, [_vp, _vp, _vp, _c_int, _c_int]),
    "dp_load_potential": (_c_int, [_vp, ctypes.c_char_p, _c_int, _c_dbl, _c_dbl, _c_int, _vp, _vp, _c_int]),
    "dp_load_potentials": (_c_int, [_vp, _c_int, ctypes.POINTER(ctypes.c_char_p), _vp, _vp, _vp, _c_int, _vp, _vp,
                                    _c_int, _vp]),
    "dp_set_field": (_c_int, [_vp, _c_int, _vp]),
    "dp_get_field": (_c_int, [_vp, _c_int, _vp]),
    "dp_build_neighbor_list": (_c_int, [_vp]),
    "dp_update_neighbor_list": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "dp_compute": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "dp_step": (_c_int, [_vp, _c_i64, _c_int, _vp]),
    "dp_integrate": (_c_int, [_vp, _vp, _c_int, _vp, _c_int]),
    "dp_set_seed": (_c_int, [_vp, _c_u64, _c_u64]),
    "dp_set_replica_seeds": (_c_int, [_vp, _vp]),
    "dp_set_tuning": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "dp_set_cell_scale": (_c_int, [_vp, _c_dbl]),
    "dp_get_plan": (_c_int, [_vp, _vp, _c_int]),
    "dp_get_neighbor_list": (_c_int, [_vp, _vp]),
    "dp_get_neighbor_counts": (_c_int, [_vp, _vp]),
    "dp_get_stats": (_c_int, [_vp, _p_i64, _p_i64, _p_i32]),
    "dp_set_graph_block": (_c_int, [_vp, _c_int]),
    "dp_set_block_rebuild": (_c_int, [_vp, _c_int]),
    "dp_prepare_step_graphs": (_c_int, [_vp, _c_int, _c_int]),
    "dp_set_kernel_timing": (_c_int, [_vp, _c_int]),
    "dp_get_kernel_time": (_c_int, [_vp, _p_dbl, _p_i64]),
    "dp_get_kernel_breakdown": (_c_int, [_vp, _p_dbl, _p_dbl]),
    "dp_rdf_counts": (_c_int, [_vp, _vp, _c_int, _vp]),
    "dp_rdf_counts_noalloc": (_c_int, [_vp, _vp, _c_int, _vp]),
    "dp_get_rdf_path": (_c_int, [_vp, _p_i32]),
    "dp_compile_check": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.c_char_p, _c_int, ctypes.c_char_p, _c_i64, _p_i64]),
    "dp_compile_check_tuned": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.c_char_p, _c_int, _c_int, _c_int,
                                        _c_int, ctypes.c_char_p, _c_i64, _p_i64]),
    "dp_compile_check_potentials": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.POINTER(ctypes.c_char_p), _vp,
                                             _c_int, ctypes.c_char_p, _c_i64, _p_i64]),
    "dp_compile_code": (_c_int, [_c_int, _c_int, _c_int, _c_int, ctypes.c_char_p, _c_int, ctypes.c_char_p, _c_i64, _p_i64]),
    "dp_sync": (_c_int, [_vp]),
    "dp_get_stream": (_c_int, [_vp, ctypes.POINTER(_vp)]),
}

_LIB = None


def load() -> ctypes.CDLL:
    """Load libdompap_hip.so (once) and declare every C-ABI signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python dompap_amd/csrc/build.py` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class DeviceError(RuntimeError):
    """HIP / hiprtc failure inside libdompap_hip.so."""


def check(rc: int, ctx=None) -> None:
    """Map a C-ABI status to the reference's exception types (ValueError for
    bad arguments and neighbour-list overflow, as Simulation.py raises)."""
    if rc == DP_OK:
        return
    lib = load()
    msg = lib.dp_last_error(ctx)
    msg = msg.decode() if msg else f"error {rc}"
    if rc in (DP_ERR_ARG, DP_ERR_OVERFLOW, DP_ERR_NONFINITE):
        raise ValueError(msg)
    raise DeviceError(f"[{rc}] {msg}")
