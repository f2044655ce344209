"""ctypes view of the C ABI declared in include/kpolrad.h.

The structures below mirror the C layouts field for field.  Loading fails
loudly when the built library is missing: there is no Python or CPU fallback
for any engine entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libkpolrad.so")

KP_OK, KP_ERR_ARG, KP_ERR_HIP, KP_ERR_ALLOC, KP_ERR_STATE, KP_ERR_UNSUPPORTED = range(6)
KP_THOMSON, KP_HG_RAYLEIGH = 0, 1
KP_F64, KP_F32 = 0, 1
KP_BAND_CELL_ALBEDO, KP_BAND_CELL_G = 1, 2

(KP_C_CROSS_MAIN, KP_C_CROSS_PEEL, KP_C_EVENTS, KP_C_ESCAPED, KP_C_KILLED_RR,
 KP_C_KILLED_STEPS, KP_C_LOST, KP_C_PACKETS) = range(8)
KP_NCOUNTERS = 8
COUNTER_NAMES = ["cross_main", "cross_peel", "events", "escaped", "killed_rr",
                 "killed_steps", "lost", "packets"]

KP_ES_X, KP_ES_D, KP_ES_I, KP_ES_Q, KP_ES_U = 0, 3, 6, 7, 8
KP_ES_EVENTS, KP_ES_CROSS, KP_ES_STATUS, KP_ES_IB, KP_ES_CELL, KP_ES_STRIDE = 9, 10, 11, 12, 13, 16
KP_END_ESCAPED, KP_END_RR, KP_END_STEPS, KP_END_LOST = 1, 2, 3, 4
KP_COORD_CARTESIAN, KP_COORD_SPHERICAL = 0, 1


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_blocks", C.c_int32),
        ("n_cell", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("x_min", C.c_double * 3),
        ("x_max", C.c_double * 3),
        ("xf0", C.POINTER(C.c_double)),
        ("dx0", C.POINTER(C.c_double)),
        ("vol0", C.POINTER(C.c_double)),
        ("level", C.POINTER(C.c_int32)),
        ("logic", C.POINTER(C.c_int32)),
        ("nb_id", C.POINTER(C.c_int32)),
        ("nb_shift", C.POINTER(C.c_uint8)),
        ("d_lvl", C.POINTER(C.c_int8)),
        ("phybnd", C.POINTER(C.c_uint8)),
        ("coord", C.c_int32),
        ("pad_", C.c_int32),
        ("xf", C.POINTER(C.c_double)),
        ("sph", C.POINTER(C.c_double)),
    ]


def sph_tab_len(n_cell: int) -> int:
    """KP_SPH_TAB: doubles per block of a spherical mesh's geometry table."""
    return 4 * (n_cell + 1) + 3 * n_cell


class Band(C.Structure):
    _fields_ = [
        ("kappa_scale", C.c_double),
        ("albedo", C.c_double),
        ("g", C.c_double),
        ("n_packets", C.c_uint64),
        ("flags", C.c_uint32),
        ("pad_", C.c_uint32),
    ]


class Source(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("weight", C.c_double)]


class Run(C.Structure):
    _fields_ = [
        ("seed", C.c_uint32),
        ("scatter_model", C.c_int32),
        ("p_max", C.c_double),
        ("n_bands", C.c_int32),
        ("precision", C.c_int32),
        ("bands", C.POINTER(Band)),
        ("packet_offset", C.c_uint64),
        ("step_max", C.c_int32),
        ("n_incl", C.c_int32),
        ("incl_deg", C.POINTER(C.c_double)),
        ("rr_weight", C.c_double),
        ("rr_survive", C.c_double),
        ("image_n", C.c_int32),
        ("pad_", C.c_int32),
        ("image_half", C.c_double),
    ]


KP_SCHED_HISTORY, KP_SCHED_PERSISTENT, KP_SCHED_BLOCKQ = 0, 1, 2


class Schedule(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("block_threads", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("chunk", C.c_int32),
        ("j_replicas", C.c_int32),
        ("img_replicas", C.c_int32),
        ("event_batch", C.c_int32),
        ("waves_per_simd", C.c_int32),
        ("coherent", C.c_int32),
        ("peel_queue", C.c_int32),
        ("vertex_slots", C.c_int32),
        ("adjoint_chunk", C.c_int32),
        ("queue_from_band", C.c_int32),
        ("queue_passes", C.c_int32),
        ("queue_chunk", C.c_int32),
        ("queue_yield", C.c_int32),
    ]


class Tallies(C.Structure):
    _fields_ = [
        ("J", C.c_void_p),
        ("image", C.c_void_p),
        ("counters", C.c_void_p),
        ("end_state", C.c_void_p),
    ]


class Grads(C.Structure):
    _fields_ = [
        ("d_logrho", C.c_void_p),
        ("d_albedo", C.c_void_p),
        ("d_g", C.c_void_p),
        ("jlin", C.c_void_p),
    ]


class ModuleInfo(C.Structure):
    _fields_ = [
        ("n_blocks", C.c_int64),
        ("n_cells", C.c_int64),
        ("n_cell", C.c_int32),
        ("n_bands", C.c_int32),
        ("n_incl", C.c_int32),
        ("image_n", C.c_int32),
        ("n_par", C.c_int64),
        ("n_cycle", C.c_int32),
        ("initialised", C.c_int32),
        ("packets_done", C.c_int64),
    ]


# Every entry point declared in include/*.h, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "kp_last_error": (C.c_char_p, []),
    "kp_mesh_create": (C.c_int, [C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(_P)]),
    "kp_mesh_destroy": (None, [_P]),
    "kp_mesh_set_coord": (C.c_int, [_P, C.c_int32, C.c_double]),
    "kp_mesh_refine": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.c_int64]),
    "kp_mesh_refine_region": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
    "kp_mesh_finalize": (C.c_int, [_P]),
    "kp_mesh_describe": (C.c_int, [_P, C.POINTER(MeshDesc)]),
    "kp_mesh_locate": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
    "kp_engine_create": (C.c_int, [C.POINTER(MeshDesc), C.c_int32, C.POINTER(_P)]),
    "kp_engine_destroy": (None, [_P]),
    "kp_set_media": (C.c_int, [_P, _P, _P, _P]),
    "kp_forward": (C.c_int, [_P, C.POINTER(Source), C.POINTER(Run), C.POINTER(Tallies), _P]),
    "kp_adjoint": (C.c_int, [_P, C.POINTER(Source), C.POINTER(Run), _P, C.POINTER(Grads), _P]),
    "kp_set_schedule": (C.c_int, [_P, C.POINTER(Schedule)]),
    "kp_autotune": (C.c_int, [_P, C.POINTER(Source), C.POINTER(Run), C.c_uint64, C.c_int32,
                              C.POINTER(Schedule), _P]),
    "kp_graph_create": (C.c_int, [_P, C.POINTER(Source), C.POINTER(Run), C.POINTER(Tallies),
                                  C.c_int32, C.POINTER(_P)]),
    "kp_graph_launch": (C.c_int, [_P, _P]),
    "kp_compact_workspace_bytes": (C.c_size_t, [C.c_int64, C.c_int32]),
    "kp_module_create": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "kp_module_destroy": (None, [_P]),
    "kp_module_init": (C.c_int, [_P, C.c_int32]),
    "kp_module_set_media": (C.c_int, [_P, _P, _P, _P]),
    "kp_module_step": (C.c_int, [_P, _P]),
    "kp_module_save": (C.c_int, [_P, C.c_char_p]),
    "kp_module_get_info": (C.c_int, [_P, C.POINTER(ModuleInfo)]),
    "kp_module_tallies": (C.c_int, [_P, _P, _P, _P]),
    "kp_module_mesh": (C.c_int, [_P, C.POINTER(_P)]),
    "kp_pool_keep": (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, _P, C.c_size_t, _P]),
    "kp_bin_by_key": (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, _P, _P, C.c_size_t, _P]),
    "kp_bin_by_key_static": (C.c_int, [_P, C.c_int64, C.c_int32, _P, _P, _P, _P]),
    "kp_gather": (C.c_int, [_P, C.c_int32, _P, _P, C.c_int64, _P, _P]),
    "kp_graph_destroy": (None, [_P]),
    "kp_get_schedule": (C.c_int, [_P, C.POINTER(Schedule)]),
}


class KPError(RuntimeError):
    """Non-zero kp_status; mirrors the reference's std::runtime_error."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"kp_status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    """The loaded engine library (raises if it was never built)."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime under the same SONAME
        # (libamdhip64.so.7); importing it first makes the dynamic loader bind
        # this library to that copy, so the process holds one HIP runtime and
        # torch-allocated device pointers are valid in our kernels.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m kratos_polrad_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != KP_OK:
        raise KPError(status, lib().kp_last_error().decode())
