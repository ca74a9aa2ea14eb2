"""ctypes binding of libmxm_hip.so (include/mxm_dg.h).

The library is loaded from this package directory; when it is missing every
call raises ``MxmError`` -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmxm_hip.so")

MXM_OK = 0
PRM_MAX = 32

# mesh table ids (MXM_T_*)
T_ELEM_NODE_XYZ, T_ELEM_INVJ, T_ELEM_INV_VOL, T_ELEM_FACES, T_ELEM_FACE_SIDE, T_ELEM_CANON_TO_REF, \
    T_FACE_ELEM, T_FACE_ELEM_NODE, T_FACE_NORMAL, T_FACE_AREA, T_FACE_BC_TYPE = range(11)


class MxmError(RuntimeError):
    """A libmxm_hip call returned a non-zero code."""


class Mesh3DDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lx", C.c_double), ("ly", C.c_double),
                ("lz", C.c_double), ("bc", C.c_int * 6)]


class Mesh2DDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("lx", C.c_double), ("ly", C.c_double), ("bc", C.c_int * 4)]


class Throughput(C.Structure):
    _fields_ = [("steps", C.c_int), ("wall_seconds", C.c_double), ("dof", C.c_longlong),
                ("state_bytes_per_step", C.c_longlong), ("dof_per_second", C.c_double),
                ("dof_stage_updates_per_second", C.c_double), ("state_bandwidth_gbs", C.c_double),
                ("stage_seconds", C.c_double), ("achieved_hbm_gbs", C.c_double), ("hbm_fraction", C.c_double),
                ("mfma_tflops", C.c_double), ("mfma_utilization", C.c_double)]


class MemReport(C.Structure):
    _fields_ = [("rk_stage", C.c_longlong), ("dg_operators", C.c_longlong), ("limiter", C.c_longlong),
                ("mesh_connectivity", C.c_longlong), ("halo_device", C.c_longlong), ("halo_pinned", C.c_longlong)]


EXPORTS = {
    "mxm_hip_init": ([C.c_int], C.c_int),
    "mxm_last_error": ([], C.c_char_p),
    "mxm_get_constant": ([C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_refelem_3d": ([C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mxm_refelem_2d": ([C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mxm_refelem_deriv": ([C.c_int, C.c_int, C.c_void_p], C.c_int),
    "mxm_mesh_create_3d": ([C.POINTER(Mesh3DDesc), C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mxm_mesh_create_2d": ([C.POINTER(Mesh2DDesc), C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mxm_mesh_info": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "mxm_mesh_get": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "mxm_mesh_destroy": ([C.c_void_p], C.c_int),
    "mxm_dg_create_3d": ([C.POINTER(Mesh3DDesc), C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mxm_dg_create_2d": ([C.POINTER(Mesh2DDesc), C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mxm_dg_info": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "mxm_dg_upload_q": ([C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "mxm_dg_download_q": ([C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "mxm_dg_download_component": ([C.c_void_p, C.c_int, C.c_void_p, C.c_longlong], C.c_int),
    "mxm_dg_rk_stage": ([C.c_void_p, C.c_int, C.c_double], C.c_int),
    "mxm_dg_step_ssprk3": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "mxm_dg_prepare_steps": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "mxm_dg_enable_limiter": ([C.c_void_p, C.c_int, C.c_double], C.c_int),
    "mxm_dg_sync": ([C.c_void_p], C.c_int),
    "mxm_dg_bench_step_loop": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(Throughput)], C.c_int),
    "mxm_dg_dof_count": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "mxm_dg_state_bytes_per_step": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "mxm_dg_memory_report": ([C.c_void_p, C.POINTER(MemReport)], C.c_int),
    "mxm_dg_time_stage_kernel": ([C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "mxm_dg_autotune": ([C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int],
                        C.c_int),
    "mxm_dg_launch_config": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_dg_tile_config": ([C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "mxm_dg_diagnostics": ([C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_dg_max_wave_speed": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "mxm_dg_max_wave_speed_noalloc": ([C.c_void_p, C.c_void_p], C.c_int),
    "mxm_dg_destroy": ([C.c_void_p], C.c_int),
    "mxm_batch_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mxm_batch_autotune": ([C.c_void_p, C.c_int], C.c_int),
    "mxm_batch_graphs": ([C.c_void_p, C.c_int], C.c_int),
    "mxm_batch_headroom": ([C.c_void_p, C.c_double], C.c_int),
    "mxm_batch_diagnostics": ([C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_max_wave_speed": ([C.c_void_p, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "mxm_batch_add_2d": ([C.c_void_p, C.POINTER(Mesh2DDesc), C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int,
                          C.c_int, C.c_double, C.c_void_p, C.c_longlong, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_add_3d": ([C.c_void_p, C.POINTER(Mesh3DDesc), C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int,
                          C.c_int, C.c_double, C.c_void_p, C.c_longlong, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_footprint_2d": ([C.POINTER(Mesh2DDesc), C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_longlong)], C.c_int),
    "mxm_batch_footprint_3d": ([C.POINTER(Mesh3DDesc), C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_longlong)], C.c_int),
    "mxm_batch_step": ([C.c_void_p, C.c_int], C.c_int),
    "mxm_batch_poll": ([C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_poll_lagged": ([C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_status": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_failed": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "mxm_batch_get_state": ([C.c_void_p, C.c_int, C.c_void_p, C.c_longlong], C.c_int),
    "mxm_batch_remove": ([C.c_void_p, C.c_int], C.c_int),
    "mxm_batch_memory": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "mxm_batch_sync": ([C.c_void_p], C.c_int),
    "mxm_batch_destroy": ([C.c_void_p], C.c_int),
}

_lib = None


def lib():
    """Load libmxm_hip.so (raises MxmError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MxmError(f"{LIB_PATH} is missing: build it with mojoxm_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (argt, rest) in EXPORTS.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = rest
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != MXM_OK:
        msg = lib().mxm_last_error()
        raise MxmError(f"{what} rc={rc}: {msg.decode() if msg else ''}")


_initialised = False


def ensure_device(device: int = 0) -> None:
    """mxm_hip_init once per process."""
    global _initialised
    if not _initialised:
        check(lib().mxm_hip_init(device), "mxm_hip_init")
        _initialised = True
