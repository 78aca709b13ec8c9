"""ctypes bindings of libsubsweep_amd.so (the C ABI in include/subsweep_amd.h)
and libssw_grid.so. Loading fails loudly when a library has not been built:
there is no CPU fallback for the sweep."""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SWEEP_LIB_PATH = os.path.join(HERE, "libsubsweep_amd.so")
GRID_LIB_PATH = os.path.join(HERE, "libssw_grid.so")

SSW_OK = 0
SSW_ERR_INVALID = -1
SSW_ERR_OOM = -2
SSW_ERR_BUDGET = -3
SSW_ERR_CYCLE = -4
SSW_ERR_HIP = -5
SSW_ERR_GRID = -6
SSW_ERR_NO_DEVICE = -7

FACE_LOCAL = 0
FACE_PERIODIC = 1
FACE_BOUNDARY = 2

TRANSPORT_FP64 = 64
TRANSPORT_FP32 = 32


class SweepParametersC(C.Structure):
    _fields_ = [
        ("num_directions", C.c_int32),
        ("explicit_directions", C.POINTER(C.c_double)),
        ("num_timestep_levels", C.c_int32),
        ("periodic", C.c_int32),
        ("max_timestep", C.c_double),
        ("significant_rate_threshold", C.c_double),
        ("timestep_safety_factor", C.c_double),
        ("chemistry_timestep_safety_factor", C.c_double),
        ("prevent_cooling", C.c_int32),
        ("transport_precision", C.c_int32),
        ("scale_factor", C.c_double),
    ]


TRANSPORT_MODE_AUTO = 0
TRANSPORT_MODE_GEOMETRIC = 3
TRANSPORT_MODE_COMPILED = 4


class InfoC(C.Structure):
    _fields_ = [
        ("num_cells", C.c_int64),
        ("num_faces", C.c_int64),
        ("num_directions", C.c_int32),
        ("max_depth", C.c_int32),
        ("updates_last_step", C.c_int64),
        ("sweeps_last_step", C.c_int64),
        ("downwind_local", C.c_double),
        ("downwind_periodic", C.c_double),
        ("downwind_boundary", C.c_double),
        ("device_bytes", C.c_int64),
        ("chemistry_failures", C.c_int64),
        ("transport_mode", C.c_int32),
        ("list_builds_last_step", C.c_int32),
        ("overflow_entries", C.c_int64),
        ("graph_capture_failures", C.c_int64),
        ("replay_block", C.c_int32),
        ("graph_replay", C.c_int32),
        ("locality_order", C.c_int32),
        ("narrow_runs", C.c_int32),
        ("peak_device_bytes", C.c_int64),
        ("group_levels", C.c_int32),
        ("group_fallback", C.c_int32),
        ("group_tasks", C.c_int64),
        ("group_block", C.c_int32),
        ("group_blocks_per_cu", C.c_int32),
        ("limiting_temperature", C.c_int64),
        ("limiting_ionization_fraction", C.c_int64),
        ("limiting_photon_rate", C.c_int64),
        ("group_launches", C.c_int64),
        ("group_workgroups", C.c_int64),
        ("compute_units", C.c_int32),
        ("direction_sweeps", C.c_int32),
    ]


class SweepError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"subsweep_amd error {code}: {message}")
        self.code = code


class RequestC(C.Structure):
    _fields_ = [
        ("params", SweepParametersC),
        ("transport_mode", C.c_int32),
        ("steps", C.c_int32),
        ("replay_block", C.c_int32),
        ("direction_sweeps", C.c_int32),
        ("num_cells", C.c_int64),
        ("face_offsets", C.c_void_p),
        ("face_neighbour", C.c_void_p),
        ("face_kind", C.c_void_p),
        ("face_area", C.c_void_p),
        ("face_normal", C.c_void_p),
        ("cell_volume", C.c_void_p),
        ("cell_size", C.c_void_p),
        ("density", C.c_void_p),
        ("ionized_fraction", C.c_void_p),
        ("temperature", C.c_void_p),
        ("source", C.c_void_p),
        ("cell_position", C.c_void_p),
        ("out_ionized_fraction", C.c_void_p),
        ("out_temperature", C.c_void_p),
        ("out_timestep", C.c_void_p),
        ("out_photon_rate", C.c_void_p),
        ("out_level", C.c_void_p),
        ("out_change_timescale", C.c_void_p),
        ("grid_key", C.c_int64),
    ]


class RequestResultC(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("transport_mode", C.c_int32),
        ("submit_s", C.c_double),
        ("start_s", C.c_double),
        ("end_s", C.c_double),
        ("updates", C.c_int64),
        ("device_bytes", C.c_int64),
        ("estimated_bytes", C.c_int64),
        ("error", C.c_char * 256),
        ("reused", C.c_int32),
        ("reserved1", C.c_int32),
    ]


P = C.c_void_p
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)
PU8 = C.POINTER(C.c_uint8)
PU32 = C.POINTER(C.c_uint32)

# Every symbol the header declares, with its signature.
SWEEP_SIGNATURES = {
    "ssw_default_parameters": (None, [C.POINTER(SweepParametersC)]),
    "ssw_create": (C.c_int, [C.POINTER(SweepParametersC), C.POINTER(P)]),
    "ssw_upload_grid": (C.c_int, [P, C.c_int64, PI64, PI32, PU8, PD, PD, PD, PD]),
    "ssw_upload_state": (C.c_int, [P, PD, PD, PD, PD]),
    "ssw_step": (C.c_int, [P, PD]),
    "ssw_assign_sources_device": (C.c_int, [C.c_int64, PD, C.c_int64, PD, PD, PD, PI64]),
    "ssw_step_batch": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, PD]),
    "ssw_packed_graph_stats": (C.c_int, [PI64, PI64]),
    "ssw_download": (C.c_int, [P, PD, PD, PD, PD, PU8, PD]),
    "ssw_download_noalloc": (C.c_int, [P, PD, PD, PD, PD, PU8, PD]),
    "ssw_rotate_directions": (C.c_int, [P, PD]),
    "ssw_rotation_matrix": (C.c_int, [PD, C.c_double, PD]),
    "ssw_random_rotation_matrix": (C.c_int, [C.c_double, C.c_double, C.c_double, PD]),
    "ssw_download_chemistry_outputs": (C.c_int, [P, PD, PD, PD, PD]),
    "ssw_download_counts": (C.c_int, [P, C.c_int32, PU32]),
    "ssw_download_levelsets": (C.c_int, [P, C.c_int32, PU32]),
    "ssw_download_fluxes": (C.c_int, [P, PD, PD, PD]),
    "ssw_get_info": (C.c_int, [P, C.POINTER(InfoC)]),
    "ssw_synchronize": (C.c_int, [P]),
    "ssw_set_transport_mode": (C.c_int, [P, C.c_int32]),
    "ssw_set_timing": (C.c_int, [P, C.c_int32]),
    "ssw_set_replay_block": (C.c_int, [P, C.c_int32]),
    "ssw_set_graph_replay": (C.c_int, [P, C.c_int32]),
    "ssw_set_narrow_runs": (C.c_int, [P, C.c_int32]),
    "ssw_set_group_levels": (C.c_int, [P, C.c_int32]),
    "ssw_set_group_block": (C.c_int, [P, C.c_int32]),
    "ssw_set_direction_sweeps": (C.c_int, [P, C.c_int32]),
    "ssw_direction_table": (C.c_int, [C.c_int32, C.POINTER(C.c_double)]),
    "ssw_checkpoint_bytes": (C.c_int64, [P]),
    "ssw_save_checkpoint": (C.c_int, [P, C.c_void_p, C.c_int64]),
    "ssw_load_checkpoint": (C.c_int, [P, C.c_void_p, C.c_int64]),
    "ssw_set_cell_positions": (C.c_int, [P, C.c_int64, PD]),
    "ssw_get_timing": (C.c_int, [P, PD, PD, PI64, PD, PD]),
    "ssw_download_transport": (C.c_int, [P, P, P]),
    "ssw_timestep_level": (C.c_int32, [C.c_int32, C.c_double, C.c_double]),
    "ssw_chemistry_timestep": (C.c_double, [PD, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int32, C.c_double, C.c_double, PI32, PI32]),
    "ssw_destroy": (None, [P]),
    "ssw_last_error": (C.c_char_p, []),
    "ssw_estimate_device_bytes": (C.c_int64, [C.POINTER(SweepParametersC), C.c_int64, C.c_int64, C.c_int64,
                                              C.c_int32]),
    "ssw_bpass_source_rate": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "ssw_time_between_scale_factors": (C.c_double, [C.c_double] * 5),
    "ssw_star_sources": (C.c_int64, [C.c_int64, PD, PD, PD] + [C.c_double] * 5 + [PD, PU8]),
    "ssw_server_create": (C.c_int, [C.c_int32, C.c_int64, C.POINTER(P)]),
    "ssw_server_create_packed": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.POINTER(P)]),
    "ssw_server_pack_stats": (C.c_int, [P, PI64, PI64]),
    "ssw_server_submit": (C.c_int, [P, C.POINTER(RequestC), PI64]),
    "ssw_server_wait": (C.c_int, [P, C.c_int64, C.POINTER(RequestResultC)]),
    "ssw_server_stats": (C.c_int, [P, PI64, PI32]),
    "ssw_server_destroy": (None, [P]),
}

GRID_SIGNATURES = {
    "ssw_points_jittered_lattice": (None, [C.c_int64, C.c_double, C.c_double, C.c_uint64, PD]),
    "ssw_points_uniform": (None, [C.c_int64, C.c_double, C.c_uint64, PD]),
    "ssw_lognormal_density": (None, [C.c_int64, C.c_double, C.c_double, C.c_uint64, PD]),
    "ssw_assign_sources": (C.c_int, [C.c_int64, PD, C.c_int64, PD, PD, PD, PI64]),
    "ssw_grid_cartesian": (None, [C.c_int64, C.c_double, C.c_int, PI64, PI32, PU8, PD, PD, PD, PD]),
    "ssw_voronoi_build": (P, [C.c_int64, PD, PD, C.c_int, PD, C.c_int]),
    "ssw_voronoi_num_faces": (C.c_int64, [P]),
    "ssw_voronoi_dropped_faces": (C.c_int64, [P]),
    "ssw_voronoi_open_cells": (C.c_int64, [P]),
    "ssw_voronoi_get": (None, [P, PI64, PI32, PU8, PD, PD, PD, PD]),
    "ssw_voronoi_free": (None, [P]),
    "ssw_grid_last_error": (C.c_char_p, []),
    "ssw_grid_set_progress": (None, [C.c_int]),
}

_sweep_lib = None
_grid_lib = None


def _load(path: str, sigs: dict):
    if not os.path.exists(path):
        raise ImportError(
            f"{os.path.basename(path)} is not built; run `python -m subsweep_amd.build` "
            f"(or __graft_entry__.build()) first"
        )
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def sweep_lib():
    global _sweep_lib
    if _sweep_lib is None:
        _sweep_lib = _load(SWEEP_LIB_PATH, SWEEP_SIGNATURES)
    return _sweep_lib


def grid_lib():
    global _grid_lib
    if _grid_lib is None:
        _grid_lib = _load(GRID_LIB_PATH, GRID_SIGNATURES)
    return _grid_lib


def check(code: int) -> None:
    if code != SSW_OK:
        msg = sweep_lib().ssw_last_error().decode(errors="replace")
        raise SweepError(code, msg)


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))
