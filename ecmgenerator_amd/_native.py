"""ctypes binding of libecmcrowd.so (include/ecm_sim.h, include/ecm_scene.h,
include/ecm_mlp.h).

The library is built in-tree by ``python -m ecmgenerator_amd.build``.  Loading
fails loudly when it is missing: there is no Python or CPU fallback for the
step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libecmcrowd.so")


class SimParams(C.Structure):
    """ecm_sim_params (include/ecm_sim.h)."""
    _fields_ = [
        ("max_agents", C.c_int),
        ("k", C.c_int),
        ("time_horizon", C.c_float),
        ("time_horizon_obst", C.c_float),
        ("max_speed", C.c_float),
        ("max_speed_from_pref", C.c_int),
        ("dt", C.c_float),
        ("mass", C.c_float),
        ("arrival_radius", C.c_float),
        ("delete_radius", C.c_float),
        ("grid_cell", C.c_float),
        ("keep_orca_velocity", C.c_int),
    ]


class SimTuning(C.Structure):
    """ecm_sim_tuning (include/ecm_sim.h)."""
    _fields_ = [("cell_scale", C.c_float), ("knn_block", C.c_int), ("orca_obst_slots", C.c_int),
                ("best_ms", C.c_float)]


REPLAN_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                        C.c_float, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int))

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_PF = C.POINTER(C.c_float)
_PI = C.POINTER(C.c_int)

# symbol -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # ecm_sim.h
    "ecm_sim_default_params": (None, [C.POINTER(SimParams)]),
    "ecm_sim_reference_params": (None, [C.POINTER(SimParams)]),
    "ecm_sim_create": (_I, [C.POINTER(SimParams), _F, _F, _F, _F, C.POINTER(_P)]),
    "ecm_sim_destroy": (None, [_P]),
    "ecm_sim_set_obstacles": (_I, [_P, _PF, _PI, _I]),
    "ecm_sim_set_ecm": (_I, [_P, _I, _PF, _I, _PI, _PF]),
    "ecm_sim_spawn": (_I, [_P, _I, _PF, _PF, _PF, _PI, _PI, _PF, _I, _PI]),
    "ecm_sim_destroy_agent": (_I, [_P, _I]),
    "ecm_sim_reset": (_I, [_P]),
    "ecm_sim_set_path": (_I, [_P, _I, _PF, _I]),
    "ecm_sim_set_replan_callback": (_I, [_P, REPLAN_FN, _P]),
    "ecm_sim_set_replan_threads": (_I, [_P, _I]),
    "ecm_sim_step": (_I, [_P, _I]),
    "ecm_sim_sync": (_I, [_P]),
    "ecm_sim_read": (_I, [_P, _I, _P, C.c_size_t]),
    "ecm_sim_write": (_I, [_P, _I, _P, C.c_size_t]),
    "ecm_sim_device_ptr": (_I, [_P, _I, C.POINTER(_P)]),
    "ecm_host_alloc": (_I, [C.c_size_t, C.POINTER(_P)]),
    "ecm_host_free": (_I, [_P]),
    "ecm_sim_num_agents": (_I, [_P, _PI]),
    "ecm_sim_last_index": (_I, [_P, _PI]),
    "ecm_sim_get_neighbors": (_I, [_P, _PI, _PI]),
    "ecm_sim_get_replans": (_I, [_P, _PI, _I, _PI]),
    "ecm_sim_read_paths": (_I, [_P, _PI, _PI, _PF, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ecm_sim_stats": (_I, [_P, C.POINTER(C.c_uint64), _I]),
    "ecm_sim_set_graph": (_I, [_P, _I]),
    "ecm_sim_graph_status": (_I, [_P, _PI, C.c_char_p, _I]),
    "ecm_sim_get_tuning": (_I, [_P, C.POINTER(SimTuning)]),
    "ecm_sim_set_tuning": (_I, [_P, C.POINTER(SimTuning)]),
    "ecm_sim_autotune": (_I, [_P, _I, C.POINTER(SimTuning)]),
    "ecm_sim_add_env": (_I, [_P, _F, _F, _F, _F, _PI]),
    "ecm_sim_num_envs": (_I, [_P, _PI]),
    "ecm_sim_env_set_obstacles": (_I, [_P, _I, _PF, _PI, _I]),
    "ecm_sim_env_set_ecm": (_I, [_P, _I, _I, _PF, _I, _PI, _PF]),
    "ecm_sim_env_spawn": (_I, [_P, _I, _I, _PF, _PF, _PF, _PI, _PI, _PF, _I, _PI]),
    "ecm_sim_env_set_replan_callback": (_I, [_P, _I, REPLAN_FN, _P]),
    "ecm_sim_agent_envs": (_I, [_P, _PI]),
    "ecm_sim_set_timing": (_I, [_P, _I]),
    "ecm_sim_last_timing": (_I, [_P, _PF, _I]),
    "ecm_sim_snapshot_save": (_I, [_P, C.c_char_p]),
    "ecm_sim_snapshot_load": (_I, [_P, C.c_char_p]),
    "ecm_last_error": (C.c_char_p, []),
    "ecm_device_available": (_I, []),
    # ecm_scene.h
    "ecm_scene_create": (_I, [_F, _F, _F, _F, C.POINTER(_P)]),
    "ecm_scene_destroy": (None, [_P]),
    "ecm_scene_add_obstacle": (_I, [_P, _PF, _I]),
    "ecm_scene_build_ecm": (_I, [_P, _F, _I]),
    "ecm_scene_ecm_counts": (_I, [_P, _PI, _PI]),
    "ecm_scene_ecm_export": (_I, [_P, _PF, _PF, _PI, _PI, _PF]),
    "ecm_scene_obstacle_counts": (_I, [_P, _PI, _PI]),
    "ecm_scene_obstacle_export": (_I, [_P, _PF, _PI]),
    "ecm_scene_locate": (_I, [_P, _F, _F]),
    "ecm_scene_plan_path": (_I, [_P, _F, _F, _F, _F, _F, _PF, _I, _PI]),
    "ecm_scene_plan_paths": (_I, [_P, _I, _PF, _PF, _PF, _I, _PI, _PF, C.c_longlong, _PI, _I]),
    "ecm_scene_inside_obstacle": (_I, [_P, _F, _F]),
    "ecm_scene_replan_callback": (REPLAN_FN, []),
    "ecm_scenario_generate": (_I, [_I, C.c_uint64, _I, C.POINTER(_P), C.POINTER(_P)]),
    "ecm_scenario_generate_box": (_I, [_F, _I, _I, _I, _I, C.c_uint64, C.POINTER(_P), C.POINTER(_P)]),
    "ecm_scenario_counts": (_I, [_P, _PI, C.POINTER(C.c_longlong)]),
    "ecm_scenario_export": (_I, [_P, _PF, _PF, _PF, _PF, _PI, _PI, _PF]),
    "ecm_scenario_destroy": (None, [_P]),
    # ecm_mlp.h
    "ecm_mlp_n_params": (_I, []),
    "ecm_mlp_create": (_I, [C.c_uint64, C.POINTER(_P)]),
    "ecm_mlp_destroy": (_I, [_P]),
    "ecm_mlp_get_params": (_I, [_P, _PF, _I]),
    "ecm_mlp_set_params": (_I, [_P, _PF, _I]),
    "ecm_mlp_dataset_reserve": (_I, [_P, C.c_int64]),
    "ecm_mlp_dataset_size": (_I, [_P, C.POINTER(C.c_int64)]),
    "ecm_mlp_dataset_clear": (_I, [_P]),
    "ecm_mlp_dataset_read": (_I, [_P, C.c_int64, _I, _PF, C.POINTER(C.c_uint32), _PF]),
    "ecm_sim_log_step": (_I, [_P, _P, _PI]),
    "ecm_mlp_train": (_I, [_P, _I, _I, _F, _F, _F, C.c_uint64, _PF]),
    "ecm_mlp_forward": (_I, [_P, _PF, _I, _PF]),
    "ecm_mlp_forward_noalloc": (_I, [_P, _P, _I, _P]),
    "ecm_mlp_loss_grad": (_I, [_P, _PF, C.POINTER(C.c_uint32), _PF, _I, _F, _F, _PF, _PF]),
    "ecm_sim_set_learned": (_I, [_P, _P, _F]),
}


class NativeError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libecmcrowd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(f"{LIB_PATH} is missing: run `python -m ecmgenerator_amd.build` first")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().ecm_last_error()
        raise NativeError(f"native call failed ({rc}): {msg.decode() if msg else ''}")


def fptr(a):
    return a.ctypes.data_as(_PF)


def iptr(a):
    return a.ctypes.data_as(_PI)


def vptr(a):
    return a.ctypes.data_as(C.c_void_p)
