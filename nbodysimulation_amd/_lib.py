"""ctypes binding of libnbody.so (the C ABI in include/nbody/nbody.h and nbody_ic.h).

The library is loaded from this package directory.  There is no fallback: if
libnbody.so is missing, importing the bindings raises, so no code path can
silently compute on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbody.so")
ABI_VERSION = 2  # NB_ABI_VERSION of the structs mirrored below

NB_OK = 0
NB_E_INVALID_ARG = 1
NB_E_NO_DEVICE = 2
NB_E_HIP = 3
NB_E_OUT_OF_MEMORY = 4
NB_E_TREE_OVERFLOW = 5
NB_E_DEPTH_OVERFLOW = 6
NB_E_NONFINITE = 7
NB_E_STATE = 8
NB_E_ADMISSION = 9
NOALLOC_MAX_BODIES = 1 << 24  # NB_NOALLOC_MAX_BODIES

# enum values of src/SimulationConfig.h:10-38
LEAPFROG, EULER = 1, 2
SPHERE_POS, UNIFORM_POS = 0, 1
RANDOM_VEL, STARTING_OUT_VEL, STARTING_IN_VEL, FUNC_ZERO_VEL, TANGENT_XZ_VEL = range(5)
BARNES_HUT, BRUTE_FORCE_GLOBAL, BRUTE_FORCE_LOCAL = 0, 1, 2
RULES_REFERENCE, RULES_STANDARD = 0, 1
LAW_PLUMMER, LAW_REFERENCE_UNIT = 0, 1
ACCUM_PLAIN, ACCUM_KAHAN = 0, 1
IC_NORMAL, IC_PLUMMER, IC_UNIFORM_BALL, IC_COLLISION = 0, 1, 2, 3

MAX_LEVEL = 64
DQ_LEN = MAX_LEVEL + 1
PATH_STRIDE = 72
NUM_STAGES = 9
STAGE_NAMES = ("bounds", "reset", "build", "summarize", "sort", "force", "update", "brute", "collision")


class NbConfig(C.Structure):
    _fields_ = [
        ("num_bodies", C.c_int32),
        ("algorithm", C.c_int32),
        ("method", C.c_int32),
        ("rules", C.c_int32),
        ("law", C.c_int32),
        ("accumulation", C.c_int32),
        ("max_octree_depth", C.c_int32),
        ("cell_capacity", C.c_int32),
        ("gravitational_constant", C.c_float),
        ("theta", C.c_float),
        ("epsilon", C.c_float),
        ("dt", C.c_float),
        ("collision", C.c_int32),
        ("use_graph", C.c_int32),
        ("time_stages", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class NbTreeInfo(C.Structure):
    _fields_ = [
        ("num_cells", C.c_int32),
        ("max_depth", C.c_int32),
        ("capacity", C.c_int32),
        ("status", C.c_int32),
    ]


class NbSystemSpec(C.Structure):
    _fields_ = [
        ("num_bodies", C.c_int32),
        ("algorithm", C.c_int32),
        ("rules", C.c_int32),
        ("cell_capacity", C.c_int32),
        ("gravitational_constant", C.c_float),
        ("theta", C.c_float),
        ("epsilon", C.c_float),
        ("dt", C.c_float),
        ("collision", C.c_int32),
    ]


class NbCheckpointHeader(C.Structure):
    _fields_ = [
        ("magic", C.c_uint32),
        ("version", C.c_uint32),
        ("num_bodies", C.c_int32),
        ("have_acc", C.c_int32),
        ("step", C.c_int64),
        ("config", NbConfig),
    ]


CHECKPOINT_MAGIC = 0x4B43424E
CHECKPOINT_VERSION = 1


class NbTuning(C.Structure):
    _fields_ = [
        ("bf_targets_per_lane", C.c_int32),
        ("bf_min_waves", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class NbServerStats(C.Structure):
    _fields_ = [
        ("active", C.c_int32),
        ("queued", C.c_int32),
        ("finished", C.c_int32),
        ("reserved", C.c_int32),
        ("arena_bytes", C.c_int64),
        ("body_steps", C.c_int64),
        ("ticks", C.c_int64),
    ]


class NbIcReferenceParams(C.Structure):
    _fields_ = [
        ("starting_speed_mul", C.c_float),
        ("starting_volume_radius", C.c_float),
        ("number_of_massive_objects", C.c_int32),
        ("massive_object_mass", C.c_float),
        ("pos_config", C.c_int32),
        ("vel_config", C.c_int32),
        ("mass_mean", C.c_float),
        ("mass_deviation", C.c_float),
    ]


class NbIcParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("total_mass", C.c_float),
        ("scale", C.c_float),
        ("mass_lo", C.c_float),
        ("mass_hi", C.c_float),
        ("cutoff", C.c_float),
        ("G", C.c_float),
        ("center", C.c_float * 3),
        ("bulk_velocity", C.c_float * 3),
    ]


_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)
_U8P = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); every entry point declared in include/nbody/*.h
SIGNATURES = {
    "nb_abi_version": (C.c_int32, []),
    "nb_config_default": (None, [C.POINTER(NbConfig)]),
    "nb_status_string": (C.c_char_p, [C.c_int]),
    "nb_last_error_detail": (C.c_char_p, []),
    "nb_device_info": (C.c_int, [_IP, C.c_char_p, C.c_size_t]),
    "nb_reference_cell_capacity": (C.c_int32, [C.c_int32, C.c_int32]),
    "nb_sim_create": (C.c_int, [C.POINTER(NbConfig), _P, C.POINTER(_P)]),
    "nb_sim_destroy": (C.c_int, [_P]),
    "nb_sim_upload": (C.c_int, [_P, _FP, _FP, _FP]),
    "nb_sim_download": (C.c_int, [_P, _FP, _FP, _FP]),
    "nb_sim_sync": (C.c_int, [_P]),
    "nb_sim_status": (C.c_int, [_P]),
    "nb_sim_clear_status": (C.c_int, [_P]),
    "nb_stage_bounds": (C.c_int, [_P, _FP]),
    "nb_stage_tree_reset": (C.c_int, [_P]),
    "nb_stage_tree_build": (C.c_int, [_P]),
    "nb_stage_summarize": (C.c_int, [_P]),
    "nb_stage_sort": (C.c_int, [_P]),
    "nb_stage_force": (C.c_int, [_P]),
    "nb_stage_update": (C.c_int, [_P]),
    "nb_stage_brute_force": (C.c_int, [_P]),
    "nb_stage_collision": (C.c_int, [_P]),
    "nb_stage_force_range": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "nb_sim_set_cell_capacity": (C.c_int, [_P, C.c_int32]),
    "nb_stage_brute_force_range": (C.c_int, [_P, C.c_int64, C.c_int64]),
    "nb_step": (C.c_int, [_P, C.c_int32]),
    "nb_tree_info_get": (C.c_int, [_P, C.POINTER(NbTreeInfo)]),
    "nb_export_bodies": (C.c_int, [_P, _IP, _U8P, _IP]),
    "nb_export_cells": (C.c_int, [_P, _IP, _IP, _IP, _FP]),
    "nb_export_force_order": (C.c_int, [_P, _IP]),
    "nb_export_dq": (C.c_int, [_P, _FP]),
    "nb_force_interactions": (C.c_int, [_P, _I64P]),
    "nb_export_interactions": (C.c_int, [_P, _IP]),
    "nb_ens_export_interactions": (C.c_int, [_P, C.c_int32, _IP]),
    "nb_stage_times": (C.c_int, [_P, _FP]),
    "nb_step_async": (C.c_int, [_P, C.c_int32]),
    "nb_sim_energy": (C.c_int, [_P, _DP]),
    "nb_energy_f64_noalloc": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_double, C.c_double, _P, _P]),
    "nb_force_vote_group": (C.c_int32, []),
    "nb_tuning_default": (None, [C.POINTER(NbTuning)]),
    "nb_set_tuning": (C.c_int, [C.POINTER(NbTuning)]),
    "nb_get_tuning": (C.c_int, [C.POINTER(NbTuning)]),
    "nb_export_vertices": (C.c_int, [_P, _FP, C.c_int32]),
    "nb_sim_step_count": (C.c_int, [_P, _I64P]),
    "nb_checkpoint_size": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "nb_checkpoint_save": (C.c_int, [_P, C.c_void_p, C.c_size_t]),
    "nb_checkpoint_load": (C.c_int, [_P, C.c_void_p, C.c_size_t]),
    "nb_ens_create": (C.c_int, [C.POINTER(NbSystemSpec), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P,
                                C.POINTER(_P)]),
    "nb_ens_destroy": (C.c_int, [_P]),
    "nb_ens_upload": (C.c_int, [_P, C.c_int32, _FP, _FP, _FP]),
    "nb_ens_download": (C.c_int, [_P, C.c_int32, _FP, _FP, _FP]),
    "nb_ens_step": (C.c_int, [_P, C.c_int32]),
    "nb_ens_status": (C.c_int, [_P, _IP]),
    "nb_ens_device_bytes": (C.c_int, [_P, _I64P]),
    "nb_server_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, C.POINTER(_P)]),
    "nb_server_destroy": (C.c_int, [_P]),
    "nb_server_submit": (C.c_int, [_P, C.POINTER(NbSystemSpec), C.c_int32, _FP, _FP, _FP, _I64P]),
    "nb_server_tick": (C.c_int, [_P, _IP]),
    "nb_server_poll": (C.c_int, [_P, _I64P, _IP]),
    "nb_server_fetch": (C.c_int, [_P, C.c_int64, _FP, _FP, _FP]),
    "nb_server_stats_get": (C.c_int, [_P, C.POINTER(NbServerStats)]),
    "nb_ic_reference_default": (None, [C.POINTER(NbIcReferenceParams)]),
    "nb_ic_params_default": (None, [C.POINTER(NbIcParams), C.c_int32]),
    "nb_ic_reference": (C.c_int, [C.POINTER(NbIcReferenceParams), C.c_int32, C.c_uint64, _FP, _FP, _FP]),
    "nb_ic_generate": (C.c_int, [C.POINTER(NbIcParams), C.c_int32, C.c_uint64, _FP, _FP, _FP]),
}

_lib = None


def load():
    """Load libnbody.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C nbodysimulation_amd)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.nb_abi_version() != ABI_VERSION:  # struct layouts of this module and the library differ
        raise RuntimeError(f"{LIB_PATH} has ABI version {lib.nb_abi_version()}, this module {ABI_VERSION}: rebuild it")
    _lib = lib
    return lib


def get_tuning() -> dict:
    t = NbTuning()
    check(load().nb_get_tuning(C.byref(t)), "nb_get_tuning")
    return {"bf_targets_per_lane": t.bf_targets_per_lane, "bf_min_waves": t.bf_min_waves}


def set_tuning(bf_targets_per_lane: int | None = None, bf_min_waves: int | None = None) -> None:
    """Process-wide launch-shape tuning; unspecified fields keep the library defaults."""
    t = NbTuning()
    load().nb_tuning_default(C.byref(t))
    if bf_targets_per_lane is not None:
        t.bf_targets_per_lane = int(bf_targets_per_lane)
    if bf_min_waves is not None:
        t.bf_min_waves = int(bf_min_waves)
    check(load().nb_set_tuning(C.byref(t)), "nb_set_tuning")


def vote_group() -> int:
    """Bodies per opening decision of the library's Barnes-Hut force kernel."""
    return int(load().nb_force_vote_group())


class NbError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = load()
        self.status = status
        detail = lib.nb_last_error_detail().decode(errors="replace")
        msg = lib.nb_status_string(status).decode()
        super().__init__(f"{where}: {msg} ({status}): {detail}")


def check(status: int, where: str) -> None:
    if status != NB_OK:
        raise NbError(status, where)
