"""ctypes binding of include/pis_hip.h (libpis_hip.so, built in-tree).

The library is the product: every compute call goes to the gfx950 kernels.  Loading
fails loudly if the shared object is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpis_hip.so")

PIS_OK = 0
STATUS_NAMES = {
    0: "PIS_OK",
    1: "PIS_ERR_NO_ATOMS",
    2: "PIS_ERR_MISSING_POTENTIAL",
    3: "PIS_ERR_NAN_ENERGY",
    4: "PIS_ERR_INVALID_ATOM_TYPE",
    5: "PIS_ERR_INVALID_ARGUMENT",
    6: "PIS_ERR_UNSUPPORTED",
    7: "PIS_ERR_HIP",
    8: "PIS_ERR_OUT_OF_MEMORY",
    9: "PIS_ERR_NO_DEVICE",
    10: "PIS_ERR_CAPACITY",
    11: "PIS_ERR_NOT_FOUND",
    12: "PIS_ERR_TIMEOUT",
    13: "PIS_ERR_IO",
}
PRECISION_FP64 = 0
PRECISION_MIXED = 1


class PairCoeff(C.Structure):
    _fields_ = [
        ("type_i", C.c_uint32),
        ("type_j", C.c_uint32),
        ("epsilon", C.c_double),
        ("sigma", C.c_double),
        ("rcut", C.c_double),
    ]


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_uint64),
        ("n_types", C.c_uint32),
        ("n_pairs", C.c_uint32),
        ("type_ids", C.POINTER(C.c_uint32)),
        ("masses", C.POINTER(C.c_double)),
        ("pairs", C.POINTER(PairCoeff)),
        ("h", C.c_double * 9),
        ("h_inv", C.c_double * 9),
        ("pbc", C.c_uint8 * 3),
        ("shift", C.c_uint8),
        ("precision", C.c_uint8),
        ("tile", C.c_uint8 * 3),
        ("skin", C.c_double),
        ("kb", C.c_double),
        ("neighbor_capacity", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class Thermo(C.Structure):
    _fields_ = [
        ("pe", C.c_double),
        ("ke", C.c_double),
        ("etot", C.c_double),
        ("temperature", C.c_double),
        ("pressure", C.c_double),
        ("virial_abs", C.c_double),
        ("virial_pair", C.c_double),
    ]


THERMO_FIELDS = [f[0] for f in Thermo._fields_]


class PlanInfo(C.Structure):
    _fields_ = [
        ("ncell", C.c_int32 * 3),
        ("tile", C.c_int32 * 3),
        ("ntile", C.c_int32 * 3),
        ("mode", C.c_int32),
        ("neighbor_capacity", C.c_int32),
        ("stage_capacity", C.c_int32),
        ("r2_list_max", C.c_double),
        ("density", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("rebuilds", C.c_uint64),
        ("force_launches", C.c_uint64),
        ("force_kernel_ms", C.c_double),
        ("timed_force_launches", C.c_uint64),
        ("neighbor_capacity", C.c_int32),
        ("max_neighbors", C.c_int32),
    ]


class MTKParams(C.Structure):
    _fields_ = [("p_start", C.c_double), ("tau", C.c_double), ("reserved", C.c_double * 2)]


class NHCParams(C.Structure):
    _fields_ = [
        ("t_start", C.c_double),
        ("t_end", C.c_double),
        ("tau", C.c_double),
        ("chain", C.c_int32),
        ("reserved", C.c_int32),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("job_id", C.c_uint64),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
        ("n_atoms", C.c_uint64),
        ("steps", C.c_uint64),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
        ("last", Thermo),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("rounds", C.c_uint64),
        ("atom_steps", C.c_uint64),
        ("resident", C.c_uint64),
        ("pending", C.c_uint64),
        ("bytes_in_use", C.c_uint64),
        ("peak_bytes", C.c_uint64),
        ("peak_resident", C.c_uint32),
        ("reserved", C.c_uint32),
        ("graph_instantiations", C.c_uint64),
        ("graph_build_s", C.c_double),
        ("pool_bytes", C.c_uint64),
    ]


# name -> (restype, argtypes); every function declared in include/pis_hip.h
_P = C.c_void_p
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "pis_abi_version": (C.c_int, []),
    "pis_last_error": (C.c_char_p, []),
    "pis_status_string": (C.c_char_p, [C.c_int]),
    "pis_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "pis_plan_system": (C.c_int, [C.POINTER(SystemDesc), C.POINTER(PlanInfo)]),
    "pis_set_tile_override": (C.c_int, [C.POINTER(C.c_int32)]),
    "pis_ctx_create": (C.c_int, [C.POINTER(SystemDesc), C.POINTER(_P)]),
    "pis_ctx_destroy": (None, [_P]),
    "pis_set_state": (C.c_int, [_P, _DP, _DP, _DP]),
    "pis_get_state": (C.c_int, [_P, _DP, _DP, _DP]),
    "pis_compute_potential": (C.c_int, [_P, _DP, _DP]),
    "pis_compute_potential_noalloc": (C.c_int, [C.POINTER(SystemDesc), _P, _P, _P, _P, _P]),
    "pis_verlet_step_nve": (C.c_int, [_P, C.c_double, C.c_uint64, C.POINTER(Thermo)]),
    "pis_get_neighbor_list": (
        C.c_int,
        [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(C.c_uint64)],
    ),
    "pis_get_list_digest": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "pis_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "pis_set_kernel_timing": (C.c_int, [_P, C.c_int32]),
    "pis_set_graph_mode": (C.c_int, [_P, C.c_int32]),
    "pis_copy_calibration": (C.c_int, [C.c_uint64, C.c_int32, _DP]),
    "pis_batch_create": (C.c_int, [C.POINTER(SystemDesc), C.c_uint32, C.POINTER(_P)]),
    "pis_batch_size": (C.c_uint32, [_P]),
    "pis_sys_set_state": (C.c_int, [_P, C.c_uint32, _DP, _DP, _DP]),
    "pis_sys_get_state": (C.c_int, [_P, C.c_uint32, _DP, _DP, _DP]),
    "pis_batch_compute_potential": (C.c_int, [_P, _DP]),
    "pis_batch_verlet_step_nve": (C.c_int, [_P, _DP, C.c_uint64, C.POINTER(Thermo)]),
    "pis_nvt_init": (C.c_int, [_P, C.POINTER(NHCParams)]),
    "pis_verlet_step_nvt": (C.c_int, [_P, C.c_double, C.c_uint64, C.POINTER(Thermo)]),
    "pis_batch_verlet_step_nvt": (C.c_int, [_P, _DP, C.c_uint64, C.POINTER(Thermo)]),
    "pis_nvt_get_chain": (C.c_int, [_P, C.c_uint32, _DP, _DP]),
    "pis_batch_verlet_step_nvt_ramp": (
        C.c_int, [_P, _DP, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(Thermo)]),
    "pis_npt_init": (C.c_int, [_P, C.POINTER(NHCParams), C.POINTER(MTKParams)]),
    "pis_batch_verlet_step_npt": (
        C.c_int, [_P, _DP, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(Thermo)]),
    "pis_npt_get_box": (C.c_int, [_P, C.c_uint32, _DP, _DP]),
    "pis_server_create": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(_P)]),
    "pis_server_submit": (
        C.c_int,
        [_P, C.POINTER(SystemDesc), _DP, _DP, C.c_double, C.c_uint64, C.POINTER(C.c_uint64)],
    ),
    "pis_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(JobResult), _DP, _DP, _DP]),
    "pis_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "pis_server_destroy": (C.c_int, [_P]),
    "pis_dump_write_block": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, _DP, C.POINTER(C.c_uint32), _DP,
                                       C.c_int32]),
    "pis_format_f64": (C.c_int, [C.c_double, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pis_snapshot_positions": (C.c_int, [_P, C.c_uint32, _DP]),
    "pis_snapshot_wait": (C.c_int, [_P]),
    "pis_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "pis_host_free": (C.c_int, [C.c_void_p]),
}

_lib = None


class PisError(RuntimeError):
    """A non-zero pis_status, with pis_last_error()'s message."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.name}: {message}")


def lib() -> C.CDLL:
    """Load libpis_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != PIS_OK:
        msg = lib().pis_last_error().decode(errors="replace")
        raise PisError(status, msg)
