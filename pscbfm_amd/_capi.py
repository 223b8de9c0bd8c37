"""ctypes binding of the C ABI in include/pscbfm_amd.h and pscbfm_amd_systems.h.

The library is pscbfm_amd/lib/libpscbfm_amd.so, built by ``python -m
pscbfm_amd.build``.  There is no fallback: without the library every call
raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpscbfm_amd.so"

BFM_OK = 0
BFM_E_INVALID_ARGUMENT = -1
BFM_E_RUNTIME = -2
BFM_E_HIP = -3
BFM_E_NO_DEVICE = -4
BFM_E_OUT_OF_MEMORY = -5
BFM_E_STATE = -6


class BfmError(RuntimeError):
    """A failed call; .code holds the BFM_E_* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class InvalidArgument(BfmError, ValueError):
    """The reference's std::invalid_argument."""


class HipError(BfmError):
    """A HIP runtime call failed."""


class NoDeviceError(BfmError):
    """No HIP device is visible."""


class StateError(BfmError):
    """Call out of order (e.g. run before initialize)."""


_ERRORS = {
    BFM_E_INVALID_ARGUMENT: InvalidArgument,
    BFM_E_RUNTIME: BfmError,
    BFM_E_HIP: HipError,
    BFM_E_NO_DEVICE: NoDeviceError,
    BFM_E_OUT_OF_MEMORY: BfmError,
    BFM_E_STATE: StateError,
}


class BfmStats(ctypes.Structure):
    _fields_ = [
        ("attempts", c_uint64),
        ("proposals", c_uint64),
        ("accepted", c_uint64),
        ("substeps", c_uint64),
        ("check_ms", c_double),
        ("perform_ms", c_double),
        ("timed_launches", c_uint64),
        ("sorts", c_uint64),
    ]


class BfmObservables(ctypes.Structure):
    _fields_ = [
        ("bonds", c_uint64),
        ("bond_hist", c_uint64 * 512),
        ("n_chains", c_uint32),
        ("reserved", c_uint32),
        ("re2_sum", c_int64),
        ("re2", c_double),
        ("rg2", c_double),
        ("box_com", c_double * 3),
    ]


class BfmChainSums(ctypes.Structure):
    _fields_ = [
        ("s", c_int64 * 3),
        ("q", c_int64),
        ("first", c_int32 * 3),
        ("middle", c_int32 * 3),
        ("last", c_int32 * 3),
        ("n", c_uint32),
    ]


class BfmMsdPoint(ctypes.Structure):
    _fields_ = [
        ("dt", c_uint64),
        ("n_values", c_double),
        ("mean", c_double * 5),
        ("sg", c_double * 5),
    ]


# name -> (restype, argtypes); every symbol declared in include/*.h
_H = c_void_p
SIGNATURES = {
    "bfm_last_error": (c_char_p, []),
    "bfm_create": (c_int, [POINTER(c_void_p)]),
    "bfm_destroy": (c_int, [_H]),
    "bfm_set_gpu": (c_int, [_H, c_int]),
    "bfm_copy_bond_set": (c_int, [_H, c_int, c_int, c_int, c_int]),
    "bfm_set_lattice_size": (c_int, [_H, c_uint64, c_uint64, c_uint64]),
    "bfm_set_nr_of_all_monomers": (c_int, [_H, c_uint32]),
    "bfm_set_attribute": (c_int, [_H, c_uint32, c_int32]),
    "bfm_set_monomer_coordinates": (c_int, [_H, c_uint32, c_int32, c_int32, c_int32]),
    "bfm_set_connectivity": (c_int, [_H, c_uint32, c_uint32]),
    "bfm_set_periodicity": (c_int, [_H, c_int, c_int, c_int]),
    "bfm_set_age": (c_int, [_H, c_int64]),
    "bfm_get_age": (c_int, [_H, POINTER(c_int64)]),
    "bfm_set_rng": (c_int, [_H, c_int]),
    "bfm_get_rng": (c_int, [_H, POINTER(c_int)]),
    "bfm_set_seed": (c_int, [_H, c_uint64]),
    "bfm_set_steps_between_sortings": (c_int, [_H, c_int64]),
    "bfm_get_steps_between_sortings": (c_int, [_H, POINTER(c_int64)]),
    "bfm_validate": (c_int, [_H]),
    "bfm_initialize": (c_int, [_H]),
    "bfm_run_simulation_on_gpu": (c_int, [_H, c_uint32]),
    "bfm_run_simulation_on_gpu_noalloc": (c_int, [_H, c_uint32]),
    "bfm_noalloc_max_monomers": (c_uint32, []),
    "bfm_get_monomer_position_in_x": (c_int, [_H, c_uint32, POINTER(c_int32)]),
    "bfm_get_monomer_position_in_y": (c_int, [_H, c_uint32, POINTER(c_int32)]),
    "bfm_get_monomer_position_in_z": (c_int, [_H, c_uint32, POINTER(c_int32)]),
    "bfm_cleanup": (c_int, [_H]),
    "bfm_set_monomer_coordinates_bulk": (c_int, [_H, c_void_p, c_uint32]),
    "bfm_set_connectivity_bulk": (c_int, [_H, c_void_p, c_uint64]),
    "bfm_get_monomer_positions": (c_int, [_H, c_void_p, c_uint32]),
    "bfm_get_colors": (c_int, [_H, c_void_p, c_uint32, POINTER(c_uint32)]),
    "bfm_get_stats": (c_int, [_H, POINTER(BfmStats)]),
    "bfm_set_kernel_timing": (c_int, [_H, c_int]),
    "bfm_set_spatial_sorting": (c_int, [_H, c_int]),
    "bfm_set_perform_group": (c_int, [_H, c_uint32]),
    "bfm_set_check_block": (c_int, [_H, c_uint32]),
    "bfm_set_segment_config": (c_int, [_H, c_uint32, c_uint32, c_uint32]),
    "bfm_set_check_items": (c_int, [_H, c_uint32]),
    "bfm_set_segment_items": (c_int, [_H, c_uint32, c_uint32]),
    "bfm_get_check_items": (c_int, [_H, c_uint32, POINTER(c_uint32)]),
    "bfm_get_launch_config": (c_int, [_H, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int), POINTER(c_int)]),
    "bfm_autotune": (c_int, [_H, c_uint32, POINTER(c_uint32), POINTER(c_int)]),
    "bfm_batch_set_perform_group": (c_int, [_H, c_uint32]),
    "bfm_batch_autotune": (c_int, [_H, c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_int)]),
    "bfm_read_bfm_file": (c_int, [_H, c_char_p]),
    "bfm_write_bfm_file": (c_int, [_H, c_char_p, c_int]),
    "bfm_set_xcd_ranges": (c_int, [_H, c_int]),
    "bfm_run_resident": (c_int, [_H, c_uint32]),
    "bfm_copy_back": (c_int, [_H]),
    "bfm_observe": (c_int, [_H, POINTER(BfmObservables)]),
    "bfm_chain_table": (c_int, [_H, POINTER(c_uint32), c_void_p, c_void_p, c_void_p, c_uint32]),
    "bfm_get_chain_sums": (c_int, [_H, c_void_p, c_uint32]),
    "bfm_msd_attach": (c_int, [_H, c_uint64, c_uint64, c_uint64, c_uint64]),
    "bfm_msd_execute": (c_int, [_H]),
    "bfm_msd_get_points": (c_int, [_H, c_void_p, c_uint32, POINTER(c_uint32)]),
    "bfm_msd_write": (c_int, [_H, c_char_p]),
    "bfm_batch_create": (c_int, [POINTER(c_void_p), c_int]),
    "bfm_batch_destroy": (c_int, [_H]),
    "bfm_batch_add": (c_int, [_H, _H, POINTER(c_uint32)]),
    "bfm_member_bytes": (c_int, [_H, POINTER(c_uint64)]),
    "bfm_batch_run": (c_int, [_H, c_uint32, c_uint32]),
    "bfm_batch_run_members": (c_int, [_H, c_void_p, c_uint32, c_uint32, c_uint32]),
    "bfm_batch_get_positions": (c_int, [_H, c_uint32, c_void_p, c_uint32]),
    "bfm_batch_get_age": (c_int, [_H, c_uint32, POINTER(c_int64)]),
    "bfm_batch_get_stats": (c_int, [_H, c_uint32, POINTER(BfmStats)]),
    "bfm_batch_observe": (c_int, [_H, c_uint32, POINTER(BfmObservables)]),
    "bfm_batch_remove": (c_int, [_H, c_uint32]),
    "bfm_batch_set_graphs": (c_int, [_H, c_int]),
    "bfm_batch_set_xcd_ranges": (c_int, [_H, c_int]),
    "bfm_batch_set_working_set_bytes": (c_int, [_H, c_uint64]),
    "bfm_batch_get_working_set_groups": (c_int, [_H, POINTER(c_uint32)]),
    "bfm_batch_synchronize": (c_int, [_H]),
    "bfm_server_create": (c_int, [POINTER(c_void_p), c_int, c_uint64, c_uint32]),
    "bfm_server_destroy": (c_int, [_H]),
    "bfm_server_submit": (c_int, [_H, _H, c_uint32, POINTER(c_uint64)]),
    "bfm_server_submit_take": (c_int, [_H, _H, c_uint32, POINTER(c_uint64)]),
    "bfm_server_step": (c_int, [_H, POINTER(c_uint32)]),
    "bfm_server_job_state": (c_int, [_H, c_uint64, POINTER(c_int), POINTER(c_uint32)]),
    "bfm_server_job_times": (c_int, [_H, c_uint64, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    "bfm_server_job_result": (c_int, [_H, c_uint64, c_void_p, c_uint32]),
    "bfm_server_job_error": (c_int, [_H, c_uint64, c_char_p, c_uint32]),
    "bfm_server_cancel": (c_int, [_H, c_uint64]),
    "bfm_server_release": (c_int, [_H, c_uint64]),
    "bfm_server_info": (c_int, [_H, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint64),
                                POINTER(c_uint64), POINTER(c_uint64)]),
    "bfm_server_host_bytes": (c_int, [_H, POINTER(c_uint64), POINTER(c_uint32), POINTER(c_uint32)]),
    "bfm_server_graph_stats": (c_int, [_H, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    "bfm_batch_graph_stats": (c_int, [_H, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    "bfm_device_count": (c_int, [POINTER(c_int)]),
    "bfm_calibration_copy": (c_int, [c_int, c_uint64, c_int, POINTER(c_double)]),
    "bfm_gen_linear_melt": (c_int, [c_uint32, c_uint32, c_uint32, c_uint64, c_void_p, c_void_p]),
    "bfm_gen_diamond_network_size": (c_int, [c_uint32, c_uint32, c_uint32, c_uint32, c_uint32,
                                             POINTER(c_uint32), POINTER(c_uint32)]),
    "bfm_gen_diamond_network": (c_int, [c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_uint64,
                                        c_void_p, c_void_p]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pscbfm_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != BFM_OK:
        msg = load().bfm_last_error()
        text = msg.decode() if msg else ""
        raise _ERRORS.get(rc, BfmError)(rc, f"status {rc}: {text}")


def calibration_copy(gpu: int, nbytes: int, repeats: int) -> float:
    """Average milliseconds of one plain copy-kernel launch over nbytes."""
    ms = c_double(0.0)
    check(load().bfm_calibration_copy(gpu, nbytes, repeats, ctypes.byref(ms)))
    return ms.value


def device_count() -> int:
    n = c_int(0)
    check(load().bfm_device_count(ctypes.byref(n)))
    return n.value
