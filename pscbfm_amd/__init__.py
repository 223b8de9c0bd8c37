"""MI355X-native bond-fluctuation-model Monte Carlo engine.

The hot path (graph-coloured species sweep) runs as hand-written HIP kernels
for gfx950 in lib/libpscbfm_amd.so; this package is the host-side face of its
C ABI (include/pscbfm_amd.h).
"""
from ._capi import BfmError, HipError, InvalidArgument, NoDeviceError, StateError, device_count, load
from .updater import GPUScBFM_AB_Type, Rng_IntHash, Rng_Philox, UpdaterGPUScBFM_AB_Type, updater_from_system
from .batch import Batch, Server, member_bytes

__all__ = [
    "Batch",
    "BfmError",
    "Server",
    "member_bytes",
    "GPUScBFM_AB_Type",
    "HipError",
    "InvalidArgument",
    "NoDeviceError",
    "Rng_IntHash",
    "Rng_Philox",
    "StateError",
    "UpdaterGPUScBFM_AB_Type",
    "device_count",
    "load",
    "updater_from_system",
]
