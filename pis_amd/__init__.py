"""pis_amd -- MI355X-native Lennard-Jones force and velocity-Verlet engine.

The compute path is libpis_hip.so (hand-written gfx950 kernels behind the C ABI of
include/pis_hip.h).  Importing this package does not load the library; the first call
that needs it does, and fails loudly if it has not been built.
"""
from ._capi import PisError, lib  # noqa: F401
from .box import SimulationBox, box_inverse  # noqa: F401
from .engine import Atoms, Engine, LennardJones, LJGpuManager, plan, thermo_to_array  # noqa: F401
from .serve import JobResult, Server  # noqa: F401
from .workloads import SystemSpec  # noqa: F401

__all__ = [
    "Server",
    "JobResult",
    "Atoms",
    "Engine",
    "LennardJones",
    "LJGpuManager",
    "PisError",
    "SimulationBox",
    "SystemSpec",
    "box_inverse",
    "lib",
    "plan",
    "thermo_to_array",
]
