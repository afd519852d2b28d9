"""MI355X (gfx950) DVM Boltzmann time step: variance-reduced collisions with
conservative remap, upwind convection, wall / specular / inflow boundaries and
FP64 moments, as hand-written HIP kernels behind the C-ABI of include/dvm_hip.h.
"""
from . import capi, config
from .batch import Batch
from .config import (SolverConfig, batched_homogeneous, homogeneous_relaxation, lid_driven_cavity, maxwellian,
                     mixture_instance, normal_shock_1d, node_velocities, rankine_hugoniot, sound_speed,
                     uniform_axis)
from .dvm import DvmError, Solver, load_library
from .serve import Request, Server, shock_instance, synthetic_requests

__all__ = ["capi", "SolverConfig", "Solver", "Batch", "DvmError", "load_library", "homogeneous_relaxation",
           "batched_homogeneous", "mixture_instance", "Request", "Server", "shock_instance",
           "synthetic_requests",
           "lid_driven_cavity", "normal_shock_1d", "maxwellian", "node_velocities", "rankine_hugoniot",
           "sound_speed", "uniform_axis", "config"]
