"""MI355X-native DGSEM right-hand side (HIP kernels for gfx950 behind a C ABI).

The compute path is libdgsem_amd.so (dgsem_amd/csrc, declared in
include/dgsem_amd.h); this package only binds it for Python callers.
"""
from ._abi import (BC_DIRICHLET, BC_MIXED_DIRICHLET_SLIP_WALL, BC_NS_NOSLIP_WALL, BC_NS_SLIP_WALL, BC_OUTFLOW, BC_PERIODIC, BC_SLIP_WALL, EULER, FLUX_CENTRAL, FLUX_CHANDRASHEKAR, FLUX_HLLC,  # noqa: F401
                   FLUX_LLF, GEOMETRY_AFFINE, GEOMETRY_NODAL, LAYOUT_LEFT, LAYOUT_RIGHT, NAVIER_STOKES, SHOCK_CAPTURING, SPLIT_FORM,
                   DgsemError, load)
from .solver import SSPRK3, Batch, SolverConfig, StructuredSolver  # noqa: F401

__all__ = ["StructuredSolver", "SSPRK3", "Batch", "SolverConfig", "load", "DgsemError"]
