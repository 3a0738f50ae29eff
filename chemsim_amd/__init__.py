"""MI355X (gfx950) implementation of taktoa/chemsim's D2Q9 lattice-Boltzmann hot path.

The compute path is libchemsim_hip.so (chemsim_amd/csrc, C ABI in
include/chemsim.h).  Importing this package loads it; there is no CPU path.
"""
from . import capi  # noqa: F401  (loads the shared library or raises)
from .lattice import Context, Lattice, Params, STORAGE  # noqa: F401
from .capi import (CS_F32, CS_F64, CS_BF16DEV, CS_F16DEV, CS_F32DEV, CS_EDGE_ZERO, CS_EDGE_PERIODIC,  # noqa: F401
                   CS_BB_REFERENCE, CS_BB_HALFWAY, CS_FLUID, CS_WALL, CS_LID, CS_INLET, CS_OUTLET)

__all__ = ["Context", "Lattice", "Params", "STORAGE"]
