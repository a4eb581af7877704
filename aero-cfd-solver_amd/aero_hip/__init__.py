"""aero_hip: MI355X (gfx950) FP64 backend of the AeroCFD 2D Euler residual path.

Importing the package loads libaero_hip.so (the HIP kernels + C ABI); a missing library raises.
"""
from . import _lib
from .mesh import Mesh, generate_mesh
from .solver import (CaseConfig, HipBatch, HipMesh, Seed, case_linearization, freestream, hip_available,
                     residual_config)

__all__ = ["Mesh", "generate_mesh", "CaseConfig", "HipBatch", "HipMesh", "Seed", "case_linearization", "freestream",
           "hip_available", "residual_config", "_lib"]

_lib.lib()
