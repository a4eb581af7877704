"""aero_hip: MI355X (gfx950) FP64 backend of the AeroCFD 2D Euler residual path.

Importing the package loads libaero_hip.so (the HIP kernels + C ABI); a missing library raises.
"""
from . import _lib
from . import adjoint
from .case_file import euler_case_from_kv, load_case, read_native_case
from .mesh import Mesh, deform_mesh, generate_mesh, load_mesh, save_mesh
from .outputs import run_case_file, run_sweep_file
from .solver import (CaseConfig, HipBatch, HipMesh, Seed, adjoint_solve_batch, case_linearization, freestream,
                     hip_available, residual_config)

__all__ = ["Mesh", "deform_mesh", "generate_mesh", "load_mesh", "save_mesh", "CaseConfig", "HipBatch", "HipMesh", "Seed", "adjoint_solve_batch",
           "case_linearization", "freestream", "hip_available", "residual_config", "adjoint", "_lib",
           "euler_case_from_kv", "load_case", "read_native_case", "run_case_file", "run_sweep_file"]

_lib.lib()
