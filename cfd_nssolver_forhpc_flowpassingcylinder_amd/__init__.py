"""MI355X (gfx950) time-step library for incompressible flow past a cylinder.

The compute path is libcfdmx.so (hand-written HIP kernels behind the C-ABI of
include/cfdmx.h).  `FlowSolver` is the ctypes view of one device context.
"""
from .api import (  # noqa: F401
    Config, FlowSolver, StepInfo, CfdmxError, default_config, load_library, LIB_PATH,
    EXPORTS, GRID_UNIFORM, GRID_SML, SOLVER_BICGSTAB, SOLVER_RBSOR, PRECOND_NONE,
    PRECOND_JACOBI, PRECISION_FP64, PRECISION_MIXED,
)

__all__ = ["Config", "FlowSolver", "StepInfo", "CfdmxError", "default_config", "load_library"]
