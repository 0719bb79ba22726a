"""jno_amd -- MI355X (gfx950) native finite-element hot path of jNO.

Lagrange P1/P2 tetrahedral assembly into CSR, symmetric Dirichlet elimination and Jacobi-PCG
(FP64, FP32-inner with FP64 residual replacement), with a hand-written adjoint for dL/dk.  All
compute runs in the HIP library ``libjno_fem.so`` (C-ABI in include/jno_fem.h) called through
ctypes over PyTorch-ROCm device tensors; there is no CPU fallback.
"""
from . import interop, precond, solve, transient  # noqa: F401
from ._lib import JnoError, load  # noqa: F401
from .fem import BCOO, CSR, FemLinearSystem, Plan, fem_poisson  # noqa: F401
from .mesh import Mesh, Shape, kuhn_box  # noqa: F401
from .transient import SemidiscreteTimeBlock, theta_march  # noqa: F401

__all__ = ["Shape", "Mesh", "kuhn_box", "Plan", "CSR", "BCOO", "FemLinearSystem", "fem_poisson", "solve", "precond",
           "interop", "transient", "SemidiscreteTimeBlock", "theta_march", "JnoError", "load"]
