"""pyaxisymflow/kernels/FastDiagonalisationPotentialSolver.py on the GPU (pyaxisymflow_amd.kernels.FastDiagonalisationPotentialSolver)."""
from pyaxisymflow_amd.kernels.FastDiagonalisationPotentialSolver import (  # noqa: F401
    FastDiagonalisationPotentialSolver,
)

__all__ = ['FastDiagonalisationPotentialSolver']
