"""pyaxisymflow/kernels/FastDiagonalisationStokesSolver.py on the GPU (pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver)."""
from pyaxisymflow_amd.kernels.FastDiagonalisationStokesSolver import (  # noqa: F401
    FastDiagonalisationStokesSolver,
)

__all__ = ['FastDiagonalisationStokesSolver']
