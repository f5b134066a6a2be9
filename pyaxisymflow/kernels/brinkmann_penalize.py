"""pyaxisymflow/kernels/brinkmann_penalize.py on the GPU (pyaxisymflow_amd.kernels.brinkmann_penalize)."""
from pyaxisymflow_amd.kernels.brinkmann_penalize import brinkmann_penalize  # noqa: F401

__all__ = ['brinkmann_penalize']
