"""pyaxisymflow/elasto_kernels/extrapolate_using_least_squares.py on the GPU (pyaxisymflow_amd.elasto_kernels.extrapolate_using_least_squares)."""
from pyaxisymflow_amd.elasto_kernels.extrapolate_using_least_squares import (  # noqa: F401
    extrapolate_eta_using_least_squares,
)

__all__ = ['extrapolate_eta_using_least_squares']
