"""pyaxisymflow/elasto_kernels/extrapolate_eta_using_least_squares_unb.py on the GPU (pyaxisymflow_amd.elasto_kernels.extrapolate_eta_using_least_squares_unb)."""
from pyaxisymflow_amd.elasto_kernels.extrapolate_eta_using_least_squares_unb import (  # noqa: F401
    extrapolate_eta_with_least_squares,
)

__all__ = ['extrapolate_eta_with_least_squares']
