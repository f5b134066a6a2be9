"""pyaxisymflow/core/extrapolate_using_least_squares.py on the GPU (pyaxisymflow_amd.core.extrapolate_using_least_squares)."""
from pyaxisymflow_amd.core.extrapolate_using_least_squares import (  # noqa: F401
    extrapolate_using_least_squares_till_first_order,
)

__all__ = ['extrapolate_using_least_squares_till_first_order']
