"""pyaxisymflow/kernels/compute_forces.py on the GPU (pyaxisymflow_amd.kernels.compute_forces)."""
from pyaxisymflow_amd.kernels.compute_forces import (  # noqa: F401
    penalization_force_sum,
    compute_force_on_body,
    max_abs_velocity_sum,
)

__all__ = ['penalization_force_sum', 'compute_force_on_body', 'max_abs_velocity_sum']
