"""pyaxisymflow/kernels/compute_vorticity_from_velocity.py on the GPU (pyaxisymflow_amd.kernels.compute_vorticity_from_velocity)."""
from pyaxisymflow_amd.kernels.compute_vorticity_from_velocity import (  # noqa: F401
    compute_vorticity_from_velocity_unb,
    compute_vorticity_from_velocity_periodic,
)

__all__ = ['compute_vorticity_from_velocity_unb', 'compute_vorticity_from_velocity_periodic']
