"""pyaxisymflow/kernels/compute_velocity_from_psi.py on the GPU (pyaxisymflow_amd.kernels.compute_velocity_from_psi)."""
from pyaxisymflow_amd.kernels.compute_velocity_from_psi import (  # noqa: F401
    compute_velocity_from_psi_unb,
    compute_velocity_from_psi_periodic,
)

__all__ = ['compute_velocity_from_psi_unb', 'compute_velocity_from_psi_periodic']
