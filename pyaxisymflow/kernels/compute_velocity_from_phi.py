"""pyaxisymflow/kernels/compute_velocity_from_phi.py on the GPU (pyaxisymflow_amd.kernels.compute_velocity_from_phi)."""
from pyaxisymflow_amd.kernels.compute_velocity_from_phi import (  # noqa: F401
    compute_velocity_from_phi_unb,
)

__all__ = ['compute_velocity_from_phi_unb']
