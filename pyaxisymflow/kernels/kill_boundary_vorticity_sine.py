"""pyaxisymflow/kernels/kill_boundary_vorticity_sine.py on the GPU (pyaxisymflow_amd.kernels.kill_boundary_vorticity_sine)."""
from pyaxisymflow_amd.kernels.kill_boundary_vorticity_sine import (  # noqa: F401
    kill_boundary_vorticity_sine_z,
    kill_boundary_vorticity_sine_r,
)

__all__ = ['kill_boundary_vorticity_sine_z', 'kill_boundary_vorticity_sine_r']
