"""pyaxisymflow/kernels/advect_particle.py on the GPU (pyaxisymflow_amd.kernels.advect_particle)."""
from pyaxisymflow_amd.kernels.advect_particle import (  # noqa: F401
    advect_vorticity_via_particles,
    advect_vorticity_via_particles_periodic,
)

__all__ = ['advect_vorticity_via_particles', 'advect_vorticity_via_particles_periodic']
