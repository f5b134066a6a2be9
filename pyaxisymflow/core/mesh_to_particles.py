"""pyaxisymflow/core/mesh_to_particles.py on the GPU (pyaxisymflow_amd.core.mesh_to_particles)."""
from pyaxisymflow_amd.core.mesh_to_particles import (  # noqa: F401
    mesh_to_particles_2D_mp4,
    mesh_to_particles_2D_unbounded_mp4,
)

__all__ = ['mesh_to_particles_2D_mp4', 'mesh_to_particles_2D_unbounded_mp4']
