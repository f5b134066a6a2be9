"""pyaxisymflow/core/particles_to_mesh.py on the GPU (pyaxisymflow_amd.core.particles_to_mesh)."""
from pyaxisymflow_amd.core.particles_to_mesh import (  # noqa: F401
    particles_to_mesh_2D_mp4,
    particles_to_mesh_2D_unbounded_mp4,
)

__all__ = ['particles_to_mesh_2D_mp4', 'particles_to_mesh_2D_unbounded_mp4']
