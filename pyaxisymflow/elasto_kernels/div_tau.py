"""pyaxisymflow/elasto_kernels/div_tau.py on the GPU (pyaxisymflow_amd.elasto_kernels.div_tau)."""
from pyaxisymflow_amd.elasto_kernels.div_tau import (  # noqa: F401
    update_vorticity_from_solid_stress,
    update_vorticity_from_solid_stress_periodic,
)

__all__ = ['update_vorticity_from_solid_stress', 'update_vorticity_from_solid_stress_periodic']
