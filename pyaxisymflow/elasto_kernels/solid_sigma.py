"""pyaxisymflow/elasto_kernels/solid_sigma.py on the GPU (pyaxisymflow_amd.elasto_kernels.solid_sigma)."""
from pyaxisymflow_amd.elasto_kernels.solid_sigma import (  # noqa: F401
    solid_sigma,
    solid_sigma_periodic,
)

__all__ = ['solid_sigma', 'solid_sigma_periodic']
