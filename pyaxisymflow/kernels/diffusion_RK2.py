"""pyaxisymflow/kernels/diffusion_RK2.py on the GPU (pyaxisymflow_amd.kernels.diffusion_RK2)."""
from pyaxisymflow_amd.kernels.diffusion_RK2 import (  # noqa: F401
    diffusion_RK2_unb,
    diffusion_RK2_periodic,
)

__all__ = ['diffusion_RK2_unb', 'diffusion_RK2_periodic']
