"""pyaxisymflow/kernels/advect_vorticity_via_eno3.py on the GPU (pyaxisymflow_amd.kernels.advect_vorticity_via_eno3)."""
from pyaxisymflow_amd.kernels.advect_vorticity_via_eno3 import (  # noqa: F401
    gen_advect_vorticity_via_eno3,
    gen_advect_vorticity_via_eno3_periodic,
)

__all__ = ['gen_advect_vorticity_via_eno3', 'gen_advect_vorticity_via_eno3_periodic']
