"""pyaxisymflow/kernels/periodic_boundary_ghost_comm.py on the GPU (pyaxisymflow_amd.kernels.periodic_boundary_ghost_comm)."""
from pyaxisymflow_amd.kernels.periodic_boundary_ghost_comm import (  # noqa: F401
    gen_periodic_boundary_ghost_comm,
    gen_periodic_boundary_ghost_comm_eta,
    plain_ghost_size,
)

__all__ = ['gen_periodic_boundary_ghost_comm', 'gen_periodic_boundary_ghost_comm_eta', 'plain_ghost_size']
