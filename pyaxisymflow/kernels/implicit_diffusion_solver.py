"""pyaxisymflow/kernels/implicit_diffusion_solver.py on the GPU (pyaxisymflow_amd.kernels.implicit_diffusion_solver)."""
from pyaxisymflow_amd.kernels.implicit_diffusion_solver import (  # noqa: F401
    ImplicitEulerDiffusionStepper,
)

__all__ = ['ImplicitEulerDiffusionStepper']
