"""pyaxisymflow/kernels/smooth_Heaviside.py on the GPU (pyaxisymflow_amd.kernels.smooth_Heaviside)."""
from pyaxisymflow_amd.kernels.smooth_Heaviside import smooth_Heaviside  # noqa: F401

__all__ = ['smooth_Heaviside']
