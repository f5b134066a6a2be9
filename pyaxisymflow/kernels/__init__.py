"""pyaxisymflow/kernels: drop-ins from pyaxisymflow_amd.kernels."""
