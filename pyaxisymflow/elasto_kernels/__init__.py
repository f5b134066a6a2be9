"""pyaxisymflow/elasto_kernels: drop-ins from pyaxisymflow_amd.elasto_kernels."""
