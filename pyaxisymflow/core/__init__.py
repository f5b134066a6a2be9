"""pyaxisymflow/core: drop-ins from pyaxisymflow_amd.core."""
