"""pyaxisymflow/utils: drop-ins from pyaxisymflow_amd.utils."""
