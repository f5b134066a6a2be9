"""pyaxisymflow/utils/dump_vtk.py (pyaxisymflow_amd.utils.dump_vtk, no vtk package needed)."""
from pyaxisymflow_amd.utils.dump_vtk import vtk_init, vtk_write, read_vti  # noqa: F401

__all__ = ['vtk_init', 'vtk_write', 'read_vti']
