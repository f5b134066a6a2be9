"""pyaxisymflow_amd -- MI355X (gfx950) kernels for PyAxisymFlow's per-timestep
vorticity-streamfunction path.

Sub-packages mirror the reference's module layout and call signatures:
``kernels`` (pyaxisymflow/kernels), ``elasto_kernels``
(pyaxisymflow/elasto_kernels) and ``core`` (pyaxisymflow/core).  Every
function runs hand-written HIP kernels from ``libaxisym_hip.so`` through a C
ABI (include/axisym.h); there is no CPU path.  ``stepper`` holds the fused,
device-resident timestep.
"""
__version__ = "0.1.0"
