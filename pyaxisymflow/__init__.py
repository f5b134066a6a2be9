"""The reference's import paths (pyaxisymflow.kernels, .elasto_kernels,
.core, .utils.dump_vtk) for the MI355X drop-ins, so that a driver written against the
reference runs unchanged: each module here re-exports the same-named
module of pyaxisymflow_amd, whose functions keep the reference's names,
signatures, in-place semantics and return values and run on the GPU through
libaxisym_hip.so."""
