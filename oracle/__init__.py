"""CPU oracle for the axisymmetric vorticity-streamfunction step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``pyaxisymflow_amd``)
imports, links or executes anything in this package.  The only permitted
users are ``tests/``, ``__graft_entry__.smoke()`` (as the checker) and the
``cpu_baseline`` leg of ``bench.py`` (as the timed CPU restatement).

Contents
--------
kernels.py   numpy restatement of the reference's numba / numpy kernels
             (pyaxisymflow/kernels/*.py, pyaxisymflow/elasto_kernels/*.py)
             and of the pystencils ENO3 stencils (pyaxisymflow/pyst_kernels).
stokes.py    the reference's dense eigen-decomposition Stokes solver, and an
             exact DCT-II + banded solve of the same discrete system.
lsq.c/.py    C restatement of the least-squares eta extrapolation
             (pyaxisymflow/core/src/extrapolate_using_least_squares.hpp).
drivers.py   the reference's time-loop step orders (examples/*).
Makefile     builds lsq.c -> _lsq.so and, where /root/reference exists, the
             reference's own header-only extrapolation into _ref/.

How each function is pinned (closed forms, invariants, reference KATs,
the compiled reference) is listed in tests/test_oracle_*.py; a function
with no independent check is marked "parity unpinned" in its docstring.
"""
