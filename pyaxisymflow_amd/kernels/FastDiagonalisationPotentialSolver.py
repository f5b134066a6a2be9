"""Drop-in for pyaxisymflow/kernels/FastDiagonalisationPotentialSolver.py:6-145."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib

_BC = {"homogenous_neumann_along_z_and_r": 0}


class FastDiagonalisationPotentialSolver:
    """Solves phi_rr + phi_r / r + phi_zz = rhs on the cell-centred (nr, nz)
    grid with the reference's discrete operators
    (FastDiagonalisationPotentialSolver.py:41-86, Neumann along z and r).

    The z-operator is diagonalised by a DCT-II (its exact eigenbasis) and each
    z-mode's tridiagonal radial operator is solved directly.  Constants are in
    the operator's null space: the reference divides mode (0, 0) by a
    round-off eigenvalue (:114-117), so its phi carries an arbitrary (for an
    incompatible rhs, huge) constant.  Here that mode is solved in the gauge
    where the last radial coefficient of z-mode 0 is zero: phi equals the
    reference's up to an additive constant when the rhs is compatible, and
    velocities (the gradient) agree.
    """

    def __init__(self, grid_size_r, grid_size_z, dx, real_dtype=np.float64,
                 bc_type="homogenous_neumann_along_z_and_r"):
        if bc_type not in _BC:
            raise NotImplementedError(f"bc_type {bc_type!r} is not supported on the GPU path")
        if real_dtype not in (np.float64, float):
            raise NotImplementedError("the drop-in solver is FP64")
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.real_dtype = real_dtype
        self.bc_type = bc_type
        # FastDiagonalisationPotentialSolver.py:20-24
        self.radial_coord = np.linspace(
            dx / 2, grid_size_r * dx - dx / 2, grid_size_r
        ).reshape(grid_size_r, 1)
        _lib.device()
        plan = ctypes.c_void_p()
        _lib.call("axs_potential_plan_create", int(grid_size_r), int(grid_size_z), float(dx),
                  _BC[bc_type], ctypes.byref(plan))
        self._plan = plan

    def solve(self, solution_field, rhs_field):
        """phi with phi_rr + phi_r / r + phi_zz = rhs
        (FastDiagonalisationPotentialSolver.py:119-145), written in place."""
        st = _lib.Staged()
        nr, nz = _lib.same_shape(solution_field, rhs_field)
        if (nr, nz) != (self.grid_size_r, self.grid_size_z):
            raise ValueError(f"field shape {(nr, nz)} != {(self.grid_size_r, self.grid_size_z)}")
        prhs = st.ptr(rhs_field, "rhs_field")
        pphi = st.ptr(solution_field, "solution_field", out=True, copy_in=False)
        _lib.call("axs_potential_solve_f64", self._plan, pphi, prhs, _lib.stream_handle())
        st.finish()

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().axs_potential_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None
