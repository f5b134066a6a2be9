"""Drop-in for pyaxisymflow/kernels/FastDiagonalisationStokesSolver.py:6-156."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib

_BC = {
    "homogenous_neumann_along_z_and_r": 0,
    "homogenous_neumann_along_r_and_periodic_along_z": 1,
    "homogenous_dirichlet_along_r_and_periodic_along_z": 2,
}


class FastDiagonalisationStokesSolver:
    """Solves -psi_rr + psi_r / r - psi_zz = r * omega on the cell-centred
    (nr, nz) grid with the reference's discrete operators
    (FastDiagonalisationStokesSolver.py:41-97).

    Instead of the reference's dense eigen-decomposition (four dense GEMMs
    per solve), the z-operator is diagonalised by its exact eigenbasis (a
    DCT-II for Neumann z, a real DFT for the periodic bc_types of :88-97) and
    each z-mode's radial operator, which is tridiagonal, is solved directly:
    same linear system, O(nr nz log nz) work, and no O(n^3) set-up.  Tables
    and the spectral buffer live on the device.
    """

    def __init__(
        self,
        grid_size_r,
        grid_size_z,
        dx,
        real_dtype=np.float64,
        bc_type="homogenous_neumann_along_z_and_r",
    ):
        if bc_type not in _BC:
            raise NotImplementedError(f"bc_type {bc_type!r} is not supported on the GPU path")
        if real_dtype not in (np.float64, float):
            raise NotImplementedError("the drop-in solver is FP64")
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.real_dtype = real_dtype
        self.bc_type = bc_type
        # FastDiagonalisationStokesSolver.py:20-24
        self.radial_coord = np.linspace(
            dx / 2, grid_size_r * dx - dx / 2, grid_size_r
        ).reshape(grid_size_r, 1)
        _lib.device()
        plan = ctypes.c_void_p()
        _lib.call("axs_stokes_plan_create", int(grid_size_r), int(grid_size_z), float(dx),
                  _BC[bc_type], ctypes.byref(plan))
        self._plan = plan

    def solve(self, solution_field, rhs_field):
        """psi = solution of the pseudo-Poisson problem with rhs r * omega
        (FastDiagonalisationStokesSolver.py:130-156), written in place."""
        st = _lib.Staged()
        nr, nz = _lib.same_shape(solution_field, rhs_field)
        if (nr, nz) != (self.grid_size_r, self.grid_size_z):
            raise ValueError(f"field shape {(nr, nz)} != {(self.grid_size_r, self.grid_size_z)}")
        prhs = st.ptr(rhs_field, "rhs_field")
        ppsi = st.ptr(solution_field, "solution_field", out=True, copy_in=False)
        _lib.call("axs_stokes_solve_f64", self._plan, ppsi, prhs, _lib.stream_handle())
        st.finish()

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().axs_stokes_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None
