"""Drop-in for pyaxisymflow/kernels/implicit_diffusion_solver.py:6-143."""
import ctypes

import numpy as np

from pyaxisymflow_amd import _lib


class ImplicitEulerDiffusionStepper:
    """One implicit Euler step of axisymmetric vorticity diffusion,
    (I - nu dt L) w' = w with the reference's discrete operators
    (implicit_diffusion_solver.py:40-72: Dirichlet [1, -2, 1] / dx^2 along z,
    P_r + D_r / r - I / r^2 along r, zero ghost values at every edge).

    Instead of the reference's dense eigen-decompositions (four dense
    products per step), the z-operator is diagonalised by its exact
    eigenbasis, the orthonormal DST-I, and each z-mode's radial operator,
    which is tridiagonal, is solved directly: the same linear system and no
    O(n^3) set-up.  Tables and the spectral buffer live on the device.
    """

    def __init__(
        self,
        time_step,
        kinematic_viscosity,
        grid_size_r,
        grid_size_z,
        dx,
        real_dtype=np.float64,
    ):
        if real_dtype not in (np.float64, float):
            raise NotImplementedError("the drop-in diffusion stepper is FP64")
        self.time_step = time_step
        self.nu_times_dt = self.time_step * kinematic_viscosity
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.real_dtype = real_dtype
        # implicit_diffusion_solver.py:22-26
        self.radial_coord = np.linspace(
            dx / 2, grid_size_r * dx - dx / 2, grid_size_r
        ).reshape(grid_size_r, 1)
        _lib.device()
        plan = ctypes.c_void_p()
        _lib.call("axs_implicit_diffusion_plan_create", int(grid_size_r), int(grid_size_z),
                  float(dx), float(self.nu_times_dt), ctypes.byref(plan))
        self._plan = plan

    def step(self, vorticity_field, dt):
        """Diffuses vorticity_field in place over dt
        (implicit_diffusion_solver.py:108-143)."""
        # :113-120: the matrices were assembled for one dt
        if dt != self.time_step:
            raise ValueError(
                "dt should be constant throughout the simulation if using "
                "implicit diffusion! Please use the value of dt being used "
                "to initialize the implicit diffusion stepper"
            )
        st = _lib.Staged()
        nr, nz = _lib.same_shape(vorticity_field)
        if (nr, nz) != (self.grid_size_r, self.grid_size_z):
            raise ValueError(f"field shape {(nr, nz)} != {(self.grid_size_r, self.grid_size_z)}")
        pw = st.ptr(vorticity_field, "vorticity_field", out=True)
        _lib.call("axs_implicit_diffusion_step_f64", self._plan, pw, _lib.stream_handle())
        st.finish()

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().axs_implicit_diffusion_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None
