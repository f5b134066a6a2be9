"""The reference's per-timestep call sequences (row A16), restated on the
oracle kernels.  TEST ORACLE ONLY.

* ``RigidBodyCase`` follows examples/FlowPastSphere/flow_past_sphere.py:39-188
  (static body, free-stream ramp, drag coefficient) and, with an oscillating
  body, examples/SpheroidStreaming/spheroid_streaming.py:129-246.
* Synthetic inputs (seeded Gaussian vorticity, sphere masks) follow the
  workload definitions used by the tests and the benchmark.
"""
import numpy as np

from oracle import kernels as K
from oracle.stokes import FastDiagonalisationStokesSolver, exact_stokes_solve


def grid(nz, nr, z_len, r_len=None):
    """Cell-centred grid as in flow_past_sphere.py:40-60 with dx = z_len/nz."""
    dx = z_len / nz
    if r_len is None:
        r_len = nr * dx
    z = np.linspace(0 + dx / 2, z_len - dx / 2, nz)
    r = np.linspace(0 + dx / 2, r_len - dx / 2, nr)
    Z, R = np.meshgrid(z, r)
    return dx, z, r, Z, R


def gaussian_vorticity(Z, R, n, seed, width_lo, width_hi, z_len, r_len):
    """Sum of n Gaussians with amplitudes N(0,1), widths U[lo,hi] and centres
    uniform in the domain (numpy.random.default_rng(seed))."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n)
    s = rng.uniform(width_lo, width_hi, n)
    cz = rng.uniform(0, z_len, n)
    cr = rng.uniform(0, r_len, n)
    w = np.zeros_like(Z)
    for i in range(n):
        w += a[i] * np.exp(-((Z - cz[i]) ** 2 + (R - cr[i]) ** 2) / (2 * s[i] ** 2))
    return w


def sphere_union_phi(Z, R, n, seed, rad_lo, rad_hi, zc_lo, zc_hi):
    """Level set (max over spheres) of n on-axis spheres."""
    rng = np.random.default_rng(seed)
    rad = rng.uniform(rad_lo, rad_hi, n)
    zc = rng.uniform(zc_lo, zc_hi, n)
    phi = np.full_like(Z, -np.inf)
    for i in range(n):
        phi = np.maximum(phi, rad[i] - np.sqrt((Z - zc[i]) ** 2 + R**2))
    return phi


class RigidBodyCase:
    """State and step of flow_past_sphere.py (static body, explicit RK2
    diffusion), generalised to the penalisation velocity U_z(t) of
    spheroid_streaming.py:207-226 when ``oscillation`` is given."""

    def __init__(self, nz, nr, z_len, nu, brink_lam, char_func, U_0=1.0,
                 T_ramp=None, r_sph=None, CFL=0.1, vorticity=None,
                 solver="eig", oscillation=None):
        self.dx, self.z, self.r, self.Z, self.R = grid(nz, nr, z_len)
        self.nz, self.nr = nz, nr
        self.nu = nu
        self.brink_lam = brink_lam
        self.char_func = char_func
        self.U_0 = U_0
        self.T_ramp = T_ramp
        self.r_sph = r_sph
        self.CFL = CFL
        self.oscillation = oscillation  # (U_amp, omega) or None
        self.eps = np.finfo(float).eps
        self.vorticity = np.zeros((nr, nz)) if vorticity is None else vorticity.copy()
        self.psi = np.zeros((nr, nz))
        self.u_z = np.zeros((nr, nz))
        self.u_r = np.zeros((nr, nz))
        self.penal_vorticity = np.zeros((nr, nz))
        self.temp_vorticity = np.zeros((nr, nz))
        self.t = 0.0
        self.it = 0
        self.drag = []
        self.dts = []
        self.solver_kind = solver
        if solver == "eig":
            self.solver = FastDiagonalisationStokesSolver(nr, nz, self.dx)

    def solve(self):
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi, rhs_field=self.vorticity)
        else:
            self.psi[...] = exact_stokes_solve(self.vorticity, self.dx)

    def step(self):
        """flow_past_sphere.py:107-188, one iteration."""
        dx = self.dx
        K.kill_boundary_vorticity_sine_z(self.vorticity, self.Z, 3, dx)
        K.kill_boundary_vorticity_sine_r(self.vorticity, self.R, 3, dx)
        self.solve()
        K.compute_velocity_from_psi_unb(self.u_z, self.u_r, self.psi, self.R, dx)
        if self.oscillation is None:
            prefac_x = 1.0
            if self.T_ramp is not None and self.t < self.T_ramp:
                prefac_x = np.sin(0.5 * np.pi * self.t / self.T_ramp)
            self.u_z[...] += self.U_0 * prefac_x
            U_pen = 0.0
        else:
            U_amp, omega = self.oscillation
            U_pen = U_amp * np.cos(omega * self.t)
        dt = min(
            0.9 * dx**2 / 4 / self.nu,
            self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps),
        )
        u_z_upen = self.u_z.copy()
        u_r_upen = self.u_r.copy()
        K.brinkmann_penalize(self.brink_lam, dt, self.char_func, U_pen, 0.0,
                             u_z_upen, u_r_upen, self.u_z, self.u_r)
        K.compute_vorticity_from_velocity_unb(
            self.penal_vorticity, self.u_z - u_z_upen, self.u_r - u_r_upen, dx
        )
        self.vorticity[...] += self.penal_vorticity
        if self.r_sph is not None:
            self.drag.append(K.drag_coefficient(self.R, self.char_func, self.u_z, dx,
                                                self.brink_lam, self.r_sph))
        K.advect_vorticity_via_eno3(self.vorticity, self.u_z, self.u_r, dt, dx)
        K.diffusion_RK2_unb(self.vorticity, self.temp_vorticity, self.R, self.nu, dt, dx)
        self.t += dt
        self.it += 1
        self.dts.append(dt)
        return dt
