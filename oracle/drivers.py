"""The reference's per-timestep call sequences (row A16), restated on the
oracle kernels.  TEST ORACLE ONLY.

* ``RigidBodyCase`` follows examples/FlowPastSphere/flow_past_sphere.py:39-188
  (static body, free-stream ramp, drag coefficient) and, with an oscillating
  body, examples/SpheroidStreaming/spheroid_streaming.py:129-246.
* ``SoftSphereCase`` follows
  examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:27-215
  and ``SoftSphereStreamingCase``
  examples/SoftSphereStreaming/soft_sphere_streaming.py:28-264.
* ``PeriodicRigidSlabCase`` follows
  examples/PeriodicRigidSlab/periodic_rigid_slab.py:33-158 (z-periodic,
  oscillating annular wall), and ``womersley_wall_driven`` restates its
  closed-form theory (theory_rigid_slab.py:9-23).
* ``PeriodicSoftSlabCase`` follows
  examples/PeriodicSoftSlab/periodic_soft_slab.py:43-404 (an elastic layer
  around the axis, the static PDE extrapolation of oracle.pde_extrapolation),
  and ``soft_slab_theory`` restates theory_soft_slab.py:15-87.
* Synthetic inputs (seeded Gaussian vorticity, sphere masks) follow the
  workload definitions used by the tests and the benchmark.
"""
import numpy as np

from oracle import kernels as K
from oracle.implicit_diffusion import ImplicitEulerDiffusionStepper, exact_implicit_diffusion
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
                 solver="eig", oscillation=None, implicit_diffusion=False):
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
        # flow_past_sphere.py:90-98: implicit Euler diffusion at the fixed
        # step dx^2 / 4 / nu (eig route, or the exact route with solver="exact")
        self.implicit_diffusion = implicit_diffusion
        if implicit_diffusion:
            self.diffusion_dt_limit = self.dx**2 / 4 / nu
            if solver == "eig":
                self.implicit_stepper = ImplicitEulerDiffusionStepper(
                    time_step=self.diffusion_dt_limit, kinematic_viscosity=nu,
                    grid_size_r=nr, grid_size_z=nz, dx=self.dx)

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
        if self.implicit_diffusion:  # :146-148
            dt = self.diffusion_dt_limit
        else:
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
        if not self.implicit_diffusion:
            K.diffusion_RK2_unb(self.vorticity, self.temp_vorticity, self.R, self.nu, dt, dx)
        elif self.solver_kind == "eig":  # :179-181
            self.implicit_stepper.step(vorticity_field=self.vorticity, dt=dt)
        else:
            self.vorticity[...] = exact_implicit_diffusion(self.vorticity, self.nu * dt, dx)
        self.t += dt
        self.it += 1
        self.dts.append(dt)
        return dt


class OscillatingSpheroidCase:
    """examples/SpheroidStreaming/spheroid_streaming.py:29-246 (explicit
    diffusion path): a spheroid oscillating along z in quiescent fluid,
    mask rebuilt every step, dt clipped to the oscillation period and to
    tEnd, psi time-averaged over each period.  Defaults are the reference's
    (womersley_square = 100, radius 0.1, spheroid_AR = 1, freq = 16)."""

    def __init__(self, grid_size_z, womersley_square=100.0, radius=0.1, spheroid_AR=1.0,
                 freq=16, domain_AR=0.5, no_cycles=25, solver="eig", brink_lam=1e4):
        # :40-68
        dx = 1.0 / grid_size_z
        grid_size_r = int(domain_AR * grid_size_z)
        self.dx, self.nz, self.nr = dx, grid_size_z, grid_size_r
        self.CFL = 0.1
        self.eps = np.finfo(float).eps
        self.brink_lam = brink_lam
        self.moll_zone = dx * 2**0.5
        self.radius_along = radius
        self.radius_normal = radius / spheroid_AR
        rad_eq = (self.radius_normal**2 * self.radius_along) ** (1 / 3)
        self.length_scale = spheroid_AR * rad_eq
        self.spheroid_AR = spheroid_AR
        self.omega = 2 * np.pi * freq
        nond_AC = 1.0 / np.sqrt(womersley_square)
        self.e = 0.1
        self.U_0 = self.e * self.length_scale * self.omega
        Rs = (self.e / nond_AC) ** 2
        self.nu = self.e * self.U_0 * self.length_scale / Rs
        self.tEnd = no_cycles / freq
        # :71-89
        self.z = np.linspace(0 + dx / 2, 1 - dx / 2, grid_size_z)
        self.r = np.linspace(0 + dx / 2, domain_AR - dx / 2, grid_size_r)
        self.Z, self.R = np.meshgrid(self.z, self.r)
        Z = self.Z
        self.vorticity = 0 * Z
        self.penal_vorticity = 0 * Z
        self.temp_vorticity = 0 * Z
        self.psi = 0 * Z
        self.u_z = 0 * Z
        self.u_r = 0 * Z
        self.avg_psi = 0 * Z
        self.Z_cm, self.R_cm = 0.5, 0.0
        self.t = 0
        self.it = 0
        self.freqTimer = 0.0
        # :104-112
        self.freqTimer_limit = 1 / freq
        diffusion_dt_limit = 0.9 * dx**2 / 4 / self.nu
        num_diffusion_substeps = int(self.freqTimer_limit / diffusion_dt_limit)
        self.diffusion_dt_limit = self.freqTimer_limit / num_diffusion_substeps
        self.solver_kind = solver
        if solver == "eig":
            self.solver = FastDiagonalisationStokesSolver(grid_size_r, grid_size_z, dx)
        self.dts = []
        self.force = []
        self.char_func = 0 * Z

    def step(self):
        """spheroid_streaming.py:129-246, one iteration (plus the Brinkman
        force sum of compute_forces.py:14 on the penalised velocity)."""
        dx, Z, R = self.dx, self.Z, self.R
        K.kill_boundary_vorticity_sine_z(self.vorticity, Z, 3, dx)
        K.kill_boundary_vorticity_sine_r(self.vorticity, R, 3, dx)
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi, rhs_field=self.vorticity)
        else:
            self.psi[...] = exact_stokes_solve(self.vorticity, dx)
        K.compute_velocity_from_psi_unb(self.u_z, self.u_r, self.psi, R, dx)
        if np.isclose(self.freqTimer, self.freqTimer_limit):
            self.freqTimer = 0.0
            self.avg_psi[...] = 0 * Z
        dt = min(
            0.9 * dx**2 / 4 / self.nu,
            self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps),
            0.01 * self.freqTimer_limit,
        )
        if self.freqTimer + dt > self.freqTimer_limit:
            dt = self.freqTimer_limit - self.freqTimer
        if self.t + dt > self.tEnd:
            dt = self.tEnd - self.t
        self.avg_psi[...] += self.psi * dt
        Z_cm_t = self.Z_cm + self.e * self.length_scale * np.sin(self.omega * self.t)
        phi = self.radius_normal - np.sqrt(
            (Z - Z_cm_t) ** 2 / self.spheroid_AR**2 + (R - self.R_cm) ** 2
        )
        self.char_func = 0 * Z
        K.smooth_Heaviside(self.char_func, phi, self.moll_zone)
        u_z_upen = self.u_z.copy()
        u_r_upen = self.u_r.copy()
        U_pen = self.U_0 * np.cos(self.omega * self.t)
        K.brinkmann_penalize(self.brink_lam, dt, self.char_func, U_pen, 0.0,
                             u_z_upen, u_r_upen, self.u_z, self.u_r)
        K.compute_vorticity_from_velocity_unb(
            self.penal_vorticity, -u_z_upen + self.u_z, -u_r_upen + self.u_r, dx
        )
        self.vorticity[...] += self.penal_vorticity
        # compute_forces.py:14 with rho_f = 1, times 2 pi dx^2 (falling_rigid_sphere.py:155)
        self.force.append(2 * np.pi * dx**2 * 1.0 * self.brink_lam
                          * np.sum(R * self.char_func * (self.u_z - U_pen)))
        K.advect_vorticity_via_eno3(self.vorticity, self.u_z, self.u_r, dt, dx)
        K.diffusion_RK2_unb(self.vorticity, self.temp_vorticity, R, self.nu, dt, dx)
        self.t += dt
        self.it += 1
        self.freqTimer = self.freqTimer + dt
        self.dts.append(dt)
        return dt


class SoftSphereCase:
    """examples/SoftSphereInTaylorGreenVortex/soft_sphere_in_taylor_green_vortex.py:27-215:
    an elastic sphere (reference map eta, neo-Hookean-like stress with
    shear modulus G) in an axisymmetric Taylor-Green-like vortex, no
    Brinkman term.  scikit-fmm's ``distance`` is replaced by
    oracle.reinit.distance (its restatement; parity with scikit-fmm itself
    is unpinned)."""

    def __init__(self, grid_size_z=256, G=1.0, nu=1e-3, rho_f=1.0, r_ball=0.15, domain_AR=0.5,
                 solver="eig", C=5e-2):
        from oracle.lsq import extrapolate_using_least_squares_till_first_order  # noqa: F401

        dx = 1.0 / grid_size_z
        grid_size_r = int(domain_AR * grid_size_z)
        self.dx, self.nz, self.nr = dx, grid_size_z, grid_size_r
        self.CFL = 0.1
        self.eps = np.finfo(float).eps
        # :36-45
        self.moll_zone = dx * 4
        self.extrap_zone = self.moll_zone + 4 * dx
        self.reinit_band = self.extrap_zone
        self.r_ball = r_ball
        self.rho_f = rho_f
        self.nu = nu
        self.G = G
        # :48-98
        self.z = np.linspace(0 + dx / 2, 1 - dx / 2, grid_size_z)
        self.r = np.linspace(0 + dx / 2, domain_AR - dx / 2, grid_size_r)
        self.Z, self.R = np.meshgrid(self.z, self.r)
        Z, R = self.Z, self.R
        self.Z_cm, self.R_cm = 0.5, 0.0
        self.ball_phi = -np.sqrt((Z - self.Z_cm) ** 2 + (R - self.R_cm) ** 2) + r_ball
        self.ball_char_func = 0 * Z
        K.smooth_Heaviside(self.ball_char_func, self.ball_phi, self.moll_zone)
        self.inside_solid = self.ball_char_func > 0.5
        self.vorticity = 0 * Z
        self.temp_vorticity = 0 * Z
        self.psi = 0 * Z
        self.u_z = 0 * Z
        self.u_r = 0 * Z
        Z_double, R_double = np.meshgrid(self.z, self.z)
        self.eta1 = Z.copy()
        self.eta2 = R.copy()
        self.eta1_double = Z_double.copy()
        self.eta2_double = R_double.copy()
        self.ball_phi_double = 0 * Z_double
        kz = 2 * np.pi
        kr = kz
        self.vorticity[...] = C * (kz**2 + kr**2) * np.sin(kz * Z) * np.sin(kr * R)
        self.t = 0
        self.it = 0
        self.bad_phi = 0 * Z
        self.phi_orig = 0 * Z
        self.sigma = [0 * Z for _ in range(3)]
        self.eta_grad = [0 * Z for _ in range(4)]  # eta1z, eta1r, eta2z, eta2r
        self.tau_z = 0 * Z
        self.tau_r = 0 * Z
        self.solver_kind = solver
        if solver == "eig":
            self.solver = FastDiagonalisationStokesSolver(grid_size_r, grid_size_z, dx)
        self.dts = []
        self.sweeps = []

    def step(self):
        """soft_sphere_in_taylor_green_vortex.py:101-213, one iteration."""
        from oracle.reinit import distance

        dx, Z, R = self.dx, self.Z, self.R
        K.kill_boundary_vorticity_sine_z(self.vorticity, Z, 3, dx)
        K.kill_boundary_vorticity_sine_r(self.vorticity, R, 3, dx)
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi, rhs_field=self.vorticity)
        else:
            self.psi[...] = exact_stokes_solve(self.vorticity, dx)
        K.compute_velocity_from_psi_unb(self.u_z, self.u_r, self.psi, R, dx)
        dt = min(
            self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps),
            self.CFL * dx / np.sqrt(self.G / self.rho_f),
            0.9 * dx**2 / 4 / self.nu,
        )
        K.advect_refmap_via_eno3(self.eta1, self.eta2, self.u_z, self.u_r, dt, dx)
        self.phi_orig[...] = (
            -np.sqrt((self.eta1 - self.Z_cm) ** 2 + (self.eta2 - self.R_cm) ** 2) + self.r_ball
        )
        band = np.where(self.ball_phi > -3 * dx)
        self.ball_phi[band] = self.phi_orig[band]
        self.bad_phi[...] = self.ball_phi
        d, masked = distance(self.ball_phi, dx, self.reinit_band)
        self.ball_phi = np.where(masked, self.bad_phi, d)
        self.ball_char_func *= 0
        K.smooth_Heaviside(self.ball_char_func, self.ball_phi, self.moll_zone)
        self.inside_solid[...] = self.ball_char_func > 0.5
        K.extrapolate_eta_with_least_squares(
            self.inside_solid, self.ball_phi, self.eta1, self.eta2, self.ball_phi_double,
            self.eta1_double, self.eta2_double, self.extrap_zone, self.nr, self.z,
        )
        s11, s12, s22 = self.sigma
        K.solid_sigma(s11, s12, s22, self.G, dx, self.eta1, self.eta2, *self.eta_grad)
        s11[...] = self.ball_char_func * s11
        s12[...] = self.ball_char_func * s12
        s22[...] = self.ball_char_func * s22
        K.update_vorticity_from_solid_stress(
            self.vorticity, self.tau_z, self.tau_r, s11, s12, s22, R, dt, dx
        )
        K.advect_vorticity_via_eno3(self.vorticity, self.u_z, self.u_r, dt, dx)
        K.diffusion_RK2_unb(self.vorticity, self.temp_vorticity, R, self.nu, dt, dx)
        self.t += dt
        self.it += 1
        self.dts.append(dt)
        return dt


class SoftSphereStreamingCase(SoftSphereCase):
    """examples/SoftSphereStreaming/soft_sphere_streaming.py:28-264: an
    elastic sphere around a rigid core (radius zeta r_ball) tethered to
    oscillate along z at U_0 cos(omega t), U_0 = e r_ball omega, in fluid at
    rest.  Against the Taylor-Green loop the vorticity is advected right
    after the reinit, the core is penalised after the stress, and dt is
    clipped to the end of each cycle and to tEnd.  Period averages and
    plots are left out."""

    def __init__(self, grid_size_z=256, domain_AR=0.5, r_ball=0.15, freq=16, nond_AC=0.125,
                 e=0.1, Cauchy=0.1, zeta=0.25, no_cycles=30, rho_f=1.0, brink_lam=1e8,
                 solver="eig"):
        omega = 2 * np.pi * freq
        U_0 = e * r_ball * omega
        Rs = (e / nond_AC) ** 2
        nu = e * U_0 * r_ball / Rs
        G = e * rho_f * (r_ball * omega) ** 2 / Cauchy
        super().__init__(grid_size_z=grid_size_z, G=G, nu=nu, rho_f=rho_f, r_ball=r_ball,
                         domain_AR=domain_AR, solver=solver, C=0.0)
        dx = self.dx
        # :40-46 (moll_zone 2 dx here)
        self.moll_zone = dx * 2
        self.extrap_zone = self.moll_zone + 4 * dx
        self.reinit_band = self.extrap_zone
        Z, R = self.Z, self.R
        self.ball_phi = -np.sqrt((Z - self.Z_cm) ** 2 + (R - self.R_cm) ** 2) + r_ball
        self.ball_char_func = 0 * Z
        K.smooth_Heaviside(self.ball_char_func, self.ball_phi, self.moll_zone)
        self.inside_solid = self.ball_char_func > 0.5
        self.vorticity = 0 * Z
        self.brink_lam = brink_lam
        self.U_0, self.omega, self.e = U_0, omega, e
        self.fixed_rad = zeta * r_ball
        self.freqTimer_limit = 1 / freq
        self.tEnd = no_cycles / freq
        self.freqTimer = 0.0
        self.tether_phi = 0 * Z
        self.tether_char_func = 0 * Z
        self.penal_vorticity = 0 * Z

    def step(self):
        """soft_sphere_streaming.py:128-264, one iteration."""
        from oracle.reinit import distance

        dx, Z, R = self.dx, self.Z, self.R
        K.kill_boundary_vorticity_sine_z(self.vorticity, Z, 3, dx)
        K.kill_boundary_vorticity_sine_r(self.vorticity, R, 3, dx)
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi, rhs_field=self.vorticity)
        else:
            self.psi[...] = exact_stokes_solve(self.vorticity, dx)
        K.compute_velocity_from_psi_unb(self.u_z, self.u_r, self.psi, R, dx)
        if self.freqTimer >= self.freqTimer_limit:
            self.freqTimer = 0.0
        dt = min(
            self.CFL * dx / np.sqrt(self.G / self.rho_f),
            self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps),
            0.9 * dx**2 / 4 / self.nu,
        )
        if self.freqTimer + dt > self.freqTimer_limit:
            dt = self.freqTimer_limit - self.freqTimer
        if self.t + dt > self.tEnd:
            dt = self.tEnd - self.t
        K.advect_refmap_via_eno3(self.eta1, self.eta2, self.u_z, self.u_r, dt, dx)
        self.phi_orig[...] = (
            -np.sqrt((self.eta1 - self.Z_cm) ** 2 + (self.eta2 - self.R_cm) ** 2) + self.r_ball
        )
        band = np.where(self.ball_phi > -3 * dx)
        self.ball_phi[band] = self.phi_orig[band]
        self.bad_phi[...] = self.ball_phi
        d, masked = distance(self.ball_phi, dx, self.reinit_band)
        self.ball_phi = np.where(masked, self.bad_phi, d)
        K.advect_vorticity_via_eno3(self.vorticity, self.u_z, self.u_r, dt, dx)
        K.smooth_Heaviside(self.ball_char_func, self.ball_phi, self.moll_zone)
        self.inside_solid[...] = self.ball_char_func > 0.5
        K.extrapolate_eta_with_least_squares(
            self.inside_solid, self.ball_phi, self.eta1, self.eta2, self.ball_phi_double,
            self.eta1_double, self.eta2_double, self.extrap_zone, self.nr, self.z,
        )
        s11, s12, s22 = self.sigma
        K.solid_sigma(s11, s12, s22, self.G, dx, self.eta1, self.eta2, *self.eta_grad)
        s11[...] = self.ball_char_func * s11
        s12[...] = self.ball_char_func * s12
        s22[...] = self.ball_char_func * s22
        K.update_vorticity_from_solid_stress(
            self.vorticity, self.tau_z, self.tau_r, s11, s12, s22, R, dt, dx
        )
        Z_cm_t = self.Z_cm + self.e * self.r_ball * np.sin(self.omega * self.t)
        self.tether_phi[...] = -np.sqrt((Z - Z_cm_t) ** 2 + (R - self.R_cm) ** 2) + \
            self.fixed_rad
        K.smooth_Heaviside(self.tether_char_func, self.tether_phi, self.moll_zone)
        u_z_upen = self.u_z.copy()
        u_r_upen = self.u_r.copy()
        K.brinkmann_penalize(self.brink_lam, dt, self.tether_char_func,
                             self.U_0 * np.cos(self.omega * self.t), 0.0, u_z_upen, u_r_upen,
                             self.u_z, self.u_r)
        K.compute_vorticity_from_velocity_unb(self.penal_vorticity, self.u_z - u_z_upen,
                                              self.u_r - u_r_upen, dx)
        self.vorticity[...] += self.penal_vorticity
        K.diffusion_RK2_unb(self.vorticity, self.temp_vorticity, R, self.nu, dt, dx)
        self.t += dt
        self.freqTimer += dt
        self.it += 1
        self.dts.append(dt)
        return dt


class PeriodicRigidSlabCase:
    """examples/PeriodicRigidSlab/periodic_rigid_slab.py:33-158: a rigid
    annular wall (|r - R_wall_center| < wall_thickness, smoothed over
    dx sqrt(2)) oscillating along z at U_0 cos(omega t) in a z-periodic
    domain of grid_size_z columns, the first and last ghost_size of them
    ghost copies; the periodic Stokes solve (Neumann r) runs on the
    interior columns.  ``solver`` "exact" (DFT + banded radial solves) or
    "eig" (the reference's dense route)."""

    def __init__(self, grid_size_z=16, max_z=0.05, max_r=0.5, brink_lam=1e4,
                 wall_thickness=0.05, R_wall_center=0.3, Re=25, U_0=1.0, freq=8.0,
                 ghost_size=2, CFL=0.1, solver="exact"):
        from oracle import periodic as P
        from oracle.stokes import NEUMANN_R_PERIODIC_Z

        self.P = P
        self.bc = NEUMANN_R_PERIODIC_Z
        domain_AR = max_r / max_z
        nz, nr = grid_size_z, int(domain_AR * grid_size_z)
        dx = max_r / nr
        self.nz, self.nr, self.dx = nz, nr, dx
        self.z = np.linspace(0 + dx / 2, max_z - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, max_r - dx / 2, nr)
        self.Z, self.R = np.meshgrid(self.z, self.r)
        self.brink_lam = brink_lam
        self.moll_zone = dx * 2**0.5
        self.R_tube = R_wall_center - wall_thickness
        self.U_0 = U_0
        self.nu = U_0 * 2 * self.R_tube / Re
        self.omega = 2 * np.pi * freq
        self.g = ghost_size
        self.CFL = CFL
        self.eps = np.finfo(float).eps
        self.comm = P.gen_periodic_boundary_ghost_comm(ghost_size)
        z0 = 0 * self.Z
        self.vorticity, self.penal_vorticity, self.temp_vorticity = z0.copy(), z0.copy(), z0.copy()
        self.psi, self.u_z, self.u_r = z0.copy(), z0.copy(), z0.copy()
        self.u_z_upen, self.u_r_upen = z0.copy(), z0.copy()
        phi0 = -np.sqrt((self.R - R_wall_center) ** 2) + wall_thickness
        self.char_func = 0 * self.Z
        K.smooth_Heaviside(self.char_func, phi0, self.moll_zone)
        self.psi_inner = self.psi[..., ghost_size:-ghost_size].copy()
        self.solver_kind = solver
        if solver == "eig":
            self.solver = FastDiagonalisationStokesSolver(nr, nz - 2 * ghost_size, dx,
                                                          bc_type=self.bc)
        self.t = 0.0
        self.it = 0
        self.dts = []

    def step(self):
        """periodic_rigid_slab.py:93-158, one iteration."""
        P, g, dx = self.P, self.g, self.dx
        K.kill_boundary_vorticity_sine_r(self.vorticity, self.R, 3, dx)
        self.psi_inner[...] = self.psi[..., g:-g]
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi_inner,
                              rhs_field=self.vorticity[:, g:-g])
        else:
            self.psi_inner[...] = exact_stokes_solve(self.vorticity[:, g:-g], dx, self.bc)
        self.psi[..., g:-g] = self.psi_inner
        P.compute_velocity_from_psi_periodic(self.u_z, self.u_r, self.psi, self.R, dx, self.comm)
        dt = min(
            0.9 * dx**2 / 4 / self.nu,
            self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps),
        )
        self.u_z_upen[...] = self.u_z.copy()
        self.u_r_upen[...] = self.u_r.copy()
        K.brinkmann_penalize(self.brink_lam, dt, self.char_func,
                             self.U_0 * np.cos(self.omega * self.t), 0.0, self.u_z_upen,
                             self.u_r_upen, self.u_z, self.u_r)
        P.compute_vorticity_from_velocity_periodic(self.penal_vorticity,
                                                   self.u_z - self.u_z_upen,
                                                   self.u_r - self.u_r_upen, dx, self.comm)
        self.vorticity[...] += self.penal_vorticity
        P.advect_vorticity_via_eno3_periodic(self.vorticity, self.u_z, self.u_r, dt, dx,
                                             self.comm)
        P.diffusion_RK2_periodic(self.vorticity, self.temp_vorticity, self.R, self.nu, dt, dx,
                                 self.comm)
        self.t += dt
        self.it += 1
        self.dts.append(dt)
        return dt


def womersley_wall_driven(r, t, U_0, omega, nu, R_tube):
    """theory_rigid_slab.py:9-23: axial velocity in a tube of radius R_tube
    whose wall moves at U_0 cos(omega t), u = U_0 Re(2 A J0(lam r) e^{-i omega t})
    with lam = sqrt(i omega / nu) and A = 1 / (2 J0(lam R_tube))."""
    from scipy.special import jv

    lam = np.sqrt(1j * (omega / nu))
    A = 0.5 / jv(0, lam * R_tube)
    return U_0 * np.real(2 * (A * jv(0, lam * r)) * np.exp(-1j * omega * t))


class TorusStreamingCase:
    """examples/TorusStreaming/torus_streaming.py:22-233 with implicit
    diffusion (the driver's setting): a torus whose cross-section centre
    oscillates in r as a sum of modes; the dilation's source flow from the
    velocity potential, the surface velocity by Brinkman penalisation.
    ``solver`` "exact" (DCT + banded radial solves for psi and phi, the
    potential's singular mode in the gauge phi_hat_0[nr-1] = 0, DST-I +
    banded for the implicit step) or "eig" (the reference's dense routes)."""

    def __init__(self, grid_size_z=256, domain_AR=0.5, solver="exact"):
        from oracle.potential import FastDiagonalisationPotentialSolver, exact_potential_solve
        from oracle.potential import compute_velocity_from_phi_unb

        self._exact_potential = exact_potential_solve
        self._vel_phi = compute_velocity_from_phi_unb
        dx = 1.0 / grid_size_z
        nz, nr = grid_size_z, int(domain_AR * grid_size_z)
        self.nz, self.nr, self.dx = nz, nr, dx
        self.CFL = 0.1
        self.brink_lam = 1e4
        self.moll_zone = dx * 2**0.5
        r_core = 0.125
        self.a = 0.34146341463414637 * r_core
        self.AR = 0.4285714285714286
        self.e = 0.0675
        self.freq_modes = np.array([16, 32])
        self.amplitude_modes = np.array([1.0, 0.3])
        freq = self.freq_modes[np.argmax(self.amplitude_modes)]
        omega = 2 * np.pi * freq
        self.omega_modes = 2 * np.pi * self.freq_modes
        self.nu = omega * (0.11235582096628798 * r_core) ** 2
        U_osc = self.e * self.a * np.sum(self.amplitude_modes * self.omega_modes)
        self.z = np.linspace(0 + dx / 2, 1 - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, domain_AR - dx / 2, nr)
        self.Z, self.R = np.meshgrid(self.z, self.r)
        z0 = 0 * self.Z
        self.vorticity, self.penal_vorticity = z0.copy(), z0.copy()
        self.psi, self.phi, self.u_z, self.u_r = z0.copy(), z0.copy(), z0.copy(), z0.copy()
        self.u_z_pen, self.u_r_pen = z0.copy(), z0.copy()
        self.vel_phi, self.u_z_divg, self.u_r_divg, self.vel_divg = (z0.copy(), z0.copy(),
                                                                     z0.copy(), z0.copy())
        self.Z_cm, self.R_cm = 0.5, r_core
        phi0 = -np.sqrt((self.Z - self.Z_cm) ** 2 + (self.R - self.R_cm) ** 2 / self.AR**2) + \
            self.a
        self.char_func = 0 * self.Z
        K.smooth_Heaviside(self.char_func, phi0, self.moll_zone)
        # :106-112, timestep aligned to the cycle
        freqTimer_limit = 1 / freq
        data_per_cycle = 100
        dataTimer_limit = 1 / freq / data_per_cycle
        diffusion_dt_limit = self.CFL / U_osc
        num_diffusion_substeps = max(int(freqTimer_limit / diffusion_dt_limit), data_per_cycle)
        diffusion_dt_limit = freqTimer_limit / num_diffusion_substeps
        self.dt_limit = min(diffusion_dt_limit, freqTimer_limit, dataTimer_limit)
        self.solver_kind = solver
        if solver == "eig":
            self.stokes = FastDiagonalisationStokesSolver(nr, nz, dx)
            self.potential = FastDiagonalisationPotentialSolver(nr, nz, dx)
            self.implicit = ImplicitEulerDiffusionStepper(
                time_step=self.dt_limit, kinematic_viscosity=self.nu, grid_size_r=nr,
                grid_size_z=nz, dx=dx)
        self.t = 0.0
        self.it = 0

    def step(self):
        """torus_streaming.py:129-233, one iteration (no plots, averages or
        data dumps)."""
        dx = self.dx
        K.kill_boundary_vorticity_sine_z(self.vorticity, self.Z, 3, dx)
        K.kill_boundary_vorticity_sine_r(self.vorticity, self.R, 3, dx)
        if self.solver_kind == "eig":
            self.stokes.solve(solution_field=self.psi, rhs_field=self.vorticity)
        else:
            self.psi[...] = exact_stokes_solve(self.vorticity, dx)
        K.compute_velocity_from_psi_unb(self.u_z, self.u_r, self.psi, self.R, dx)
        e, a, t = self.e, self.a, self.t
        R_cm_t = self.R_cm + e * a * np.sum(self.amplitude_modes * np.sin(self.omega_modes * t))
        self.phi[...] = a - np.sqrt((self.Z - self.Z_cm) ** 2 + (self.R - R_cm_t) ** 2 /
                                    self.AR**2)
        K.smooth_Heaviside(self.char_func, self.phi, self.moll_zone)
        self.vel_divg[...] = (
            e * a * np.sum(self.amplitude_modes * self.omega_modes *
                           np.cos(self.omega_modes * t)) / self.R)
        rhs = self.char_func * self.vel_divg
        if self.solver_kind == "eig":
            self.potential.solve(solution_field=self.vel_phi, rhs_field=rhs)
        else:
            self.vel_phi[...] = self._exact_potential(rhs, dx)
        self._vel_phi(self.u_z_divg, self.u_r_divg, self.vel_phi, dx)
        self.u_z[...] += self.u_z_divg
        self.u_r[...] += self.u_r_divg
        dt = self.dt_limit
        K.brinkmann_penalize(
            self.brink_lam, dt, self.char_func, 0.0,
            e * a * np.sum(self.amplitude_modes * self.omega_modes * np.cos(self.omega_modes * t)),
            self.u_z, self.u_r, self.u_z_pen, self.u_r_pen)
        K.compute_vorticity_from_velocity_unb(self.penal_vorticity, self.u_z_pen - self.u_z,
                                              self.u_r_pen - self.u_r, dx)
        self.vorticity[...] += self.penal_vorticity
        K.advect_vorticity_via_eno3(self.vorticity, self.u_z, self.u_r, dt, dx)
        if self.solver_kind == "eig":
            self.implicit.step(vorticity_field=self.vorticity, dt=dt)
        else:
            self.vorticity[...] = exact_implicit_diffusion(self.vorticity, self.nu * dt, dx)
        self.t += dt
        self.it += 1
        return dt


def soft_slab_theory(L_f, L_s, shear_rate, omega, G, V_wall, rho_f, nu_f, rho_ratio=1.0,
                     nu_ratio=1.0, resolution=None):
    """theory_soft_slab.py:15-87: the oscillatory axial velocity of a
    Kelvin-Voigt solid (r < L_s, shear modulus G, viscosity mu_s) inside a
    fluid layer (L_s < r < L_s + L_f) whose outer wall moves at
    V_wall cos(omega t).  Fluid velocity 2 Re((A J0(lam1 y) + B Y0(lam1 y))
    e^{-i omega t}), solid displacement Re-part of 2 C J0(lam2 y) e^{-i omega t}
    (velocity -i omega times it), lam1 = sqrt(i omega / nu_f), lam2 = omega /
    sqrt(-i omega nu_s + G / rho_s); A, B, C from the wall velocity, the
    velocity and the shear stress matched at y = L_s.  The reference solves
    the 3 x 3 system symbolically; here numerically (same equations).
    Returns (velocity(t) at the positions Y, Y, (fluid(y, t), solid(y, t),
    displacement(y, t))) with y = 1e-20 .. L_s + L_f (30 points) unless
    ``resolution`` gives them."""
    from scipy.special import jv, yv

    rho_s = rho_ratio * rho_f
    nu_s = nu_ratio * nu_f
    mu_f = nu_f * rho_f
    mu_s = nu_s * rho_s
    lam1 = np.sqrt(1j * omega / nu_f)
    lam2 = omega / np.sqrt(-1j * omega * nu_s + G / rho_s)
    L = L_s + L_f
    k1, k2 = jv(0, lam1 * L), yv(0, lam1 * L)
    k3, k4 = jv(0, lam1 * L_s), yv(0, lam1 * L_s)
    k5 = -1j * omega * jv(0, lam2 * L_s)
    k6, k7 = jv(1, lam1 * L_s), yv(1, lam1 * L_s)
    k8 = lam2 * jv(1, lam2 * L_s) / (mu_f * lam1) * (G - mu_s * 1j * omega)
    M = np.array([[k1, k2, 0.0], [k3, k4, -k5], [k6, k7, -k8]], dtype=complex)
    A, B, C = np.linalg.solve(M, np.array([V_wall / 2, 0.0, 0.0], dtype=complex))

    def fluid(y, t):
        return 2 * np.real((A * jv(0, lam1 * y) + B * yv(0, lam1 * y)) * np.exp(-1j * omega * t))

    def displacement(y, t):
        return 2 * np.real(C * jv(0, lam2 * y) * np.exp(-1j * omega * t))

    def solid(y, t):
        return 2 * np.real(-1j * omega * C * jv(0, lam2 * y) * np.exp(-1j * omega * t))

    Y = np.linspace(1e-20, L, 30) if resolution is None else np.array(resolution, dtype=float)

    def temporal(time):
        return (Y < L_s) * solid(Y, time) + (Y >= L_s) * fluid(Y, time)

    return temporal, Y, (fluid, solid, displacement)


class PeriodicSoftSlabCase:
    """examples/PeriodicSoftSlab/periodic_soft_slab.py:43-404: an elastic
    layer r < L_s (eta extrapolated by the static PDE scheme) inside fluid,
    driven by an annular wall at R_wall_center oscillating at
    V_wall cos(omega t), z-periodic with ghost_size ghost columns on each
    side; Re and Er fix nu_f and G (:80-102).  ``solver`` "exact" or "eig"
    for the interior Stokes solve.  ``samples`` collects (t, u_z at column
    nz / 2 for r < R_tube) where the driver compares with the theory
    (:229-238)."""

    def __init__(self, grid_size_r=256, Re=10.0, Er=0.25, domain_AR=32, zeta=1.0,
                 solver="exact", t_end=None):
        from oracle import periodic as P
        from oracle.pde_extrapolation import StaticPDEExtrapolation
        from oracle.stokes import NEUMANN_R_PERIODIC_Z

        self.P = P
        self.bc = NEUMANN_R_PERIODIC_Z
        max_r = 0.5
        max_z = max_r / domain_AR
        nz = int(grid_size_r / domain_AR)
        nr = grid_size_r
        dx = max_r / grid_size_r
        self.nr, self.nz, self.dx, self.max_z = nr, nz, dx, max_z
        self.z = np.linspace(0 + dx / 2, max_z - dx / 2, nz)
        self.r = np.linspace(0 + dx / 2, max_r - dx / 2, nr)
        self.Z, self.R = np.meshgrid(self.z, self.r)
        g = self.g = 2
        self.comm = P.gen_periodic_boundary_ghost_comm(g)
        self.comm_eta = P.gen_periodic_boundary_ghost_comm_eta(g, max_z, dx)
        self.CFL = 0.1
        self.eps = np.finfo(float).eps
        self.brink_lam = 1e4
        self.moll_zone = np.sqrt(2) * dx
        self.extrap_zone = self.moll_zone + 3 * dx
        self.extrap_tol = 1e-3
        wall_half = 0.05
        self.R_wall_center = 0.5 * max_r
        self.R_tube = self.R_wall_center - wall_half
        L = 2 * self.R_tube
        self.L_f = self.R_tube * zeta / (1 + zeta)
        self.L_s = self.R_tube / (1 + zeta)
        self.freq = 1.0
        self.freq_limit = 0.1 / self.freq
        self.omega = 2 * np.pi * self.freq
        self.freq_timer = self.freq_limit
        self.V_wall = 1.0
        self.rho_f = 1.0
        self.shear_rate = 2 * self.V_wall / self.omega / L
        self.nu = self.shear_rate * self.omega * self.L_f**2 / Re
        self.G = self.rho_f * self.nu * self.shear_rate * self.omega / Er
        self.t_end = 15 / self.freq if t_end is None else t_end
        self.solid_phi = self.L_s - self.R
        self.solid_char_func = 0 * self.Z
        K.smooth_Heaviside(self.solid_char_func, self.solid_phi, self.moll_zone)
        self.inside_solid = self.solid_char_func > 0.5
        wall_phi = -np.sqrt((self.R - self.R_wall_center) ** 2) + wall_half
        self.wall_char_func = 0 * self.Z
        K.smooth_Heaviside(self.wall_char_func, wall_phi, self.moll_zone)
        z0 = 0 * self.Z
        self.vorticity, self.penal_vorticity, self.temp_vorticity = z0.copy(), z0.copy(), z0.copy()
        self.psi, self.u_z, self.u_r = z0.copy(), z0.copy(), z0.copy()
        self.u_z_upen, self.u_r_upen = z0.copy(), z0.copy()
        self.eta1, self.eta2 = self.Z.copy(), self.R.copy()
        self.s11, self.s12, self.s22 = z0.copy(), z0.copy(), z0.copy()
        self.e1z, self.e2z, self.e1r, self.e2r = z0.copy(), z0.copy(), z0.copy(), z0.copy()
        self.tau_z, self.tau_r = z0.copy(), z0.copy()
        self.psi_inner = self.psi[..., g:-g].copy()
        self.solver_kind = solver
        if solver == "eig":
            self.solver = FastDiagonalisationStokesSolver(nr, nz - 2 * g, dx, bc_type=self.bc)
        self.extrapolator = StaticPDEExtrapolation(
            dx=dx, grid_size_r=nr, grid_size_z=nz, extrap_tol=self.extrap_tol,
            extrap_band=self.extrap_zone, periodic=True, per_communicator_gen=self.comm,
            per_communicator_eta=self.comm_eta)
        self.sample_rows = int(nr * self.R_tube / max_r)
        self.samples = []
        self.t = 0.0
        self.it = 0
        self.dts = []

    def step(self):
        """:214-396, one iteration."""
        P, g, dx = self.P, self.g, self.dx
        dt = min(self.CFL * dx / np.sqrt(self.G / self.rho_f), 0.9 * dx**2 / 4 / self.nu,
                 self.CFL * dx / (np.amax(np.fabs(self.u_z) + np.fabs(self.u_r)) + self.eps))
        if self.freq_timer + dt > self.freq_limit:
            dt = self.freq_limit - self.freq_timer
        if self.t + dt > self.t_end:
            dt = self.t_end - self.t
        K.kill_boundary_vorticity_sine_r(self.vorticity, self.R, 3, dx)
        self.psi_inner[...] = self.psi[..., g:-g]
        if self.solver_kind == "eig":
            self.solver.solve(solution_field=self.psi_inner, rhs_field=self.vorticity[:, g:-g])
        else:
            self.psi_inner[...] = exact_stokes_solve(self.vorticity[:, g:-g], dx, self.bc)
        self.psi[..., g:-g] = self.psi_inner
        P.compute_velocity_from_psi_periodic(self.u_z, self.u_r, self.psi, self.R, dx, self.comm)
        if self.freq_timer >= self.freq_limit:
            self.freq_timer = 0.0
            self.samples.append((self.t, self.u_z[:self.sample_rows, self.nz // 2].copy()))
        P.advect_refmap_via_eno3_periodic(self.eta1, self.eta2, self.u_z, self.u_r, dt, dx,
                                          self.comm, self.comm_eta)
        self.u_z_upen[...] = self.u_z
        self.u_r_upen[...] = self.u_r
        K.brinkmann_penalize(self.brink_lam, dt, self.wall_char_func,
                             self.V_wall * np.cos(self.omega * self.t), 0.0, self.u_z_upen,
                             self.u_r_upen, self.u_z, self.u_r)
        P.compute_vorticity_from_velocity_periodic(self.penal_vorticity, self.u_z - self.u_z_upen,
                                                   self.u_r - self.u_r_upen, dx, self.comm)
        self.vorticity[...] += self.penal_vorticity
        self.eta1[...] = self.inside_solid * self.eta1
        self.eta2[...] = self.inside_solid * self.eta2
        self.extrapolator.extrapolate(self.eta1, self.solid_phi)
        self.extrapolator.extrapolate(self.eta2, self.solid_phi)
        P.solid_sigma_periodic(self.s11, self.s12, self.s22, self.G, dx, self.eta1, self.eta2,
                               self.e1z, self.e1r, self.e2z, self.e2r, self.comm, self.comm_eta)
        self.s11[...] = self.solid_char_func * self.s11
        self.s12[...] = self.solid_char_func * self.s12
        self.s22[...] = self.solid_char_func * self.s22
        P.update_vorticity_from_solid_stress_periodic(self.vorticity, self.tau_z, self.tau_r,
                                                      self.s11, self.s12, self.s22, self.R, dt,
                                                      dx, self.comm)
        P.advect_vorticity_via_eno3_periodic(self.vorticity, self.u_z, self.u_r, dt, dx,
                                             self.comm)
        P.diffusion_RK2_periodic(self.vorticity, self.temp_vorticity, self.R, self.nu, dt, dx,
                                 self.comm)
        self.t += dt
        self.freq_timer += dt
        self.it += 1
        self.dts.append(dt)
        return dt
