"""Implicit Euler diffusion oracle (test oracle only).

Two CPU solvers of the same discrete system  (I - nu dt L) w' = w,  with
L w = A_r w + w P_z^T,  P_z = [1, -2, 1] / dx^2 (Dirichlet at both z ends),
A_r = P_r + D_r / r - I / r^2,  P_r = [1, -2, 1] / dx^2,  D_r = [-1, 0, 1] / (2 dx)
(every matrix truncated at the array ends: zero ghost values):

* ``ImplicitEulerDiffusionStepper`` restates the reference's dense
  eigen-decomposition route (kernels/implicit_diffusion_solver.py:6-143):
  scipy.sparse.diags -> dense, numpy.linalg.eig, descending sort,
  numpy.linalg.inv, numpy.linalg.multi_dot, the dt check of :115-120.
* ``exact_implicit_diffusion`` solves the same matrices by the DST-I along z
  (the eigenvectors of the Dirichlet P_z in closed form, scipy.fft.dst
  type 1) followed by one banded radial solve per z-mode
  (scipy.linalg.solve_banded).  It is the exact discrete solution that the
  eig route approximates.

``implicit_diffusion_residual`` applies (I - nu dt L) by its stencil, with
no matrix shared with either solver, so it checks both.
"""
import numpy as np
import numpy.linalg as la
import scipy.fft as sfft
import scipy.sparse as spp
from scipy.linalg import solve_banded


def radial_coord(grid_size_r, dx, real_dtype=np.float64):
    """implicit_diffusion_solver.py:22-26."""
    return (
        np.linspace(dx / 2, grid_size_r * dx - dx / 2, grid_size_r)
        .astype(real_dtype)
        .reshape(grid_size_r, 1)
    )


def fdm_matrices(grid_size_r, grid_size_z, dx, real_dtype=np.float64):
    """implicit_diffusion_solver.py:40-72: (P_z, P_r, D_r / r - I / r^2)."""
    r = radial_coord(grid_size_r, dx, real_dtype)
    inv_dx2 = real_dtype(1 / dx / dx)
    inv_2dx = real_dtype(1 / 2 / dx)
    P_z = inv_dx2 * spp.diags([1, -2, 1], [-1, 0, 1], shape=(grid_size_z, grid_size_z), format="csr")
    P_z = P_z.toarray().astype(real_dtype)
    P_r = inv_dx2 * spp.diags([1, -2, 1], [-1, 0, 1], shape=(grid_size_r, grid_size_r), format="csr")
    P_r = P_r.toarray().astype(real_dtype)
    N_r = inv_2dx * spp.diags([-1, 1], [-1, 1], shape=(grid_size_r, grid_size_r), format="csr")
    N_r = N_r.toarray().astype(real_dtype)
    N_r[...] = N_r / r - np.identity(grid_size_r) / (r**2)
    return P_z, P_r, N_r


class ImplicitEulerDiffusionStepper:
    """implicit_diffusion_solver.py:6-143 (the eig route)."""

    def __init__(self, time_step, kinematic_viscosity, grid_size_r, grid_size_z, dx,
                 real_dtype=np.float64):
        self.time_step = time_step
        self.nu_times_dt = self.time_step * kinematic_viscosity
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.real_dtype = real_dtype
        P_z, P_r, N_r = fdm_matrices(grid_size_r, grid_size_z, dx, real_dtype)
        # :84-106
        eig_vals_r, eig_vecs_r = la.eig(P_r + N_r)
        idx = eig_vals_r.argsort()[::-1]
        eig_vals_r[...] = eig_vals_r[idx]
        eig_vecs_r[...] = eig_vecs_r[:, idx]
        self.eig_vecs_r = eig_vecs_r
        self.inv_of_eig_vecs_r = la.inv(eig_vecs_r)
        eig_vals_z, eig_vecs_z = la.eig(P_z)
        idx = eig_vals_z.argsort()[::-1]
        eig_vals_z[...] = eig_vals_z[idx]
        eig_vecs_z[...] = eig_vecs_z[:, idx]
        self.tranpose_of_eig_vecs_z = np.transpose(eig_vecs_z)
        self.tranpose_of_inv_of_eig_vecs_z = np.transpose(la.inv(eig_vecs_z))
        eig_val_matrix = real_dtype(1.0) - self.nu_times_dt * (
            np.tile(eig_vals_z.reshape(1, grid_size_z), reps=(grid_size_r, 1))
            + np.tile(eig_vals_r.reshape(grid_size_r, 1), reps=(1, grid_size_z))
        )
        self.inv_eig_val_matrix = real_dtype(1) / eig_val_matrix
        self.spectral_field_buffer = np.zeros_like(self.inv_eig_val_matrix)

    def step(self, vorticity_field, dt):
        """:108-143, in place on vorticity_field."""
        if dt != self.time_step:
            raise ValueError(
                "dt should be constant throughout the simulation if using "
                "implicit diffusion! Please use the value of dt being used "
                "to initialize the implicit diffusion stepper"
            )
        la.multi_dot(
            [self.inv_of_eig_vecs_r, vorticity_field, self.tranpose_of_inv_of_eig_vecs_z],
            out=self.spectral_field_buffer,
        )
        np.multiply(self.spectral_field_buffer, self.inv_eig_val_matrix,
                    out=self.spectral_field_buffer)
        la.multi_dot(
            [self.eig_vecs_r, self.spectral_field_buffer, self.tranpose_of_eig_vecs_z],
            out=vorticity_field,
        )


def dirichlet_z_eigenvalues(grid_size_z, dx):
    """Eigenvalues of the Dirichlet P_z for the DST-I modes k = 1..nz:
    -(4 / dx^2) sin^2(pi k / (2 (nz + 1)))."""
    k = np.arange(1, grid_size_z + 1)
    return -(4.0 / dx**2) * np.sin(np.pi * k / (2.0 * (grid_size_z + 1))) ** 2


def radial_bands(grid_size_r, dx):
    """(sub, diag, super) of A_r = P_r + D_r / r - I / r^2 (sub[0] and
    super[-1] unused), entries formed as the reference forms them."""
    r = radial_coord(grid_size_r, dx)[:, 0]
    inv_dx2 = 1 / dx / dx
    inv_2dx = 1 / 2 / dx
    sub = inv_dx2 * 1.0 + (inv_2dx * -1.0) / r
    sup = inv_dx2 * 1.0 + (inv_2dx * 1.0) / r
    diag = inv_dx2 * -2.0 - 1.0 / (r**2)
    return sub, diag, sup


def exact_implicit_diffusion(omega, nu_dt, dx):
    """The exact solution of (I - nu dt L) w' = omega (DST-I along z, banded
    radial solves)."""
    nr, nz = omega.shape
    lam_z = dirichlet_z_eigenvalues(nz, dx)
    spec = sfft.dst(omega, type=1, axis=1, norm="ortho")
    sub, diag, sup = radial_bands(nr, dx)
    out = np.empty_like(spec)
    ab = np.zeros((3, nr))
    for k in range(nz):
        ab[0, 1:] = -nu_dt * sup[:-1]
        ab[1, :] = 1.0 - nu_dt * (diag + lam_z[k])
        ab[2, :-1] = -nu_dt * sub[1:]
        out[:, k] = solve_banded((1, 1), ab, spec[:, k])
    return sfft.dst(out, type=1, axis=1, norm="ortho")


def implicit_diffusion_residual(w_new, w_old, nu_dt, dx):
    """max |(I - nu dt L) w_new - w_old| / max |w_old|, L applied by its
    5-point stencil with zero ghost values."""
    nr, nz = w_new.shape
    r = (np.arange(nr) + 0.5) * dx
    g = np.zeros((nr + 2, nz + 2))
    g[1:-1, 1:-1] = w_new
    c = g[1:-1, 1:-1]
    lap_z = (g[1:-1, 2:] - 2 * c + g[1:-1, :-2]) / dx**2
    lap_r = (g[2:, 1:-1] - 2 * c + g[:-2, 1:-1]) / dx**2
    d_r = (g[2:, 1:-1] - g[:-2, 1:-1]) / (2 * dx)
    Lw = lap_z + lap_r + d_r / r[:, None] - c / (r[:, None] ** 2)
    res = w_new - nu_dt * Lw - w_old
    return np.max(np.abs(res)) / max(np.max(np.abs(w_old)), 1e-300)
