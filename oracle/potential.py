"""Velocity-potential Poisson oracle (test oracle only).

The reference's FastDiagonalisationPotentialSolver
(kernels/FastDiagonalisationPotentialSolver.py:6-145) solves
    A_r phi + phi P_z^T = rhs,
P_z = [1, -2, 1] / dx^2 with Neumann corners (-1 / dx^2, :82-83),
A_r = P_r + D_r / r with P_r[-1, -1] = -1 / dx^2 and D_r[-1, -2] = 0 (:84-86).
Both operators annihilate constants (A_r 1 = 0 row by row, P_z 1 = 0), so the
(k = 0, radial null) mode is singular: the reference divides by its
round-off eigenvalue (:114-117), which adds an arbitrary constant to phi.

* ``FastDiagonalisationPotentialSolver`` restates the eig route exactly.
* ``exact_potential_solve`` is the well-posed form: DCT-II along z (the
  eigenvectors of the Neumann P_z in closed form), one banded radial solve
  per z-mode, and for mode 0 the gauge phi_hat_0[nr - 1] = 0 (rows
  0 .. nr-2 of the singular radial system, which are nonsingular).  For a
  compatible rhs (``potential_null_weights``) it solves the system exactly;
  phi then differs from the eig route's by a constant, the velocity not at
  all.
* ``compute_velocity_from_phi_unb`` restates kernels/compute_velocity_from_phi.py:4-17.
"""
import numpy as np
import numpy.linalg as la
import scipy.fft as sfft
import scipy.sparse as spp
from scipy.linalg import solve_banded


def fdm_matrices(grid_size_r, grid_size_z, dx, real_dtype=np.float64):
    """FastDiagonalisationPotentialSolver.py:20-24, 41-86: (radial_coord, P_z, P_r, D_r)
    after the Neumann modifications."""
    r = (np.linspace(dx / 2, grid_size_r * dx - dx / 2, grid_size_r)
         .astype(real_dtype).reshape(grid_size_r, 1))
    inv_dx2 = real_dtype(1 / dx / dx)
    inv_2dx = real_dtype(1 / 2 / dx)
    P_z = inv_dx2 * spp.diags([1, -2, 1], [-1, 0, 1], shape=(grid_size_z, grid_size_z), format="csr")
    P_z = P_z.toarray().astype(real_dtype)
    P_r = inv_dx2 * spp.diags([1, -2, 1], [-1, 0, 1], shape=(grid_size_r, grid_size_r), format="csr")
    P_r = P_r.toarray().astype(real_dtype)
    D_r = inv_2dx * spp.diags([-1, 1], [-1, 1], shape=(grid_size_r, grid_size_r), format="csr")
    D_r = D_r.toarray().astype(real_dtype)
    D_r[...] = D_r / r
    P_z[0, 0] = -inv_dx2
    P_z[-1, -1] = -inv_dx2
    P_r[-1, -1] = -inv_dx2
    D_r[-1, -2] = 0
    return r, P_z, P_r, D_r


class FastDiagonalisationPotentialSolver:
    """FastDiagonalisationPotentialSolver.py:6-145 (the eig route)."""

    def __init__(self, grid_size_r, grid_size_z, dx, real_dtype=np.float64,
                 bc_type="homogenous_neumann_along_z_and_r"):
        if bc_type != "homogenous_neumann_along_z_and_r":
            raise ValueError(bc_type)
        self.radial_coord, P_z, P_r, D_r = fdm_matrices(grid_size_r, grid_size_z, dx, real_dtype)
        eig_vals_r, eig_vecs_r = la.eig(P_r + D_r)
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
        eig_val_matrix = np.tile(eig_vals_z.reshape(1, grid_size_z), reps=(grid_size_r, 1)) + \
            np.tile(eig_vals_r.reshape(grid_size_r, 1), reps=(1, grid_size_z))
        with np.errstate(divide="ignore"):
            self.inv_eig_val_matrix = real_dtype(1) / eig_val_matrix
        self.spectral_field_buffer = np.zeros_like(self.inv_eig_val_matrix)

    def solve(self, solution_field, rhs_field):
        """:119-145."""
        la.multi_dot([self.inv_of_eig_vecs_r, rhs_field, self.tranpose_of_inv_of_eig_vecs_z],
                     out=self.spectral_field_buffer)
        np.multiply(self.spectral_field_buffer, self.inv_eig_val_matrix,
                    out=self.spectral_field_buffer)
        solution_field[...] = la.multi_dot(
            [self.eig_vecs_r, self.spectral_field_buffer, self.tranpose_of_eig_vecs_z])


def radial_bands(grid_size_r, dx):
    """(sub, diag, super) of A_r = P_r + D_r / r after :84-86 (sub[0] and
    super[-1] unused), entries formed as the reference forms them."""
    r = np.linspace(dx / 2, grid_size_r * dx - dx / 2, grid_size_r)
    inv_dx2 = 1 / dx / dx
    inv_2dx = 1 / 2 / dx
    sub = inv_dx2 * 1.0 + (inv_2dx * -1.0) / r
    sup = inv_dx2 * 1.0 + (inv_2dx * 1.0) / r
    diag = np.full(grid_size_r, inv_dx2 * -2.0)
    diag[-1] = -inv_dx2
    sub[-1] = inv_dx2 * 1.0 + 0.0
    return sub, diag, sup


def neumann_z_eigenvalues(grid_size_z, dx):
    """Eigenvalues of the Neumann P_z for the DCT-II modes k = 0..nz-1."""
    k = np.arange(grid_size_z)
    return -(4.0 / dx**2) * np.sin(np.pi * k / (2.0 * grid_size_z)) ** 2


def exact_potential_solve(rhs, dx):
    """The well-posed solution (gauge phi_hat_0[nr-1] = 0 for the singular mode)."""
    nr, nz = rhs.shape
    spec = sfft.dct(rhs, type=2, axis=1, norm="ortho")
    lam = neumann_z_eigenvalues(nz, dx)
    sub, diag, sup = radial_bands(nr, dx)
    out = np.empty_like(spec)
    ab = np.zeros((3, nr))
    for k in range(nz):
        ab[0, 1:] = sup[:-1]
        ab[1, :] = diag + lam[k]
        ab[2, :-1] = sub[1:]
        if k == 0:
            n = nr - 1
            out[:n, 0] = solve_banded((1, 1), ab[:, :n], spec[:n, 0])
            out[n, 0] = 0.0
        else:
            out[:, k] = solve_banded((1, 1), ab, spec[:, k])
    return sfft.idct(out, type=2, axis=1, norm="ortho")


def potential_null_weights(grid_size_r, dx):
    """Left null vector w of A_r (w^T A_r = 0, w[0] = 1): a rhs is compatible
    when sum_ij w_i rhs_ij = 0 (the z left null vector of P_z is 1)."""
    sub, diag, sup = radial_bands(grid_size_r, dx)
    # w^T A_r = 0 column by column: w_{j-1} sup_{j-1} + w_j diag_j + w_{j+1} sub_{j+1} = 0
    w = np.zeros(grid_size_r)
    w[0] = 1.0
    w[1] = -w[0] * diag[0] / sub[1]
    for j in range(1, grid_size_r - 1):
        w[j + 1] = -(w[j - 1] * sup[j - 1] + w[j] * diag[j]) / sub[j + 1]
    return w


def make_compatible(rhs, dx):
    """rhs minus the constant that makes sum_ij w_i rhs_ij vanish."""
    w = potential_null_weights(rhs.shape[0], dx)
    c = np.sum(w[:, None] * rhs) / (np.sum(w) * rhs.shape[1])
    return rhs - c


def potential_residual(phi, rhs, dx):
    """max |A_r phi + phi P_z^T - rhs| / max |rhs|, by the stencil (Neumann
    ghost values equal to the edge value; the axis needs none)."""
    nr, nz = phi.shape
    r = (np.arange(nr) + 0.5) * dx
    gz = np.concatenate([phi[:, :1], phi, phi[:, -1:]], axis=1)
    lap_z = (gz[:, 2:] - 2 * phi + gz[:, :-2]) / dx**2
    up = np.concatenate([phi[1:], phi[-1:]], axis=0)      # Neumann at r_max
    dn = np.concatenate([phi[:1] * 0, phi[:-1]], axis=0)  # axis: ghost coefficient is 0
    lap_r = (up - 2 * phi + dn) / dx**2 + (up - dn) / (2 * dx * r[:, None])
    lap_r[-1] = (phi[-2] - phi[-1]) / dx**2  # last row as :84-86 set it
    res = lap_z + lap_r - rhs
    return np.max(np.abs(res)) / max(np.max(np.abs(rhs)), 1e-300)


def compute_velocity_from_phi_unb(u_z, u_r, phi, dx):
    """kernels/compute_velocity_from_phi.py:4-17."""
    u_z[:, 1:-1] = (phi[:, 2:] - phi[:, :-2]) / (2 * dx)
    u_z[:, 0] = (-phi[:, 2] + 4 * phi[:, 1] - 3 * phi[:, 0]) / (2 * dx)
    u_z[:, -1] = (phi[:, -3] - 4 * phi[:, -2] + 3 * phi[:, -1]) / (2 * dx)
    u_r[1:-1, :] = (phi[2:, :] - phi[:-2, :]) / (2 * dx)
    u_r[0, :] = (-phi[2, :] + 4 * phi[1, :] - 3 * phi[0, :]) / (2 * dx)
    u_r[-1, :] = (phi[-3, :] - 4 * phi[-2, :] + 3 * phi[-1, :]) / (2 * dx)
