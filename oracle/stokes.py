"""Stream-function Poisson oracle (test oracle only).

Two CPU solvers of the same discrete system  A_r psi + psi P_z^T = r * omega:

* ``FastDiagonalisationStokesSolver`` restates the reference's dense
  eigen-decomposition route (kernels/FastDiagonalisationStokesSolver.py:6-156)
  exactly: scipy.sparse.diags -> dense, numpy.linalg.eig, descending sort,
  numpy.linalg.inv, numpy.linalg.multi_dot.
* ``exact_stokes_solve`` solves the same matrices by a DCT-II (Neumann z;
  eigenvectors of P_z in closed form) or a real DFT (periodic z) followed by
  banded radial solves with scipy.linalg.solve_banded.  It is the exact
  discrete solution that the eig route approximates (the eig route carries
  the conditioning of the non-normal radial eigenvector matrix).
"""
import numpy as np
import numpy.linalg as la
import scipy.sparse as spp
import scipy.fft as sfft
from scipy.linalg import solve_banded

NEUMANN_ZR = "homogenous_neumann_along_z_and_r"
NEUMANN_R_PERIODIC_Z = "homogenous_neumann_along_r_and_periodic_along_z"
DIRICHLET_R_PERIODIC_Z = "homogenous_dirichlet_along_r_and_periodic_along_z"


def fdm_matrices(grid_size_r, grid_size_z, dx, bc_type=NEUMANN_ZR, real_dtype=np.float64):
    """FastDiagonalisationStokesSolver.py:20-24, 41-97.

    Returns (radial_coord (nr,1), P_z, P_r, D_r) after boundary conditions."""
    radial_coord = (
        np.linspace(dx / 2, grid_size_r * dx - dx / 2, grid_size_r)
        .astype(real_dtype)
        .reshape(grid_size_r, 1)
    )
    inv_dx2 = real_dtype(1 / dx / dx)
    inv_2dx = real_dtype(1 / 2 / dx)
    P_z = inv_dx2 * spp.diags(
        [-1, 2, -1], [-1, 0, 1], shape=(grid_size_z, grid_size_z), format="csr"
    )
    P_z = P_z.toarray().astype(real_dtype)
    P_r = inv_dx2 * spp.diags(
        [-1, 2, -1], [-1, 0, 1], shape=(grid_size_r, grid_size_r), format="csr"
    )
    P_r = P_r.toarray().astype(real_dtype)
    D_r = inv_2dx * spp.diags(
        [1, -1], [-1, 1], shape=(grid_size_r, grid_size_r), format="csr"
    )
    D_r = D_r.toarray().astype(real_dtype)
    D_r[...] = D_r / radial_coord
    # :69-97
    if bc_type == NEUMANN_ZR:
        P_z[0, 0] = inv_dx2
        P_z[-1, -1] = inv_dx2
        P_r[-1, -1] = inv_dx2
        D_r[-1, -2] = 0
    elif bc_type == NEUMANN_R_PERIODIC_Z:
        P_z[0, -1] = P_z[0, 1]
        P_z[-1, 0] = P_z[-1, -2]
        P_r[-1, -1] = inv_dx2
        D_r[-1, -2] = 0
    elif bc_type == DIRICHLET_R_PERIODIC_Z:
        P_z[0, -1] = P_z[0, 1]
        P_z[-1, 0] = P_z[-1, -2]
    else:
        raise ValueError(f"unknown bc_type {bc_type!r}")
    return radial_coord, P_z, P_r, D_r


class FastDiagonalisationStokesSolver:
    """FastDiagonalisationStokesSolver.py:6-156 (the eig route)."""

    def __init__(self, grid_size_r, grid_size_z, dx, real_dtype=np.float64,
                 bc_type=NEUMANN_ZR):
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.real_dtype = real_dtype
        self.bc_type = bc_type
        self.radial_coord, P_z, P_r, D_r = fdm_matrices(
            grid_size_r, grid_size_z, dx, bc_type, real_dtype
        )
        # :99-128
        eig_vals_r, eig_vecs_r = la.eig(P_r - D_r)
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
        eig_val_matrix = np.tile(
            eig_vals_z.reshape(1, grid_size_z), reps=(grid_size_r, 1)
        ) + np.tile(eig_vals_r.reshape(grid_size_r, 1), reps=(1, grid_size_z))
        self.inv_eig_val_matrix = real_dtype(1) / eig_val_matrix
        self.spectral_field_buffer = np.zeros_like(self.inv_eig_val_matrix)

    def solve(self, solution_field, rhs_field):
        """:130-156"""
        la.multi_dot(
            [
                self.inv_of_eig_vecs_r,
                np.multiply(rhs_field, self.radial_coord),
                self.tranpose_of_inv_of_eig_vecs_z,
            ],
            out=self.spectral_field_buffer,
        )
        np.multiply(
            self.spectral_field_buffer,
            self.inv_eig_val_matrix,
            out=self.spectral_field_buffer,
        )
        solution_field[...] = la.multi_dot(
            [self.eig_vecs_r, self.spectral_field_buffer, self.tranpose_of_eig_vecs_z],
        )


def radial_operator(grid_size_r, dx, bc_type=NEUMANN_ZR):
    """A_r = P_r - D_r as (sub, diag, super) vectors of length nr, built with
    the same floating-point operations as fdm_matrices (entries agree bit for
    bit with the dense matrix)."""
    radial_coord, _, P_r, D_r = fdm_matrices(grid_size_r, 3, dx, bc_type)
    A = P_r - D_r
    nr = grid_size_r
    sub = np.zeros(nr)
    sup = np.zeros(nr)
    sub[1:] = np.diagonal(A, -1)
    sup[:-1] = np.diagonal(A, 1)
    return sub, np.diagonal(A).copy(), sup


def z_eigenvalues(grid_size_z, dx, bc_type=NEUMANN_ZR):
    """Closed-form eigenvalues of P_z in transform order.

    Neumann: DCT-II basis cos(pi k (2n+1) / (2N)), lambda_k = (4/dx^2) sin^2(pi k/(2N)).
    Periodic: DFT basis, lambda_k = (4/dx^2) sin^2(pi k / N)."""
    k = np.arange(grid_size_z)
    if bc_type == NEUMANN_ZR:
        return (4.0 / dx**2) * np.sin(np.pi * k / (2 * grid_size_z)) ** 2
    return (4.0 / dx**2) * np.sin(np.pi * k / grid_size_z) ** 2


def exact_stokes_solve(rhs, dx, bc_type=NEUMANN_ZR):
    """Exact discrete solution of A_r psi + psi P_z^T = r*rhs (Neumann z or
    periodic z), by z-diagonalisation and banded radial solves."""
    nr, nz = rhs.shape
    radial_coord = np.linspace(dx / 2, nr * dx - dx / 2, nr).reshape(nr, 1)
    f = rhs * radial_coord
    sub, diag, sup = radial_operator(nr, dx, bc_type)
    lam = z_eigenvalues(nz, dx, bc_type)
    if bc_type == NEUMANN_ZR:
        fh = sfft.dct(f, type=2, axis=1)
    else:
        fh = sfft.fft(f, axis=1)
    out = np.empty_like(fh)
    ab = np.zeros((3, nr))
    ab[0, 1:] = sup[:-1]
    ab[2, :-1] = sub[1:]
    for k in range(nz):
        ab[1, :] = diag + lam[k]
        out[:, k] = solve_banded((1, 1), ab, fh[:, k])
    if bc_type == NEUMANN_ZR:
        return sfft.idct(out, type=2, axis=1)
    return np.real(sfft.ifft(out, axis=1))


def stokes_residual(psi, rhs, dx, bc_type=NEUMANN_ZR):
    """Relative residual ||A_r psi + psi P_z^T - r rhs|| / ||r rhs||, with the
    matrices assembled by fdm_matrices."""
    nr, nz = rhs.shape
    radial_coord, P_z, P_r, D_r = fdm_matrices(nr, nz, dx, bc_type)
    res = (P_r - D_r) @ psi + psi @ P_z.T - rhs * radial_coord
    return np.linalg.norm(res) / np.linalg.norm(rhs * radial_coord)
