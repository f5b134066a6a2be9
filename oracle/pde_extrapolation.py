"""Static PDE extrapolation of a reference-map component out of a solid
(test oracle only): examples/PeriodicSoftSlab/bounded_static_PDE_extrapolation.py:5-235,
the scheme the periodic soft-slab driver uses in place of the least-squares
extrapolation (periodic_soft_slab.py:160-169, 367-370).

In the bounding box of {phi + band >= 0} (two cells wider where the grid
allows), with phi negated so that the solid is upwind:
  1. grad_eta_n = n . grad(eta) inside the solid (central differences),
     n the upwind second-order normal of phi (:159-183);
  2. Jacobi iterations of n . grad(grad_eta_n) = 0 in the extrapolation
     zone (second-order one-sided differences along each axis, split by
     the sign of n), until the L2 norm of one iteration's change is at
     most the tolerance (:185-218);
  3. the same iterations for n . grad(eta) = grad_eta_n (:100-111);
  4. eta = 0 everywhere, then the box (less its two-cell margin on sides
     that are not at the grid edge) takes the result (:220-235).
Periodic grids exchange phi's and eta's ghost columns first (:58-62; both
with the communicator the driver passes as per_communicator_eta).

The arithmetic follows the reference expression by expression (numpy, FP64).
"""
import numpy as np

_EPS = np.finfo(float).eps


class StaticPDEExtrapolation:
    """:5-56; the bounding box and its edge flags persist between calls as
    in the reference (:43-51, 125-150)."""

    def __init__(self, dx, grid_size_r, grid_size_z, extrap_tol, extrap_band, periodic=False,
                 per_communicator_gen=None, per_communicator_eta=None):
        if periodic and (not per_communicator_gen) and (not per_communicator_eta):
            raise ValueError("periodic extrapolation needs the ghost communicators")
        self.dx = dx
        self.grid_size_r = grid_size_r
        self.grid_size_z = grid_size_z
        self.extrap_tol = extrap_tol
        self.extrap_band = extrap_band
        self.periodic = periodic
        self.per_communicator_gen = per_communicator_gen
        self.per_communicator_eta = per_communicator_eta
        self.offset = np.sqrt(2) * dx
        self.r_start, self.r_end = 0, grid_size_r
        self.z_start, self.z_end = 0, grid_size_z
        self.r_start_at_boundary = True
        self.r_end_at_boundary = True
        self.z_start_at_boundary = True
        self.z_end_at_boundary = True
        self.iterations = []  # Jacobi iterations of each loop, in call order

    def extrapolate(self, eta, phi):
        """:58-123"""
        if self.periodic:
            self.per_communicator_gen(phi)
            self.per_communicator_eta(eta)
        self._find_bounding_box(phi + self.extrap_band >= 0)
        rs, re, zs, ze = self.r_start, self.r_end, self.z_start, self.z_end
        bphi = -phi[rs:re, zs:ze].copy()
        beta = eta[rs:re, zs:ze].copy()
        inside = bphi[2:-2, 2:-2] <= -self.offset
        zone = (bphi[2:-2, 2:-2] > -self.offset) & (bphi[2:-2, 2:-2] < self.extrap_band)
        n_r, n_z = self.normal_upwind(bphi)
        two_dx = 2 * self.dx
        g_r = (beta[3:-1, 2:-2] - beta[1:-3, 2:-2]) / two_dx
        g_z = (beta[2:-2, 3:-1] - beta[2:-2, 1:-3]) / two_dx
        grad_n = beta * 0
        grad_n[2:-2, 2:-2] = inside * (n_r * g_r + n_z * g_z)
        n_rp = (n_r >= 0) * n_r
        n_zp = (n_z >= 0) * n_z
        n_rm = n_r - n_rp
        n_zm = n_z - n_zp
        denom = 3.0 * (np.abs(n_r) + np.abs(n_z)) + _EPS
        self._jacobi(grad_n, bphi * 0, zone, inside, denom, n_rp, n_rm, n_zp, n_zm)
        self._jacobi(beta, grad_n, zone, inside, denom, n_rp, n_rm, n_zp, n_zm)
        self._restore(eta, beta)

    def _find_bounding_box(self, box):
        """:125-150"""
        r_any = np.any(box, axis=1)
        z_any = np.any(box, axis=0)
        self.r_start, self.r_end = np.where(r_any)[0][[0, -1]]
        self.z_start, self.z_end = np.where(z_any)[0][[0, -1]]
        self.r_end += 1
        self.z_end += 1
        if self.r_start >= 2:
            self.r_start -= 2
            self.r_start_at_boundary = False
        if self.r_end <= self.grid_size_r - 2:
            self.r_end += 2
            self.r_end_at_boundary = False
        if self.z_start >= 2:
            self.z_start -= 2
            self.z_start_at_boundary = False
        if self.z_end <= self.grid_size_z - 2:
            self.z_end += 2
            self.z_end_at_boundary = False
        if min(self.r_end - self.r_start, self.z_end - self.z_start) < 6:
            raise ValueError("too few grid points in the extrapolation box")

    def normal_upwind(self, phi):
        """:159-183: the upwind second-order gradient of phi, normalised."""
        c = phi[2:-2, 2:-2]
        n_r = (np.maximum(3.0 * c - 4.0 * phi[1:-3, 2:-2] + phi[:-4, 2:-2], 0)
               + np.minimum(-3.0 * c + 4.0 * phi[3:-1, 2:-2] - phi[4:, 2:-2], 0)) / (2 * self.dx)
        n_z = (np.maximum(3.0 * c - 4.0 * phi[2:-2, 1:-3] + phi[2:-2, :-4], 0)
               + np.minimum(-3.0 * c + 4.0 * phi[2:-2, 3:-1] - phi[2:-2, 4:], 0)) / (2 * self.dx)
        mag = np.sqrt(n_r**2 + n_z**2)
        n_r /= mag + _EPS
        n_z /= mag + _EPS
        return n_r, n_z

    def _jacobi(self, soln, rhs, zone, inside, denom, n_rp, n_rm, n_zp, n_zm):
        """:185-218: n . grad(soln) = rhs in the zone, soln fixed inside."""
        residual = 1 + self.extrap_tol
        original = soln[2:-2, 2:-2].copy() * inside
        previous = soln[2:-2, 2:-2].copy()
        its = 0
        while residual > self.extrap_tol:
            num_r = n_rp * (4.0 * soln[1:-3, 2:-2] - soln[:-4, 2:-2]) - n_rm * (
                4.0 * soln[3:-1, 2:-2] - soln[4:, 2:-2])
            num_z = n_zp * (4.0 * soln[2:-2, 1:-3] - soln[2:-2, :-4]) - n_zm * (
                4.0 * soln[2:-2, 3:-1] - soln[2:-2, 4:])
            soln[2:-2, 2:-2] = original + zone * (2 * self.dx * rhs[2:-2, 2:-2] + num_r
                                                  + num_z) / denom
            residual = np.linalg.norm(soln[2:-2, 2:-2] - previous)
            previous[...] = soln[2:-2, 2:-2]
            its += 1
        self.iterations.append(its)

    def _restore(self, eta, beta):
        """:220-235"""
        rs0 = 0 if self.r_start_at_boundary else 2
        re0 = None if self.r_end_at_boundary else -2
        zs0 = 0 if self.z_start_at_boundary else 2
        ze0 = None if self.z_end_at_boundary else -2
        eta[...] *= 0
        eta[self.r_start + rs0:self.r_end + (0 if re0 is None else re0),
            self.z_start + zs0:self.z_end + (0 if ze0 is None else ze0)] = beta[rs0:re0, zs0:ze0]
