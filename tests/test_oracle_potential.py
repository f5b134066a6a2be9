"""The velocity-potential oracle (kernels/FastDiagonalisationPotentialSolver.py,
kernels/compute_velocity_from_phi.py) pinned by what the mathematics fixes:
the DCT-II eigenpairs of the Neumann P_z, the radial bands equal to the
reference's dense matrix bit for bit, the left null vector of A_r, the exact
route satisfying the discrete system by its stencil for a compatible rhs,
the eig route agreeing up to the null-space constant, and velocity
exactness on quadratics."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle import potential as O

SIZES = [(64, 32), (128, 64), (90, 37)]  # (nz, nr)


def test_dct2_diagonalises_neumann_Pz_and_constants_are_null():
    nz, nr = 40, 12
    dx = 1.0 / nz
    _, P_z, P_r, D_r = O.fdm_matrices(nr, nz, dx)
    C = sfft.dct(np.eye(nz), type=2, axis=0, norm="ortho").T  # columns = basis vectors
    lam = O.neumann_z_eigenvalues(nz, dx)
    assert np.max(np.abs(P_z @ C - C * lam[None, :])) * dx**2 < 1e-12
    assert np.max(np.abs((P_r + D_r) @ np.ones(nr))) * dx**2 < 1e-12
    w = O.potential_null_weights(nr, dx)
    assert np.max(np.abs(w @ (P_r + D_r))) / np.max(np.abs(w)) * dx**2 < 1e-12


def test_radial_bands_bitwise_match_dense():
    for nr in (5, 37, 128):
        dx = 1.0 / 256
        _, _, P_r, D_r = O.fdm_matrices(nr, 16, dx)
        A = P_r + D_r
        sub, diag, sup = O.radial_bands(nr, dx)
        assert np.array_equal(np.diag(A), diag)
        assert np.array_equal(np.diag(A, -1), sub[1:])
        assert np.array_equal(np.diag(A, 1), sup[:-1])


@pytest.mark.parametrize("nz,nr", SIZES)
def test_exact_route_solves_compatible_system_and_eig_route_agrees(nz, nr):
    dx = 1.0 / nz
    rhs = O.make_compatible(np.random.default_rng(0).standard_normal((nr, nz)), dx)
    ex = O.exact_potential_solve(rhs, dx)
    assert O.potential_residual(ex, rhs, dx) < 1e-12
    eig = np.zeros_like(rhs)
    O.FastDiagonalisationPotentialSolver(nr, nz, dx).solve(eig, rhs)
    d = eig - ex
    assert np.linalg.norm(d - d.mean()) / np.linalg.norm(ex - ex.mean()) < 1e-11
    # the gauge: z-mode 0's last radial coefficient vanishes
    assert abs(sfft.dct(ex, type=2, axis=1, norm="ortho")[-1, 0]) < 1e-12 * np.abs(ex).max()


def test_incompatible_rhs_blows_up_the_eig_route_only():
    nz, nr = 64, 32
    dx = 1.0 / nz
    rhs = np.ones((nr, nz))
    eig = np.zeros_like(rhs)
    O.FastDiagonalisationPotentialSolver(nr, nz, dx).solve(eig, rhs)
    ex = O.exact_potential_solve(rhs, dx)
    assert np.isfinite(ex).all() and np.abs(ex).max() < 1e3
    assert not np.isfinite(eig).all() or np.abs(eig).max() > 1e6


def test_velocity_from_phi_exact_on_quadratics_and_planted_error_caught():
    nz, nr = 24, 16
    dx = 0.1
    z = (np.arange(nz) + 0.5) * dx
    r = (np.arange(nr) + 0.5) * dx
    Z, R = np.meshgrid(z, r)
    phi = 1.5 * Z**2 - 0.7 * R**2 + 0.3 * Z * R + 2.0 * Z
    u_z, u_r = np.zeros_like(phi), np.zeros_like(phi)
    O.compute_velocity_from_phi_unb(u_z, u_r, phi, dx)
    assert np.allclose(u_z, 3.0 * Z + 0.3 * R + 2.0, rtol=0, atol=1e-11)
    assert np.allclose(u_r, -1.4 * R + 0.3 * Z, rtol=0, atol=1e-11)
    # planted mistake: the one-sided edge formula with the wrong sign of its first term
    bad = u_z.copy()
    bad[:, 0] = (phi[:, 2] + 4 * phi[:, 1] - 3 * phi[:, 0]) / (2 * dx)
    assert not np.allclose(bad, 3.0 * Z + 0.3 * R + 2.0, rtol=0, atol=1e-3)
