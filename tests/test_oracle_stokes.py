"""Pins the Poisson oracle (oracle/stokes.py): the eig route (a literal
restatement of FastDiagonalisationStokesSolver.py) and the exact DCT-II +
banded route, against the discrete operator itself, closed-form
eigenpairs and a manufactured solution.  CPU only."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle.stokes import (
    FastDiagonalisationStokesSolver,
    exact_stokes_solve,
    fdm_matrices,
    radial_operator,
    stokes_residual,
    z_eigenvalues,
)


def test_fdm_matrices_structure():
    """FastDiagonalisationStokesSolver.py:41-86: P_z Neumann rows, c_0 = 0
    exactly for r_0 = dx/2 (power-of-two dx), D_r[-1,-2] = 0."""
    nr, nz, dx = 8, 16, 1.0 / 16
    rc, Pz, Pr, Dr = fdm_matrices(nr, nz, dx)
    assert Pz[0, 0] == Pz[-1, -1] == 1 / dx**2
    assert Pr[-1, -1] == 1 / dx**2 and Dr[-1, -2] == 0
    A = Pr - Dr
    assert A[0, 1] == 0.0
    np.testing.assert_allclose(A[3, 2], -1 / dx**2 - 1 / (2 * dx * rc[3, 0]))
    np.testing.assert_allclose(A[3, 4], -1 / dx**2 + 1 / (2 * dx * rc[3, 0]))


def test_dct2_diagonalises_neumann_Pz():
    """Closed-form eigenpairs: P_z cos(pi k (2n+1)/(2N)) = lambda_k cos(...)."""
    nz, dx = 24, 0.1
    _, Pz, _, _ = fdm_matrices(4, nz, dx)
    lam = z_eigenvalues(nz, dx)
    n = np.arange(nz)
    for k in (0, 1, 5, 23):
        v = np.cos(np.pi * k * (2 * n + 1) / (2 * nz))
        np.testing.assert_allclose(Pz @ v, lam[k] * v, atol=1e-9 * lam.max())
    # and scipy's DCT-II is that transform (unnormalised, factor 2)
    x = np.random.default_rng(0).standard_normal(nz)
    X = np.array([np.sum(x * np.cos(np.pi * k * (2 * n + 1) / (2 * nz))) for k in range(nz)])
    np.testing.assert_allclose(sfft.dct(x, type=2) / 2, X, atol=1e-12)


@pytest.mark.parametrize("nz,nr", [(32, 16), (128, 64), (256, 128)])
def test_exact_route_residual_and_manufactured(nz, nr):
    dx = 1.0 / nz
    rng = np.random.default_rng(nz)
    w = rng.standard_normal((nr, nz))
    psi = exact_stokes_solve(w, dx)
    assert stokes_residual(psi, w, dx) < 1e-12
    # manufactured: pick psi, form rhs with the dense operator, recover psi
    rc, Pz, Pr, Dr = fdm_matrices(nr, nz, dx)
    psi_true = rng.standard_normal((nr, nz))
    rhs = ((Pr - Dr) @ psi_true + psi_true @ Pz.T) / rc
    np.testing.assert_allclose(exact_stokes_solve(rhs, dx), psi_true, atol=1e-9)


@pytest.mark.parametrize("nz,nr", [(64, 32), (256, 128)])
def test_eig_route_matches_exact_route(nz, nr):
    """The reference algorithm (eig route) and the exact solve of the same
    system agree to the eig route's conditioning (2.5e-12 at 256x128)."""
    dx = 1.0 / nz
    w = np.random.default_rng(1).standard_normal((nr, nz))
    s = FastDiagonalisationStokesSolver(nr, nz, dx)
    p = np.zeros((nr, nz))
    s.solve(p, w)
    e = exact_stokes_solve(w, dx)
    assert np.linalg.norm(p - e) / np.linalg.norm(e) < 1e-11
    assert stokes_residual(p, w, dx) < 1e-11


def test_radial_operator_bitwise_matches_dense():
    sub, diag, sup = radial_operator(40, 1.0 / 80)
    _, _, Pr, Dr = fdm_matrices(40, 3, 1.0 / 80)
    A = Pr - Dr
    assert np.array_equal(sub[1:], np.diagonal(A, -1))
    assert np.array_equal(sup[:-1], np.diagonal(A, 1))
    assert np.array_equal(diag, np.diagonal(A))
